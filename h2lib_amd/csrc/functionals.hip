// Boundary data on the device: L2 projection of the Dirichlet or Neumann
// data of a point source onto the piecewise linear basis.
//
// Restates projectL2_bem3d_l_avector (Library/bem3d.c:9009-9088): the load
// vector b_i = int phi_i f (integrate_bem3d_l_avector,
// Library/bem3d.c:8631-8702, with the single-triangle Gauss rule of order q)
// and conjugate gradients on the mass matrix (addeval_mass_linear_bem3d,
// Library/bem3d.c:8878-8908) from x = 0 to |r| <= 1e-15 |b|, at most 500
// steps.  f is eval_dirichlet_fundamental_laplacebem3d or
// eval_neumann_fundamental_laplacebem3d (Library/laplacebem3d.c:1390-1430)
// for the source point x0.
//
// The projection runs over the whole device: one launch for the triangle
// integrals, one for the load vector, then three launches per CG step
// (product with the mass matrix and <p, Mp>; the updates of x and r and
// |r|^2; the new direction), each reduction summed deterministically by the
// last block to finish (fixed partial sums per block, summed in block
// order).  The stopping test runs on the device: the step that meets it
// sets a done flag and the launches after it return at once.  The host
// enqueues CHUNK steps at a time and reads the flag once per chunk while
// the next chunk is already queued.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dmath.h"
#include "h2hip_internal.h"

namespace {

constexpr int PT = 256;      // threads per block
constexpr int PW = PT / 64;
constexpr int PNB = 512;     // blocks at most (grid-stride loops beyond)
constexpr int PCHUNK = 8;    // CG steps between two looks at the flag
constexpr int MAXQ2 = 64;  // points of the triangle rule (q <= 8)

struct ProjArgs {
  const double *x, *n, *g;
  const uint32_t *t;
  uint32_t nv, nt;
  const uint32_t *vt_ptr, *vt;      // vertex -> 4 * triangle + local vertex
  const uint32_t *mrow, *mcol;      // mass matrix, CSR
  const double *mval;
  double rx0[MAXQ2], rx1[MAXQ2], rw[MAXQ2];
  uint32_t nq;
  double src[3];
  int kind;
  double *tri;                      // 3 nt
  double *b, *r, *p, *ap;           // nv each
  double *out;
  double *scal;   // [0] |b|^2, [1] |r|^2, [2] alpha, [3] beta
  double *part;   // 2 partial sums per block
  uint32_t *flag; // [0] done, [1] steps, [2] block counter
};

// Sums two per-thread values over the grid: every block stores its sums
// in part, the last block to finish adds them in block order and leaves
// the totals in tot[0..1] (LDS); true in that block only.  The barrier
// completes the block's stores; the agent-scope fence publishes them before
// the counter increment, and in the last block makes the other blocks'
// sums visible.
__device__ bool grid_sum2(double v0, double v1, double *part, uint32_t *cnt,
                          double *tot) {
  __shared__ double red[2][PW];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
  }
  if (lane == 0) {
    red[0][wv] = v0;
    red[1][wv] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < PW; w++) {
      s0 += red[0][w];
      s1 += red[1][w];
    }
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
    __threadfence();
    last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    if (last) __threadfence();
  }
  __syncthreads();
  if (!last) return false;
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += PT) {
    s0 += part[2 * b];
    s1 += part[2 * b + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  __syncthreads();  // red is reused
  if (lane == 0) {
    red[0][wv] = s0;
    red[1][wv] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < PW; w++) {
      t0 += red[0][w];
      t1 += red[1][w];
    }
    tot[0] = t0;
    tot[1] = t1;
    *cnt = 0;
  }
  __syncthreads();
  return true;
}

// per triangle: g_t sum_k w_k f(x_k) phi_l(x_k), l = 0, 1, 2
__device__ __forceinline__ void tri_integrals(const ProjArgs &a, uint32_t t,
                                              double (&s)[3]) {
  const double *A = a.x + 3 * a.t[3 * t], *B = a.x + 3 * a.t[3 * t + 1],
               *C = a.x + 3 * a.t[3 * t + 2];
  const double *nn = a.n + 3 * t;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t k = 0; k < a.nq; k++) {
    const double u = a.rx0[k], v = a.rx1[k];
    double d[3];
    for (int c = 0; c < 3; c++)
      d[c] = A[c] * (1.0 - u) + B[c] * (u - v) + C[c] * v - a.src[c];
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double rr = sqrt(r2);
    const double f =
        a.kind == 0 ? 1.0 / rr
                    : -(d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2]) /
                          (r2 * rr);
    const double wf = a.rw[k] * f;
    s0 += wf * (1.0 - u);
    s1 += wf * (u - v);
    s2 += wf * v;
  }
  s[0] = s0 * a.g[t];
  s[1] = s1 * a.g[t];
  s[2] = s2 * a.g[t];
}

__global__ __launch_bounds__(PT) void proj_tri(ProjArgs a) {
  for (uint32_t t = blockIdx.x * PT + threadIdx.x; t < a.nt;
       t += gridDim.x * PT) {
    double s[3];
    tri_integrals(a, t, s);
    a.tri[3 * t] = s[0];
    a.tri[3 * t + 1] = s[1];
    a.tri[3 * t + 2] = s[2];
  }
}

// The load vector alone (out = or += b), one thread per vertex over its
// triangles in the order of the vertex list, the triangle integrals
// recomputed per vertex (no workspace, so calls need no lock)
__global__ __launch_bounds__(PT) void load_vector(ProjArgs a, int acc) {
  for (uint32_t i = blockIdx.x * PT + threadIdx.x; i < a.nv;
       i += gridDim.x * PT) {
    double sum = 0.0;
    for (uint32_t e = a.vt_ptr[i]; e < a.vt_ptr[i + 1]; e++) {
      const uint32_t code = a.vt[e];
      double s[3];
      tri_integrals(a, code >> 2, s);
      sum += s[code & 3];
    }
    a.out[i] = acc ? a.out[i] + sum : sum;
  }
}

// b_i = sum over the triangles of vertex i; CG start x = 0, r = p = b,
// and the test before the first step
__global__ __launch_bounds__(PT) void proj_init(ProjArgs a) {
  double part = 0.0;
  for (uint32_t i = blockIdx.x * PT + threadIdx.x; i < a.nv;
       i += gridDim.x * PT) {
    double s = 0.0;
    for (uint32_t e = a.vt_ptr[i]; e < a.vt_ptr[i + 1]; e++) {
      const uint32_t code = a.vt[e];
      s += a.tri[3 * (code >> 2) + (code & 3)];
    }
    a.b[i] = s;
    a.r[i] = s;
    a.p[i] = s;
    a.out[i] = 0.0;
    part += s * s;
  }
  __shared__ double tot[2];
  if (!grid_sum2(part, 0.0, a.part, a.flag + 2, tot)) return;
  if (threadIdx.x == 0) {
    a.scal[0] = tot[0];
    a.scal[1] = tot[0];
    a.flag[1] = 0;
    a.flag[0] = tot[0] > 1e-30 * tot[0] ? 0u : 1u;
  }
}

// ap = M p, alpha = |r|^2 / <p, ap>
__global__ __launch_bounds__(PT) void proj_spmv(ProjArgs a) {
  if (a.flag[0]) return;  // uniform over the grid
  double part = 0.0;
  for (uint32_t i = blockIdx.x * PT + threadIdx.x; i < a.nv;
       i += gridDim.x * PT) {
    double s = 0.0;
    for (uint32_t e = a.mrow[i]; e < a.mrow[i + 1]; e++)
      s += a.mval[e] * a.p[a.mcol[e]];
    a.ap[i] = s;
    part += a.p[i] * s;
  }
  __shared__ double tot[2];
  if (!grid_sum2(part, 0.0, a.part, a.flag + 2, tot)) return;
  if (threadIdx.x == 0) a.scal[2] = a.scal[1] / tot[0];
}

// x += alpha p, r -= alpha ap, then |r|^2, beta, the step count and the
// stopping test (at most 500 steps, |r|^2 <= 1e-30 |b|^2)
__global__ __launch_bounds__(PT) void proj_update(ProjArgs a) {
  if (a.flag[0]) return;
  const double alpha = a.scal[2];
  double part = 0.0;
  for (uint32_t i = blockIdx.x * PT + threadIdx.x; i < a.nv;
       i += gridDim.x * PT) {
    a.out[i] += alpha * a.p[i];
    const double ri = a.r[i] - alpha * a.ap[i];
    a.r[i] = ri;
    part += ri * ri;
  }
  __shared__ double tot[2];
  if (!grid_sum2(part, 0.0, a.part, a.flag + 2, tot)) return;
  if (threadIdx.x == 0) {
    const double rr = tot[0];
    a.scal[3] = rr / a.scal[1];
    a.scal[1] = rr;
    const uint32_t it = a.flag[1] + 1;
    a.flag[1] = it;
    if (!(it < 500 && rr > 1e-30 * a.scal[0])) a.flag[0] = 1;
  }
}

// p = r + beta p
__global__ __launch_bounds__(PT) void proj_dir(ProjArgs a) {
  if (a.flag[0]) return;
  const double beta = a.scal[3];
  for (uint32_t i = blockIdx.x * PT + threadIdx.x; i < a.nv;
       i += gridDim.x * PT)
    a.p[i] = a.r[i] + beta * a.p[i];
}

// vertex -> (triangle, local vertex) lists and the CSR mass matrix
// M_ij = sum over triangles containing i and j of g_t (1/12 if i = j,
// 1/24 otherwise) (Library/laplacebem3d.c:985-994)
int build_projection_data(const h2hip_surface *s) {
  const uint32_t nv = s->nv, nt = s->nt;
  std::vector<uint32_t> vt_ptr(nv + 1, 0), vt(3 * (size_t)nt);
  for (size_t i = 0; i < 3 * (size_t)nt; i++) vt_ptr[s->t[i] + 1]++;
  for (uint32_t v = 0; v < nv; v++) vt_ptr[v + 1] += vt_ptr[v];
  std::vector<uint32_t> fill(vt_ptr.begin(), vt_ptr.end() - 1);
  for (uint32_t t = 0; t < nt; t++)
    for (uint32_t l = 0; l < 3; l++) vt[fill[s->t[3 * t + l]]++] = 4 * t + l;
  std::vector<uint32_t> mrow(nv + 1, 0), mcol;
  std::vector<double> mval;
  std::vector<int64_t> pos(nv, -1);
  std::vector<uint32_t> touched;
  for (uint32_t i = 0; i < nv; i++) {
    touched.clear();
    for (uint32_t e = vt_ptr[i]; e < vt_ptr[i + 1]; e++) {
      const uint32_t t = vt[e] >> 2;
      for (uint32_t k = 0; k < 3; k++) {
        const uint32_t j = s->t[3 * t + k];
        const double m = s->g[t] * (j == i ? 1.0 / 12.0 : 1.0 / 24.0);
        if (pos[j] < 0) {
          pos[j] = (int64_t)mcol.size();
          mcol.push_back(j);
          mval.push_back(m);
          touched.push_back(j);
        } else {
          mval[pos[j]] += m;
        }
      }
    }
    for (uint32_t j : touched) pos[j] = -1;
    mrow[i + 1] = (uint32_t)mcol.size();
  }
  if (s->d_vt_ptr.upload(vt_ptr.data(), vt_ptr.size()) != hipSuccess ||
      s->d_vt.upload(vt.data(), vt.size()) != hipSuccess ||
      s->d_mrow.upload(mrow.data(), mrow.size()) != hipSuccess ||
      s->d_mcol.upload(mcol.data(), mcol.size()) != hipSuccess ||
      s->d_mval.upload(mval.data(), mval.size()) != hipSuccess ||
      s->d_fwork.alloc(3 * (size_t)nt + 4 * (size_t)nv + 16 + 2 * PNB) !=
          hipSuccess ||
      s->d_pflag.alloc(4) != hipSuccess ||
      hipHostMalloc((void **)&s->h_plook, 4 * sizeof(uint32_t),
                    hipHostMallocDefault) != hipSuccess) {
    h2hip::set_error("h2hip_projectL2: device allocation failed");
    return H2HIP_ENOMEM;
  }
  for (auto &e : s->pev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      h2hip::set_error("h2hip_projectL2: event creation failed");
      return H2HIP_EDEVICE;
    }
  // the block counter starts at zero; every reduction resets it
  if (hipMemset(s->d_pflag.p, 0, 4 * sizeof(uint32_t)) != hipSuccess) {
    h2hip::set_error("h2hip_projectL2: device error");
    return H2HIP_EDEVICE;
  }
  return H2HIP_OK;
}

}  // namespace

int h2hip::surface_vertex_lists(const h2hip_surface *s) {
  std::lock_guard<std::mutex> guard(s->fl_lock);
  if (!s->fl_ready) {
    const int rc = build_projection_data(s);
    if (rc != H2HIP_OK) return rc;
    s->fl_ready = true;
  }
  return H2HIP_OK;
}

extern "C" int h2hip_projectL2_point_source(const h2hip_surface *s,
                                            uint32_t q, int kind,
                                            const double *x0, double *d_out,
                                            void *stream) {
  if (!s || !x0 || !d_out || q < 1 || q > 8 || (kind != 0 && kind != 1)) {
    h2hip::set_error("h2hip_projectL2_point_source: invalid arguments");
    return H2HIP_EINVAL;
  }
  {
    std::lock_guard<std::mutex> guard(s->fl_lock);
    if (!s->fl_ready) {
      const int rc = build_projection_data(s);
      if (rc != H2HIP_OK) return rc;
      s->fl_ready = true;
    }
  }
  h2hip::TriRule tr = h2hip::make_triangle_rule(q);
  ProjArgs a;
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.t = s->d_t.p;
  a.nv = s->nv;
  a.nt = s->nt;
  a.vt_ptr = s->d_vt_ptr.p;
  a.vt = s->d_vt.p;
  a.mrow = s->d_mrow.p;
  a.mcol = s->d_mcol.p;
  a.mval = s->d_mval.p;
  a.nq = (uint32_t)tr.w.size();
  for (uint32_t k = 0; k < a.nq; k++) {
    a.rx0[k] = tr.x0[k];
    a.rx1[k] = tr.x1[k];
    a.rw[k] = tr.w[k];
  }
  for (int c = 0; c < 3; c++) a.src[c] = x0[c];
  a.kind = kind;
  double *w = s->d_fwork.p;
  a.tri = w;
  a.b = w + 3 * (size_t)s->nt;
  a.r = a.b + s->nv;
  a.p = a.r + s->nv;
  a.ap = a.p + s->nv;
  a.out = d_out;
  a.scal = a.ap + s->nv;
  a.part = a.scal + 16;
  a.flag = s->d_pflag.p;
  const hipStream_t st = (hipStream_t)stream;
  const dim3 gt((uint32_t)std::min<uint64_t>(PNB, (s->nt + PT - 1) / PT));
  const dim3 gv((uint32_t)std::min<uint64_t>(PNB, (s->nv + PT - 1) / PT));
  // one projection per surface at a time (shared workspace and looks)
  std::lock_guard<std::mutex> guard(s->proj_lock);
  hipLaunchKernelGGL(proj_tri, gt, dim3(PT), 0, st, a);
  hipLaunchKernelGGL(proj_init, gv, dim3(PT), 0, st, a);
  auto chunk = [&](int slot) -> hipError_t {
    for (int j = 0; j < PCHUNK; j++) {
      hipLaunchKernelGGL(proj_spmv, gv, dim3(PT), 0, st, a);
      hipLaunchKernelGGL(proj_update, gv, dim3(PT), 0, st, a);
      hipLaunchKernelGGL(proj_dir, gv, dim3(PT), 0, st, a);
    }
    hipError_t e = hipMemcpyAsync(s->h_plook + 2 * slot, s->d_pflag.p,
                                  2 * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(s->pev[slot], st);
    return e;
  };
  // while the host waits for the look after one chunk, the next chunk is
  // already queued; at most 500 steps, so the loop ends
  int slot = 0;
  H2HIP_CHECK(chunk(0));
  for (;;) {
    H2HIP_CHECK(chunk(slot ^ 1));
    H2HIP_CHECK(hipEventSynchronize(s->pev[slot]));
    if (s->h_plook[2 * slot]) break;
    slot ^= 1;
  }
  // the chunk queued behind the last look returns at once; it finishes
  // before the workspace can be used again
  H2HIP_CHECK(hipEventSynchronize(s->pev[slot ^ 1]));
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

extern "C" int h2hip_integrate_point_source(const h2hip_surface *s,
                                            uint32_t q, int kind,
                                            const double *x0, int accumulate,
                                            double *d_out, void *stream) {
  if (!s || !x0 || !d_out || q < 1 || q > 8 || (kind != 0 && kind != 1)) {
    h2hip::set_error("h2hip_integrate_point_source: invalid arguments");
    return H2HIP_EINVAL;
  }
  const int rc = h2hip::surface_vertex_lists(s);
  if (rc != H2HIP_OK) return rc;
  h2hip::TriRule tr = h2hip::make_triangle_rule(q);
  ProjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.t = s->d_t.p;
  a.nv = s->nv;
  a.nt = s->nt;
  a.vt_ptr = s->d_vt_ptr.p;
  a.vt = s->d_vt.p;
  a.nq = (uint32_t)tr.w.size();
  for (uint32_t k = 0; k < a.nq; k++) {
    a.rx0[k] = tr.x0[k];
    a.rx1[k] = tr.x1[k];
    a.rw[k] = tr.w[k];
  }
  for (int c = 0; c < 3; c++) a.src[c] = x0[c];
  a.kind = kind;
  a.out = d_out;
  const dim3 gv((uint32_t)std::min<uint64_t>(PNB, (s->nv + PT - 1) / PT));
  hipLaunchKernelGGL(load_vector, gv, dim3(PT), 0, (hipStream_t)stream, a,
                     accumulate ? 1 : 0);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// ---- Helmholtz point-source data and error norms ---------------------------
//
// Dirichlet data e^{ikr}/r (kind 0) or Neumann data e^{ikr}/r^3 (ikr - 1)
// <n, x - x0> (kind 1) of a point source x0, r = |x - x0|, real k
// (rhs_dirichlet_point_helmholtzbem3d / rhs_neumann_point_helmholtzbem3d,
// Library/helmholtzbem3d.c:1521-1578), against the single-triangle rule of
// order q:
//   mode 0  projection onto the constants: out[t] = 2 sum_q w_q f(x_q)
//           (projectL2_bem3d_c_avector, Library/bem3d.c:8827-8876; the
//           rule's weights sum to 1/2)
//   mode 1  integrals against the constants: out[t] = g_t sum_q w_q f
//           (integrate_bem3d_c_avector, Library/bem3d.c:8579-8629)
//   mode 2  integrals against the linear basis: out[i] = sum over the
//           triangles t at vertex i of g_t sum_q w_q phi_i(x_q) f(x_q)
//           (integrate_bem3d_l_avector, Library/bem3d.c:8631-8702), one
//           thread per vertex over its triangles (no atomics)
// and the norms |x - f|_L2, |f|_L2 of a constant-basis x
// (normL2diff_c_bem3d, Library/bem3d.c:8776-8825, normL2_bem3d,
// Library/bem3d.c:8704-8751).

namespace {

constexpr int HD_NT = 256;
constexpr int HD_MAXQ = 64;

struct HelmArgs {
  const double *x, *n, *g;
  const uint32_t *t;
  const uint32_t *vt_ptr, *vt;
  uint32_t nt, nv;
  uint32_t nq;
  double rx0[HD_MAXQ], rx1[HD_MAXQ], rw[HD_MAXQ];
  double src[3];
  double k;
  int kind, accumulate;
  double *out;
  const double *xv;  // error norms: the constant-basis vector
  double *part;      // error norms: 2 per block
};

__device__ __forceinline__ void helm_point(const HelmArgs &a,
                                           const double *nn, double px,
                                           double py, double pz, double &re,
                                           double &im) {
  const double d0 = px - a.src[0], d1 = py - a.src[1], d2 = pz - a.src[2];
  const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
  const double rs = 1.0 / sqrt(r2), r = r2 * rs;
  double sn, cs;
  h2hip::sincos_cw(a.k * r, sn, cs);
  re = cs * rs;
  im = sn * rs;
  if (a.kind == 1) {
    const double f = (nn[0] * d0 + nn[1] * d1 + nn[2] * d2) * rs * rs;
    const double kr = a.k * r;
    const double nr = -re - im * kr, ni = re * kr - im;
    re = nr * f;
    im = ni * f;
  }
}

// sum_q w_q phi(x_q) f(x_q) on triangle t; corner < 0: phi = 1
__device__ void helm_tri(const HelmArgs &a, uint32_t t, int corner,
                         double &sre, double &sim) {
  const double *A = a.x + 3 * a.t[3 * t], *B = a.x + 3 * a.t[3 * t + 1],
               *C = a.x + 3 * a.t[3 * t + 2];
  const double *nn = a.n + 3 * t;
  sre = 0.0;
  sim = 0.0;
  for (uint32_t q = 0; q < a.nq; q++) {
    const double u = a.rx0[q], v = a.rx1[q];
    const double ca = 1.0 - u, cb = u - v, cc = v;
    double re, im;
    helm_point(a, nn, A[0] * ca + B[0] * cb + C[0] * cc,
               A[1] * ca + B[1] * cb + C[1] * cc,
               A[2] * ca + B[2] * cb + C[2] * cc, re, im);
    const double w = a.rw[q] * (corner < 0 ? 1.0 : corner == 0 ? ca
                                                 : corner == 1 ? cb : cc);
    sre += w * re;
    sim += w * im;
  }
}

template <int MODE>
__global__ __launch_bounds__(HD_NT) void helm_data(HelmArgs a) {
  const uint32_t i = blockIdx.x * HD_NT + threadIdx.x;
  double re = 0.0, im = 0.0;
  if constexpr (MODE == 2) {
    if (i >= a.nv) return;
    for (uint32_t e = a.vt_ptr[i]; e < a.vt_ptr[i + 1]; e++) {
      const uint32_t code = a.vt[e], t = code >> 2;
      double sr, si;
      helm_tri(a, t, (int)(code & 3), sr, si);
      re += sr * a.g[t];
      im += si * a.g[t];
    }
  } else {
    if (i >= a.nt) return;
    helm_tri(a, i, -1, re, im);
    const double f = MODE == 0 ? 2.0 : a.g[i];
    re *= f;
    im *= f;
  }
  if (a.accumulate) {
    a.out[2 * i] += re;
    a.out[2 * i + 1] += im;
  } else {
    a.out[2 * i] = re;
    a.out[2 * i + 1] = im;
  }
}

__global__ __launch_bounds__(HD_NT) void helm_errors(HelmArgs a) {
  const uint32_t t = blockIdx.x * HD_NT + threadIdx.x;
  double sd = 0.0, sn = 0.0;
  if (t < a.nt) {
    const double *A = a.x + 3 * a.t[3 * t], *B = a.x + 3 * a.t[3 * t + 1],
                 *C = a.x + 3 * a.t[3 * t + 2];
    const double *nn = a.n + 3 * t;
    const double xr = a.xv[2 * t], xi = a.xv[2 * t + 1];
    for (uint32_t q = 0; q < a.nq; q++) {
      const double u = a.rx0[q], v = a.rx1[q];
      const double ca = 1.0 - u, cb = u - v, cc = v;
      double re, im;
      helm_point(a, nn, A[0] * ca + B[0] * cb + C[0] * cc,
                 A[1] * ca + B[1] * cb + C[1] * cc,
                 A[2] * ca + B[2] * cb + C[2] * cc, re, im);
      const double er = re - xr, ei = im - xi;
      sd += a.rw[q] * (er * er + ei * ei);
      sn += a.rw[q] * (re * re + im * im);
    }
    sd *= a.g[t];
    sn *= a.g[t];
  }
  __shared__ double red[2][HD_NT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    sd += __shfl_xor(sd, off, 64);
    sn += __shfl_xor(sn, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sd;
    red[1][threadIdx.x >> 6] = sn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d = 0.0, m = 0.0;
    for (int w = 0; w < HD_NT / 64; w++) {
      d += red[0][w];
      m += red[1][w];
    }
    a.part[2 * blockIdx.x] = d;
    a.part[2 * blockIdx.x + 1] = m;
  }
}

int helm_setup(const h2hip_surface *s, uint32_t q, int kind, double k,
               const double *x0, HelmArgs &a) {
  const h2hip::TriRule tr = h2hip::make_triangle_rule(q);
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.t = s->d_t.p;
  a.vt_ptr = s->d_vt_ptr.p;
  a.vt = s->d_vt.p;
  a.nt = s->nt;
  a.nv = s->nv;
  a.nq = (uint32_t)tr.w.size();
  if (a.nq > (uint32_t)HD_MAXQ) return H2HIP_EINVAL;
  for (uint32_t i = 0; i < a.nq; i++) {
    a.rx0[i] = tr.x0[i];
    a.rx1[i] = tr.x1[i];
    a.rw[i] = tr.w[i];
  }
  for (int c = 0; c < 3; c++) a.src[c] = x0[c];
  a.k = k;
  a.kind = kind;
  a.accumulate = 0;
  a.out = nullptr;
  a.xv = nullptr;
  a.part = nullptr;
  return H2HIP_OK;
}

}  // namespace

extern "C" int h2hip_helmholtz_point_data(const h2hip_surface *s, uint32_t q,
                                          int kind, int mode, double k,
                                          const double *x0, int accumulate,
                                          void *d_out, void *stream) {
  if (!s || !x0 || !d_out || q < 1 || q > 8 || (kind != 0 && kind != 1) ||
      mode < 0 || mode > 2) {
    h2hip::set_error("h2hip_helmholtz_point_data: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (mode == 2) {
    std::lock_guard<std::mutex> guard(s->fl_lock);
    if (!s->fl_ready) {
      const int rc = build_projection_data(s);
      if (rc != H2HIP_OK) return rc;
      s->fl_ready = true;
    }
  }
  HelmArgs a;
  if (helm_setup(s, q, kind, k, x0, a) != H2HIP_OK) {
    h2hip::set_error("h2hip_helmholtz_point_data: rule too large");
    return H2HIP_EINVAL;
  }
  a.accumulate = accumulate ? 1 : 0;
  a.out = (double *)d_out;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t n = mode == 2 ? s->nv : s->nt;
  const dim3 grid((n + HD_NT - 1) / HD_NT);
  if (mode == 0)
    hipLaunchKernelGGL(helm_data<0>, grid, dim3(HD_NT), 0, st, a);
  else if (mode == 1)
    hipLaunchKernelGGL(helm_data<1>, grid, dim3(HD_NT), 0, st, a);
  else
    hipLaunchKernelGGL(helm_data<2>, grid, dim3(HD_NT), 0, st, a);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

extern "C" int h2hip_helmholtz_point_errors_c(const h2hip_surface *s,
                                              uint32_t q, int kind, double k,
                                              const double *x0,
                                              const void *d_x, double *out,
                                              void *stream) {
  if (!s || !x0 || !d_x || !out || q < 1 || q > 8 ||
      (kind != 0 && kind != 1)) {
    h2hip::set_error("h2hip_helmholtz_point_errors_c: invalid arguments");
    return H2HIP_EINVAL;
  }
  HelmArgs a;
  if (helm_setup(s, q, kind, k, x0, a) != H2HIP_OK) {
    h2hip::set_error("h2hip_helmholtz_point_errors_c: rule too large");
    return H2HIP_EINVAL;
  }
  const uint32_t nb = (s->nt + HD_NT - 1) / HD_NT;
  h2hip::DevBuf<double> part;
  H2HIP_CHECK(part.alloc(2 * (size_t)nb));
  a.xv = (const double *)d_x;
  a.part = part.p;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(helm_errors, dim3(nb), dim3(HD_NT), 0, st, a);
  H2HIP_CHECK(hipGetLastError());
  std::vector<double> h(2 * (size_t)nb);
  H2HIP_CHECK(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(),
                             hipMemcpyDeviceToHost, st));
  H2HIP_CHECK(hipStreamSynchronize(st));
  double d = 0.0, m = 0.0;
  for (uint32_t b = 0; b < nb; b++) {
    d += h[2 * b];
    m += h[2 * b + 1];
  }
  out[0] = std::sqrt(d);
  out[1] = std::sqrt(m);
  return H2HIP_OK;
}

// ---- exterior potential of the Brakhage-Werner density ---------------------
//
// u(x) = kernel_const sum_t g_t w_t sum_q ww_q e^{ik r}/r^3 ((1 - ik r)
// <x - y_q, n_t> - i eta r^2), r = |x - y_q|, y_q the points of the
// single-triangle rule of order q on triangle t, as the reference's
// Helmholtz test evaluates its solution (eval_brakhage_werner_c,
// Tests/test_helmholtzbem3d.c:67-132; for a piecewise linear density w
// enters through the corner values weighted by the basis functions,
// eval_brakhage_werner_l, Tests/test_helmholtzbem3d.c:133-221).  One
// workgroup per evaluation point, the triangles split over its threads,
// one reduction at the end.

namespace {

constexpr int BW_NT = 256;
constexpr int BW_MAXQ = 16;

struct BWArgs {
  const double *x, *n, *g;
  const uint32_t *t;
  uint32_t nt;
  const double *w;  // complex density, one per triangle
  const double *pts;
  uint32_t nq;
  double rx0[BW_MAXQ], rx1[BW_MAXQ], rw[BW_MAXQ];
  double kr, ki, er, ei;  // k and eta
  double *out;            // complex, one per point
};

// LIN: piecewise linear density (one value per vertex, w_i weighted by
// the basis function of corner i at each point) instead of constant
template <int LIN>
__global__ __launch_bounds__(BW_NT) void bw_potential(BWArgs a) {
  const uint32_t p = blockIdx.x;
  const double px = a.pts[3 * p], py = a.pts[3 * p + 1], pz = a.pts[3 * p + 2];
  double sre = 0.0, sim = 0.0;
  for (uint32_t t = threadIdx.x; t < a.nt; t += BW_NT) {
    const double *A = a.x + 3 * a.t[3 * t], *B = a.x + 3 * a.t[3 * t + 1],
                 *C = a.x + 3 * a.t[3 * t + 2];
    const double nx = a.n[3 * t], ny = a.n[3 * t + 1], nz = a.n[3 * t + 2];
    double qre = 0.0, qim = 0.0;
    double lre[3] = {0.0, 0.0, 0.0}, lim[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = 0; q < a.nq; q++) {
      const double tx = a.rx0[q], sx = a.rx1[q];
      const double ca = 1.0 - tx, cb = tx - sx, cc = sx;
      const double dx = px - (A[0] * ca + B[0] * cb + C[0] * cc);
      const double dy = py - (A[1] * ca + B[1] * cb + C[1] * cc);
      const double dz = pz - (A[2] * ca + B[2] * cb + C[2] * cc);
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double rs = 1.0 / sqrt(r2);
      const double r = r2 * rs;
      const double rs3 = rs * rs * rs;
      // e^{ikr} = e^{-ki r} (cos kr r + i sin kr r)
      double sn, cs;
      h2hip::sincos_cw(a.kr * r, sn, cs);
      const double damp = a.ki != 0.0 ? exp(-a.ki * r) : 1.0;
      const double ere = damp * cs, eim = damp * sn;
      const double dn = dx * nx + dy * ny + dz * nz;
      // (1 - ik r) dn - i eta r^2
      const double fre = dn * (1.0 + a.ki * r) + a.ei * r2;
      const double fim = -dn * a.kr * r - a.er * r2;
      const double vre = (ere * fre - eim * fim) * rs3;
      const double vim = (ere * fim + eim * fre) * rs3;
      if constexpr (LIN) {
        const double phi[3] = {ca, cb, cc};
#pragma unroll
        for (int i = 0; i < 3; i++) {
          lre[i] += a.rw[q] * phi[i] * vre;
          lim[i] += a.rw[q] * phi[i] * vim;
        }
      } else {
        qre += a.rw[q] * vre;
        qim += a.rw[q] * vim;
      }
    }
    const double gw = a.g[t];
    if constexpr (LIN) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const uint32_t v = a.t[3 * t + i];
        const double wr = a.w[2 * v], wi = a.w[2 * v + 1];
        sre += gw * (lre[i] * wr - lim[i] * wi);
        sim += gw * (lre[i] * wi + lim[i] * wr);
      }
    } else {
      const double wr = a.w[2 * t], wi = a.w[2 * t + 1];
      sre += gw * (qre * wr - qim * wi);
      sim += gw * (qre * wi + qim * wr);
    }
  }
  __shared__ double red[2][BW_NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sre += __shfl_xor(sre, off, 64);
    sim += __shfl_xor(sim, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sre;
    red[1][threadIdx.x >> 6] = sim;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double re = 0.0, im = 0.0;
    for (int v = 0; v < BW_NT / 64; v++) {
      re += red[0][v];
      im += red[1][v];
    }
    constexpr double KCONST = 0.0795774715459476679;
    a.out[2 * p] = KCONST * re;
    a.out[2 * p + 1] = KCONST * im;
  }
}

}  // namespace

namespace {

int eval_bw(const char *name, int lin, const h2hip_surface *s, uint32_t q,
            double kr, double ki, const double *eta, const void *d_w,
            uint32_t npts, const double *pts, void *out, void *stream) {
  if (!s || !eta || !d_w || (npts && (!pts || !out)) || q < 1 || q > 4) {
    h2hip::set_error(std::string(name) + ": invalid arguments");
    return H2HIP_EINVAL;
  }
  if (npts == 0) return H2HIP_OK;
  const h2hip::TriRule tr = h2hip::make_triangle_rule(q);
  BWArgs a;
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.t = s->d_t.p;
  a.nt = s->nt;
  a.w = (const double *)d_w;
  a.nq = (uint32_t)tr.w.size();
  for (uint32_t i = 0; i < a.nq; i++) {
    a.rx0[i] = tr.x0[i];
    a.rx1[i] = tr.x1[i];
    a.rw[i] = tr.w[i];
  }
  a.kr = kr;
  a.ki = ki;
  a.er = eta[0];
  a.ei = eta[1];
  h2hip::DevBuf<double> d_pts, d_out;
  H2HIP_CHECK(d_pts.upload(pts, 3 * (size_t)npts));
  H2HIP_CHECK(d_out.alloc(2 * (size_t)npts));
  a.pts = d_pts.p;
  a.out = d_out.p;
  hipStream_t st = (hipStream_t)stream;
  if (lin)
    hipLaunchKernelGGL(bw_potential<1>, dim3(npts), dim3(BW_NT), 0, st, a);
  else
    hipLaunchKernelGGL(bw_potential<0>, dim3(npts), dim3(BW_NT), 0, st, a);
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipMemcpyAsync(out, d_out.p, sizeof(double) * 2 * npts,
                             hipMemcpyDeviceToHost, st));
  H2HIP_CHECK(hipStreamSynchronize(st));
  return H2HIP_OK;
}

}  // namespace

extern "C" int h2hip_eval_brakhage_werner_c(const h2hip_surface *s,
                                            uint32_t q, const double *k,
                                            const double *eta,
                                            const void *d_w, uint32_t npts,
                                            const double *pts, void *out,
                                            void *stream) {
  if (!k) {
    h2hip::set_error("h2hip_eval_brakhage_werner_c: invalid arguments");
    return H2HIP_EINVAL;
  }
  return eval_bw("h2hip_eval_brakhage_werner_c", 0, s, q, k[0], k[1], eta,
                 d_w, npts, pts, out, stream);
}

// The reference's linear variant declares its wave number real (real k =
// bem->k, Tests/test_helmholtzbem3d.c:144), so only Re k enters.
extern "C" int h2hip_eval_brakhage_werner_l(const h2hip_surface *s,
                                            uint32_t q, const double *k,
                                            const double *eta,
                                            const void *d_w, uint32_t npts,
                                            const double *pts, void *out,
                                            void *stream) {
  if (!k) {
    h2hip::set_error("h2hip_eval_brakhage_werner_l: invalid arguments");
    return H2HIP_EINVAL;
  }
  return eval_bw("h2hip_eval_brakhage_werner_l", 1, s, q, k[0], 0.0, eta,
                 d_w, npts, pts, out, stream);
}

// ---- linear combinations of operators --------------------------------------

struct h2hip_combined {
  h2hip_addeval_t f1 = nullptr, f2 = nullptr;
  void *A1 = nullptr, *A2 = nullptr;
  double c1[2] = {1.0, 0.0}, c2[2] = {0.0, 0.0};
};

extern "C" int h2hip_combined_create(h2hip_combined **out, h2hip_addeval_t f1,
                                     void *A1, const double *c1,
                                     h2hip_addeval_t f2, void *A2,
                                     const double *c2) {
  if (!out || !f1 || !c1 || (f2 && !c2)) {
    h2hip::set_error("h2hip_combined_create: invalid arguments");
    return H2HIP_EINVAL;
  }
  h2hip_combined *C = new h2hip_combined;
  C->f1 = f1;
  C->A1 = A1;
  C->c1[0] = c1[0];
  C->c1[1] = c1[1];
  C->f2 = f2;
  C->A2 = A2;
  if (f2) {
    C->c2[0] = c2[0];
    C->c2[1] = c2[1];
  }
  *out = C;
  return H2HIP_OK;
}

extern "C" void h2hip_combined_destroy(h2hip_combined *C) { delete C; }

extern "C" int h2hip_addeval_combined(const double *alpha, void *A,
                                      const void *x, void *y, void *stream) {
  const h2hip_combined *C = (const h2hip_combined *)A;
  if (!alpha || !C) return H2HIP_EINVAL;
  // alpha c (complex products)
  const double a1[2] = {alpha[0] * C->c1[0] - alpha[1] * C->c1[1],
                        alpha[0] * C->c1[1] + alpha[1] * C->c1[0]};
  int rc = C->f1(a1, C->A1, x, y, stream);
  if (rc || !C->f2) return rc;
  const double a2[2] = {alpha[0] * C->c2[0] - alpha[1] * C->c2[1],
                        alpha[0] * C->c2[1] + alpha[1] * C->c2[0]};
  return C->f2(a2, C->A2, x, y, stream);
}
