// Restarted GMRES on device vectors.
//
// Semantics of the reference's solve_gmres_avector
// (Library/krylovsolvers.c:183-227) with init/step/finish_gmres
// (Library/krylov.c:612-770): the Krylov space is restarted when its
// dimension k reaches kmax - 1, the iteration stops when the least-squares
// residual |rhat_k| <= eps |b| or after maxiter - 1 steps, and the solution
// is updated at every restart and at the end.  The reference orthogonalises
// with Householder reflections; here the Arnoldi vectors are
// orthogonalised by classical Gram-Schmidt with one reorthogonalisation
// (two passes, each one multi-dot kernel and one update kernel), which
// spans the same Krylov spaces and agrees to rounding.  All inner products
// are FP64; the small Hessenberg least-squares problem (Givens rotations)
// is solved on the host.
#include <cmath>
#include <complex>
#include <vector>

#include "h2hip_internal.h"

namespace {

constexpr int NT = 256;
constexpr int NB = 240;  // blocks of the reduction kernels

template <int C>
struct Val {
  double re, im;
};

// partial[b * m + j] = sum over the block's share of conj(Q_j) . w
template <int C>
__global__ __launch_bounds__(NT) void multidot(const double *__restrict__ Q,
                                               uint64_t ldq, int m,
                                               const double *__restrict__ w,
                                               uint64_t n,
                                               double *__restrict__ part) {
  __shared__ double red[NT / 64][2 * 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j0 = 0; j0 < m; j0 += 32) {
    const int mj = min(32, m - j0);
    double acc_re[32], acc_im[32];
#pragma unroll
    for (int j = 0; j < 32; j++) acc_re[j] = acc_im[j] = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * NT) {
      double wr, wi = 0.0;
      if constexpr (C) {
        wr = w[2 * i];
        wi = w[2 * i + 1];
      } else {
        wr = w[i];
      }
#pragma unroll
      for (int j = 0; j < 32; j++) {
        if (j < mj) {
          const uint64_t q = (uint64_t)(j0 + j) * ldq + i;
          if constexpr (C) {
            const double qr = Q[2 * q], qi = Q[2 * q + 1];
            acc_re[j] = fma(qr, wr, fma(qi, wi, acc_re[j]));
            acc_im[j] = fma(qr, wi, fma(-qi, wr, acc_im[j]));
          } else {
            acc_re[j] = fma(Q[q], wr, acc_re[j]);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 32; j++) {
      if (j < mj) {
        double r = acc_re[j], im = acc_im[j];
        for (int off = 32; off > 0; off >>= 1) {
          r += __shfl_xor(r, off, 64);
          if constexpr (C) im += __shfl_xor(im, off, 64);
        }
        if (lane == 0) {
          red[wv][2 * j] = r;
          red[wv][2 * j + 1] = im;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < 2 * mj) {
      double s = 0.0;
      for (int v = 0; v < NT / 64; v++) s += red[v][threadIdx.x];
      part[((uint64_t)blockIdx.x * 64 + (j0 + threadIdx.x / 2)) * 2 +
           (threadIdx.x & 1)] = s;
    }
    __syncthreads();
  }
}

// w -= Q h (h: m coefficients, complex pairs when C)
template <int C>
__global__ __launch_bounds__(NT) void multiaxpy(const double *__restrict__ Q,
                                                uint64_t ldq, int m,
                                                const double *__restrict__ h,
                                                double *__restrict__ w,
                                                uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  double ar = 0.0, ai = 0.0;
  for (int j = 0; j < m; j++) {
    const uint64_t q = (uint64_t)j * ldq + i;
    if constexpr (C) {
      const double qr = Q[2 * q], qi = Q[2 * q + 1];
      const double hr = h[2 * j], hi = h[2 * j + 1];
      ar = fma(qr, hr, fma(-qi, hi, ar));
      ai = fma(qr, hi, fma(qi, hr, ai));
    } else {
      ar = fma(Q[q], h[j], ar);
    }
  }
  if constexpr (C) {
    w[2 * i] -= ar;
    w[2 * i + 1] -= ai;
  } else {
    w[i] -= ar;
  }
}

// dst = s * src (real scale)
__global__ void scale_copy(const double *src, double *dst, double s,
                           uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = s * src[i];
}

__global__ void copy_kernel(const double *src, double *dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

using cd = std::complex<double>;

struct Workspace {
  int C;
  uint64_t n, ldq;  // ldq in elements (complex: pairs)
  int kmax;
  h2hip::DevBuf<double> Q, w, part, h;
  std::vector<double> hpart;
  hipStream_t st;
};

// h[0..m) = Q[:, 0..m)^* w, sum of the block partials on the host
template <int C>
int mdot(Workspace &W, int m, const double *w, std::vector<cd> &out) {
  hipLaunchKernelGGL(multidot<C>, dim3(NB), dim3(NT), 0, W.st, W.Q.p, W.ldq,
                     m, w, W.n, W.part.p);
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipMemcpyAsync(W.hpart.data(), W.part.p,
                             sizeof(double) * NB * 64 * 2,
                             hipMemcpyDeviceToHost, W.st));
  H2HIP_CHECK(hipStreamSynchronize(W.st));
  out.assign(m, cd(0.0, 0.0));
  for (int b = 0; b < NB; b++)
    for (int j = 0; j < m; j++)
      out[j] += cd(W.hpart[((uint64_t)b * 64 + j) * 2],
                   W.hpart[((uint64_t)b * 64 + j) * 2 + 1]);
  return H2HIP_OK;
}

template <int C>
int maxpy(Workspace &W, int m, const std::vector<cd> &h, double *w) {
  std::vector<double> hh(2 * m);
  for (int j = 0; j < m; j++) {
    hh[C ? 2 * j : j] = h[j].real();
    if (C) hh[2 * j + 1] = h[j].imag();
  }
  H2HIP_CHECK(hipMemcpyAsync(W.h.p, hh.data(), sizeof(double) * (C ? 2 : 1) * m,
                             hipMemcpyHostToDevice, W.st));
  hipLaunchKernelGGL(multiaxpy<C>, dim3((W.n + NT - 1) / NT), dim3(NT), 0,
                     W.st, W.Q.p, W.ldq, m, W.h.p, w, W.n);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// norm of a device vector: dot with itself via the multi-dot kernel
template <int C>
int vnorm(Workspace &W, const double *v, double &nrm) {
  hipLaunchKernelGGL(multidot<C>, dim3(NB), dim3(NT), 0, W.st, v, (uint64_t)0,
                     1, v, W.n, W.part.p);
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipMemcpyAsync(W.hpart.data(), W.part.p,
                             sizeof(double) * NB * 64 * 2,
                             hipMemcpyDeviceToHost, W.st));
  H2HIP_CHECK(hipStreamSynchronize(W.st));
  double s = 0.0;
  for (int b = 0; b < NB; b++) s += W.hpart[(uint64_t)b * 128];
  nrm = std::sqrt(s);
  return H2HIP_OK;
}

// Givens rotation that maps (a, b) to (r, 0)
void givens(cd a, cd b, cd &c, cd &s) {
  const double na = std::abs(a), nb = std::abs(b);
  if (nb == 0.0) {
    c = 1.0;
    s = 0.0;
    return;
  }
  if (na == 0.0) {
    c = 0.0;
    s = std::conj(b) / nb;
    return;
  }
  const double nr = std::hypot(na, nb);
  c = na / nr;
  s = (a / na) * std::conj(b) / nr;
}

template <int C>
int gmres(h2hip_addeval_t addeval, void *A, uint64_t n, const double *b,
          double *x, double eps, uint32_t maxiter, uint32_t kmax,
          hipStream_t st, uint32_t *iters, double *resid) {
  const int cm = C ? 2 : 1;
  Workspace W;
  W.C = C;
  W.n = n;
  W.ldq = n;
  W.kmax = (int)kmax;
  W.st = st;
  if (W.Q.alloc((uint64_t)cm * n * kmax) != hipSuccess ||
      W.w.alloc((uint64_t)cm * n) != hipSuccess ||
      W.part.alloc((uint64_t)NB * 64 * 2) != hipSuccess ||
      W.h.alloc(2 * 64) != hipSuccess) {
    h2hip::set_error("h2hip_solve_gmres: device allocation failed");
    return H2HIP_ENOMEM;
  }
  W.hpart.resize((uint64_t)NB * 64 * 2);
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  const dim3 gv((uint32_t)((cm * n + NT - 1) / NT));
  double normb;
  int rc = vnorm<C>(W, b, normb);
  if (rc) return rc;

  std::vector<cd> g(kmax + 1), cs(kmax), sn(kmax), h, h2;
  std::vector<std::vector<cd>> R(kmax, std::vector<cd>(kmax));
  int k = 0;
  double error = 0.0;

  // r = b - A x in column 0 of Q, normalised
  auto init = [&]() -> int {
    double *q0 = W.Q.p;
    hipLaunchKernelGGL(copy_kernel, gv, dim3(NT), 0, st, b, q0,
                       (uint64_t)cm * n);
    int r = addeval(mone, A, x, q0, st);
    if (r) return r;
    double beta;
    r = vnorm<C>(W, q0, beta);
    if (r) return r;
    if (beta > 0.0)
      hipLaunchKernelGGL(scale_copy, gv, dim3(NT), 0, st, q0, q0, 1.0 / beta,
                         (uint64_t)cm * n);
    std::fill(g.begin(), g.end(), cd(0.0, 0.0));
    g[0] = beta;
    k = 0;
    error = beta;
    return H2HIP_OK;
  };
  // x += Q y with R y = g (upper triangular, k x k)
  auto finish = [&]() -> int {
    if (k == 0) return H2HIP_OK;
    std::vector<cd> y(k);
    for (int i = k - 1; i >= 0; i--) {
      cd s = g[i];
      for (int j = i + 1; j < k; j++) s -= R[i][j] * y[j];
      y[i] = s / R[i][i];
    }
    for (auto &v : y) v = -v;  // maxpy subtracts
    int r = maxpy<C>(W, k, y, x);
    if (r) return r;
    return init();
  };

  rc = init();
  if (rc) return rc;
  uint32_t iter = 0;
  while (error > eps * normb && iter + 1 != maxiter) {
    if (k + 1 >= (int)kmax) {
      rc = finish();
      if (rc) return rc;
    }
    // step: w = A q_k, orthogonalise against q_0 .. q_k (two passes)
    double *qk = W.Q.p + (uint64_t)cm * k * n;
    double *w = W.w.p;
    hipLaunchKernelGGL(scale_copy, gv, dim3(NT), 0, st, qk, w, 0.0,
                       (uint64_t)cm * n);
    rc = addeval(one, A, qk, w, st);
    if (rc) return rc;
    rc = mdot<C>(W, k + 1, w, h);
    if (rc) return rc;
    rc = maxpy<C>(W, k + 1, h, w);
    if (rc) return rc;
    rc = mdot<C>(W, k + 1, w, h2);
    if (rc) return rc;
    rc = maxpy<C>(W, k + 1, h2, w);
    if (rc) return rc;
    for (int j = 0; j <= k; j++) h[j] += h2[j];
    double hn;
    rc = vnorm<C>(W, w, hn);
    if (rc) return rc;
    double *qn = W.Q.p + (uint64_t)cm * (k + 1) * n;
    if (hn > 0.0)
      hipLaunchKernelGGL(scale_copy, gv, dim3(NT), 0, st, w, qn, 1.0 / hn,
                         (uint64_t)cm * n);
    // previous rotations, then the new one
    for (int i = 0; i < k; i++) {
      const cd t = cs[i] * h[i] + sn[i] * h[i + 1];
      h[i + 1] = -std::conj(sn[i]) * h[i] + cs[i] * h[i + 1];
      h[i] = t;
    }
    cd c, s;
    givens(h[k], cd(hn, 0.0), c, s);
    cs[k] = c;
    sn[k] = s;
    const cd rkk = c * h[k] + s * hn;
    for (int i = 0; i < k; i++) R[i][k] = h[i];
    R[k][k] = rkk;
    g[k + 1] = -std::conj(s) * g[k];
    g[k] = c * g[k];
    k++;
    error = std::abs(g[k]);
    iter++;
  }
  rc = finish();
  if (rc) return rc;
  H2HIP_CHECK(hipStreamSynchronize(st));
  if (iters) *iters = iter;
  if (resid) *resid = error;
  return H2HIP_OK;
}

}  // namespace

extern "C" int h2hip_solve_gmres(h2hip_addeval_t addeval, void *A,
                                 int is_complex, uint64_t n, const void *d_b,
                                 void *d_x, double eps, uint32_t maxiter,
                                 uint32_t kmax, void *stream,
                                 uint32_t *iterations, double *residual) {
  if (!addeval || !d_b || !d_x || n == 0 || kmax < 2 || kmax > 64) {
    h2hip::set_error("h2hip_solve_gmres: invalid arguments (kmax 2..64)");
    return H2HIP_EINVAL;
  }
  return is_complex
             ? gmres<1>(addeval, A, n, (const double *)d_b, (double *)d_x, eps,
                        maxiter, kmax, (hipStream_t)stream, iterations,
                        residual)
             : gmres<0>(addeval, A, n, (const double *)d_b, (double *)d_x, eps,
                        maxiter, kmax, (hipStream_t)stream, iterations,
                        residual);
}
