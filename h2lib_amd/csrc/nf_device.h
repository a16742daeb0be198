// Device code of the near-field quadrature (included by the per-family
// translation units nf_laplace.hip and nf_helmholtz.hip).
//
// Two kernels per operator, run in this order on one stream:
//
//  nf_regular_kernel  one workgroup (four wave64s) per tile of at most
//      64 x 64 DOFs.  Every pair (row support triangle, column support
//      triangle) without common vertices is evaluated by one lane with the
//      tensor-product Gauss rule from quadrature points precomputed per
//      triangle (16 kernel evaluations at q = 2, all from registers).  The
//      element matrix (1x1, 1x3, 3x1 or 3x3) is added into an LDS tile that
//      is stored once, coalesced; the constant-basis case stores entries
//      directly.  Pairs with common vertices are skipped (cc: stored as 0).
//
//  nf_singular_kernel one row of 16 lanes per pair with common vertices,
//      one launch per rule class (identical, edge, vertex) over the pair
//      list built at planning time.  The 2 q2^4 or 6 q2^4 Sauter-Schwab
//      points are formed on the fly from the q2 Gauss nodes (no rule tables
//      are read), the sums are reduced across the row and up to nine lanes
//      add the element matrix into the output with FP64 atomics.  x - y is
//      formed from differences of the permuted vertices, so the shared
//      vertices cancel exactly.
#pragma once

#include <algorithm>

#include "dmath.h"
#include "h2hip_internal.h"

namespace h2hip {
namespace nf {

constexpr int NTHREADS = 256;
constexpr int NWAVES = NTHREADS / 64;

enum { K_LAP_SLP = 0, K_LAP_DLP = 1, K_HELM_SLP = 2, K_HELM_DLP = 3 };

// 1/sqrt(x): hardware estimate and one third-order correction
// y (1 + e/2 + 3 e^2/8), e = 1 - x y^2.  Positive finite x only (x = 0
// gives NaN where 1/sqrt gives inf; both are non-finite).
__device__ __forceinline__ double rsqrt_fast(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Kernel function values without the 1/(4 pi) factor.
// Laplace SLP 1/r, DLP <d, n_y>/r^3 (Library/laplacebem3d.c:35-108);
// Helmholtz SLP e^{ikr}/r, DLP e^{ikr} (1 - ikr) <d, n_y>/r^3
// (Library/helmholtzbem3d.c:26-170), d = x - y, r = |d|,
// k = kr + i ki' with ki = -ki' (damping e^{ki r}).
template <int KER>
__device__ __forceinline__ void kval(double dx, double dy, double dz,
                                     double nx, double ny, double nz,
                                     double kr, double ki, double &re,
                                     double &im) {
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rs = rsqrt_fast(r2);
  if constexpr (KER == K_LAP_SLP) {
    re = rs;
    im = 0.0;
  } else if constexpr (KER == K_LAP_DLP) {
    re = fma(dx, nx, fma(dy, ny, dz * nz)) * (rs * rs * rs);
    im = 0.0;
  } else if constexpr (KER == K_HELM_SLP) {
    const double r = r2 * rs;
    double amp = rs;
    if (ki != 0.0) amp *= exp(ki * r);
    double s, c;
    h2hip::sincos_cw(kr * r, s, c);
    re = amp * c;
    im = amp * s;
  } else {
    const double r = r2 * rs;
    double amp = fma(dx, nx, fma(dy, ny, dz * nz)) * (rs * rs * rs);
    if (ki != 0.0) amp *= exp(ki * r);
    const double kr_r = kr * r;
    double s, c;
    h2hip::sincos_cw(kr_r, s, c);
    re = fma(kr_r, s, c) * amp;
    im = fma(-c, kr_r, s) * amp;
  }
}

// The same kernel values in FP32 (mixed precision), from a difference
// vector d that the caller formed in FP64 (singular pairs) or from an FP64
// difference of anchor points plus FP32 offsets (regular pairs).  1/r is
// the hardware reciprocal square root (1 ulp); the phase k r is reduced to
// [-1/2, 1/2] revolutions before the hardware sine and cosine, so their
// argument is always small; the damping e^{ki r} is the hardware 2^x of
// ki log2(e) r.
template <int KER>
__device__ __forceinline__ void kval32(float dx, float dy, float dz, float nx,
                                       float ny, float nz, float krf,
                                       float krev, float kil2, float &re,
                                       float &im) {
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const float rs = __builtin_amdgcn_rsqf(r2);
  if constexpr (KER == K_LAP_SLP) {
    re = rs;
    im = 0.0f;
  } else if constexpr (KER == K_LAP_DLP) {
    re = fmaf(dx, nx, fmaf(dy, ny, dz * nz)) * (rs * rs * rs);
    im = 0.0f;
  } else {
    const float r = r2 * rs;
    float rev = r * krev;
    rev -= __builtin_rintf(rev);
    const float s = __builtin_amdgcn_sinf(rev);
    const float c = __builtin_amdgcn_cosf(rev);
    float amp;
    if constexpr (KER == K_HELM_SLP)
      amp = rs;
    else
      amp = fmaf(dx, nx, fmaf(dy, ny, dz * nz)) * (rs * rs * rs);
    if (kil2 != 0.0f) amp *= __builtin_amdgcn_exp2f(kil2 * r);
    if constexpr (KER == K_HELM_SLP) {
      re = amp * c;
      im = amp * s;
    } else {
      const float kr_r = krf * r;
      re = fmaf(kr_r, s, c) * amp;
      im = fmaf(-c, kr_r, s) * amp;
    }
  }
}

// Normal vector of the double layer kernels for the pair (row triangle
// rt, column triangle ct): n_y of the column triangle (double layer), or
// -n_x of the row triangle (adjoint double layer, <y - x, n_x> =
// <x - y, -n_x>; the reference's integrators pass the row triangle's
// normal as n_x also where its lc far rule swaps the point roles,
// Library/bem3d.c:2637-2649).
__device__ __forceinline__ void layer_normal(const KArgs &a, uint32_t rt,
                                             uint32_t ct, double (&nv)[3]) {
  const uint32_t t = a.adj ? rt : ct;
  const double sg = a.adj ? -1.0 : 1.0;
#pragma unroll
  for (int d = 0; d < 3; d++) nv[d] = sg * a.n[3 * t + d];
}

static __global__ void zero_counter(uint32_t *c) { *c = 0u; }

__device__ __forceinline__ uint32_t loc_of(uint32_t packed, int l) {
  return (packed >> (8 * l)) & 0xffu;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Kahan's compensated sum s += x in FP32 (c carries the rounding error of
// the previous additions; the compiler must not reassociate, and does not
// without fast-math)
__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// One tile of the regular kernel, run by all NTHREADS threads of a
// workgroup.  MP: 0 FP64; 1 mixed (FP32 kernel values and inner sums, FP64
// outer sums); 2 mixed with FP32 accumulation throughout, the outer sums
// compensated (Kahan).
template <int KER, int BR, int BC, int CLS, int Q2, bool MP>
__device__ __forceinline__ void tile_singular(const KArgs &a, const Tile &T,
                                              const SPair *list, uint32_t n,
                                              double *acc_re, double *acc_im);

// FUSED: the tile's workgroup also evaluates the tile's pairs with common
// vertices (tlist + T.sbeg, T.ns[c] per class) before the tile is stored,
// so they add into the LDS accumulator instead of the output (constant
// basis: the regular pass leaves their entries to the singular pass).
template <int KER, int BR, int BC, int Q, int MP, bool FUSED = false>
__device__ __forceinline__ void regular_tile(const KArgs &a, uint32_t tile,
                                             const SPair *tlist = nullptr) {
  constexpr bool CPLX = KER >= K_HELM_SLP;
  constexpr bool DLP = (KER == K_LAP_DLP) || (KER == K_HELM_DLP);
  constexpr bool CC = (BR == 0) && (BC == 0);
  constexpr int NR = BR ? 3 : 1;
  constexpr int NC = BC ? 3 : 1;
  constexpr int QQ = Q * Q;

  __shared__ double acc_re[CC ? 1 : TILE * TILE];
  __shared__ double acc_im[(CC || !CPLX) ? 1 : TILE * TILE];

  const int tid = threadIdx.x;
  const Tile T = a.tiles[tile];
  const uint32_t nR = T.nrsup, nC = T.ncsup;
  const Supp *RS = a.supp + T.rsup;
  const Supp *CS = a.supp + T.csup;
  const double *__restrict__ W = a.regw;  // w | phi0 w | phi1 w | phi2 w
  const bool conj = T.ntrans && (BR == BC);

  if constexpr (!CC) {
    for (int i = tid; i < TILE * TILE; i += NTHREADS) {
      acc_re[i] = 0.0;
      if constexpr (CPLX) acc_im[i] = 0.0;
    }
    __syncthreads();
  }

  const uint32_t npairs = nR * nC;
  // pair p = r + nR c, advanced by NTHREADS without a division per pair
  const uint32_t dc = nR ? NTHREADS / nR : 0, dr = NTHREADS - dc * nR;
  uint32_t rr = nR ? tid % nR : 0, cc = nR ? tid / nR : 0;
  for (uint32_t p = tid; p < npairs; p += NTHREADS) {
    const uint32_t r = rr, c = cc;
    rr += dr;
    cc += dc;
    if (rr >= nR) {
      rr -= nR;
      cc++;
    }
    const Supp sr = RS[r], sc = CS[c];
    int shared = 0;
    if (!a.far) {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) shared += (sr.v[i] == sc.v[j]);
    }
    double Sre[NR][NC], Sim[NR][NC];
#pragma unroll
    for (int l = 0; l < NR; l++)
#pragma unroll
      for (int k = 0; k < NC; k++) Sre[l][k] = Sim[l][k] = 0.0;
    double f2 = 0.0;
    if (MP && !shared) {
      // Mixed precision: x_i - y_b = (P_t,i - P_s,0) - (P_s,b - P_s,0), the
      // first difference in FP64 (then rounded), the offsets of the column
      // points in FP32; kernel values and the sums over i in FP32, the sums
      // over b in FP64.
      const bool swapgeo = (BR && !BC) && a.far;
      const uint32_t tt = swapgeo ? sc.tri : sr.tri;
      const uint32_t ts = swapgeo ? sr.tri : sc.tri;
      const double *__restrict__ Pt = a.regpts + (size_t)tt * (3 * QQ);
      const double *__restrict__ Ps = a.regpts + (size_t)ts * (3 * QQ);
      const float *__restrict__ Os = a.regoff + (size_t)ts * (3 * QQ);
      const float *__restrict__ WF = a.regwf;
      const double sx0 = Ps[0], sy0 = Ps[1], sz0 = Ps[2];
      float ptx[QQ], pty[QQ], ptz[QQ];
#pragma unroll
      for (int i = 0; i < QQ; i++) {
        ptx[i] = (float)(Pt[3 * i] - sx0);
        pty[i] = (float)(Pt[3 * i + 1] - sy0);
        ptz[i] = (float)(Pt[3 * i + 2] - sz0);
      }
      float nx = 0.0f, ny = 0.0f, nz = 0.0f;
      if constexpr (DLP) {
        double n64[3];
        layer_normal(a, sr.tri, sc.tri, n64);
        nx = (float)n64[0];
        ny = (float)n64[1];
        nz = (float)n64[2];
      }
      f2 = a.kc * a.g[sr.tri] * a.g[sc.tri];
      // MP 2: the outer sums in FP32 with their compensations
      float Fre[NR][NC], Fim[NR][NC], Cre[NR][NC], Cim[NR][NC];
#pragma unroll
      for (int l = 0; l < NR; l++)
#pragma unroll
        for (int k = 0; k < NC; k++)
          Fre[l][k] = Fim[l][k] = Cre[l][k] = Cim[l][k] = 0.0f;
#pragma unroll
      for (int b = 0; b < QQ; b++) {
        const float qx = Os[3 * b], qy = Os[3 * b + 1], qz = Os[3 * b + 2];
        float Ure[NR], Uim[NR];
#pragma unroll
        for (int l = 0; l < NR; l++) Ure[l] = Uim[l] = 0.0f;
#pragma unroll
        for (int i = 0; i < QQ; i++) {
          float vre, vim;
          kval32<KER>(ptx[i] - qx, pty[i] - qy, ptz[i] - qz, nx, ny, nz,
                      a.krf, a.krev, a.kil2, vre, vim);
#pragma unroll
          for (int l = 0; l < NR; l++) {
            const float wr = BR ? WF[QQ * (1 + l) + i] : WF[i];
            Ure[l] = fmaf(wr, vre, Ure[l]);
            if constexpr (CPLX) Uim[l] = fmaf(wr, vim, Uim[l]);
          }
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
          if constexpr (MP == 2) {
            const float wc = BC ? WF[QQ * (1 + k) + b] : WF[b];
#pragma unroll
            for (int l = 0; l < NR; l++) {
              kahan_add(Fre[l][k], Cre[l][k], wc * Ure[l]);
              if constexpr (CPLX) kahan_add(Fim[l][k], Cim[l][k], wc * Uim[l]);
            }
          } else {
            const double wc = BC ? W[QQ * (1 + k) + b] : W[b];
#pragma unroll
            for (int l = 0; l < NR; l++) {
              Sre[l][k] = fma(wc, (double)Ure[l], Sre[l][k]);
              if constexpr (CPLX)
                Sim[l][k] = fma(wc, (double)Uim[l], Sim[l][k]);
            }
          }
        }
      }
      if constexpr (MP == 2) {
#pragma unroll
        for (int l = 0; l < NR; l++)
#pragma unroll
          for (int k = 0; k < NC; k++) {
            Sre[l][k] = (double)Fre[l][k] - (double)Cre[l][k];
            Sim[l][k] = (double)Fim[l][k] - (double)Cim[l][k];
          }
      }
    } else if (!MP && !shared) {
      // The reference's lc far integrator maps the first rule coordinates
      // onto the column triangle (Library/bem3d.c:2637-2649); kept as is.
      const bool swapgeo = (BR && !BC) && a.far;
      const double *__restrict__ Pt =
          a.regpts + (size_t)(swapgeo ? sc.tri : sr.tri) * (3 * QQ);
      const double *__restrict__ Ps =
          a.regpts + (size_t)(swapgeo ? sr.tri : sc.tri) * (3 * QQ);
      double ptx[QQ], pty[QQ], ptz[QQ];
#pragma unroll
      for (int i = 0; i < QQ; i++) {
        ptx[i] = Pt[3 * i];
        pty[i] = Pt[3 * i + 1];
        ptz[i] = Pt[3 * i + 2];
      }
      double nx = 0.0, ny = 0.0, nz = 0.0;
      if constexpr (DLP) {
        double n64[3];
        layer_normal(a, sr.tri, sc.tri, n64);
        nx = n64[0];
        ny = n64[1];
        nz = n64[2];
      }
      f2 = a.kc * a.g[sr.tri] * a.g[sc.tri];
#pragma unroll
      for (int b = 0; b < QQ; b++) {
        const double qx = Ps[3 * b], qy = Ps[3 * b + 1], qz = Ps[3 * b + 2];
        double Ure[NR], Uim[NR];
#pragma unroll
        for (int l = 0; l < NR; l++) Ure[l] = Uim[l] = 0.0;
#pragma unroll
        for (int i = 0; i < QQ; i++) {
          double vre, vim;
          kval<KER>(ptx[i] - qx, pty[i] - qy, ptz[i] - qz, nx, ny, nz, a.kr,
                    a.ki, vre, vim);
#pragma unroll
          for (int l = 0; l < NR; l++) {
            const double wr = BR ? W[QQ * (1 + l) + i] : W[i];
            Ure[l] = fma(wr, vre, Ure[l]);
            if constexpr (CPLX) Uim[l] = fma(wr, vim, Uim[l]);
          }
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
          const double wc = BC ? W[QQ * (1 + k) + b] : W[b];
#pragma unroll
          for (int l = 0; l < NR; l++) {
            Sre[l][k] = fma(wc, Ure[l], Sre[l][k]);
            if constexpr (CPLX) Sim[l][k] = fma(wc, Uim[l], Sim[l][k]);
          }
        }
      }
    }
    if constexpr (CC) {
      // entry (r, c) of the tile; pairs with common vertices are stored as
      // zero and completed by the singular kernel (fused: stored by the
      // singular pass below)
      if (FUSED && shared) continue;
      const size_t pos = T.out_off + (T.ntrans ? (size_t)c + (size_t)r * T.ld
                                               : (size_t)r + (size_t)c * T.ld);
      if constexpr (CPLX) {
        a.out[2 * pos] = f2 * Sre[0][0];
        a.out[2 * pos + 1] = conj ? -(f2 * Sim[0][0]) : f2 * Sim[0][0];
      } else {
        a.out[pos] = f2 * Sre[0][0];
      }
    } else if (!shared) {
#pragma unroll
      for (int l = 0; l < NR; l++) {
        const uint32_t lr = loc_of(sr.loc, l);
        if (lr == 0xffu) continue;
#pragma unroll
        for (int k = 0; k < NC; k++) {
          const uint32_t lc = loc_of(sc.loc, k);
          if (lc == 0xffu) continue;
          const int e = lr + TILE * lc;
          atomicAdd(&acc_re[e], f2 * Sre[l][k]);
          if constexpr (CPLX) atomicAdd(&acc_im[e], f2 * Sim[l][k]);
        }
      }
    }
  }

  if constexpr (FUSED) {
    if (!a.far) {
      constexpr int Q2 = 4;
      constexpr bool SMP = MP != 0;
      const SPair *L = tlist + T.sbeg;
      tile_singular<KER, BR, BC, CLS_IDENTICAL, Q2, SMP>(a, T, L, T.ns[0],
                                                         acc_re, acc_im);
      L += T.ns[0];
      tile_singular<KER, BR, BC, CLS_EDGE, Q2, SMP>(a, T, L, T.ns[1], acc_re,
                                                    acc_im);
      L += T.ns[1];
      tile_singular<KER, BR, BC, CLS_VERTEX, Q2, SMP>(a, T, L, T.ns[2],
                                                      acc_re, acc_im);
    }
  }
  if constexpr (!CC) {
    __syncthreads();
    const int nr = T.nr, nc = T.nc;
    for (int e = tid; e < nr * nc; e += NTHREADS) {
      const int i = e % nr, j = e / nr;
      const size_t pos =
          T.out_off + (T.ntrans ? (size_t)j + (size_t)i * T.ld
                                : (size_t)i + (size_t)j * T.ld);
      const int s = i + TILE * j;
      if constexpr (CPLX) {
        a.out[2 * pos] = acc_re[s];
        a.out[2 * pos + 1] = conj ? -acc_im[s] : acc_im[s];
      } else {
        a.out[pos] = acc_re[s];
      }
    }
  }
}

template <int KER, int BR, int BC, int Q, int MP>
__global__ __launch_bounds__(NTHREADS) void nf_regular_kernel(KArgs a) {
  regular_tile<KER, BR, BC, Q, MP>(a, blockIdx.x);
}

// Persistent form: a grid sized to the resident workgroups takes tiles from
// a counter (zeroed before the launch) until all ntiles are taken; every
// workgroup leaves the loop once the counter has passed ntiles.
template <int KER, int BR, int BC, int Q, int MP>
__global__ __launch_bounds__(NTHREADS) void nf_regular_persistent(
    KArgs a, uint32_t ntiles, uint32_t *counter) {
  __shared__ uint32_t cur;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = cur;
    // every thread has read cur (and finished the previous tile's LDS
    // stores) before thread 0 takes the next one
    __syncthreads();
    if (t >= ntiles) break;
    regular_tile<KER, BR, BC, Q, MP>(a, t);
  }
}

// Regular and singular pairs of each tile in one workgroup (singular rule
// order 4): one workgroup per tile, or, persistent, a resident grid taking
// tiles from a counter as nf_regular_persistent does.  tlist: the singular
// pairs grouped by tile.
template <int KER, int BR, int BC, int Q, int MP>
__global__ __launch_bounds__(NTHREADS) void nf_fused_kernel(
    KArgs a, const SPair *__restrict__ tlist) {
  regular_tile<KER, BR, BC, Q, MP, true>(a, blockIdx.x, tlist);
}

template <int KER, int BR, int BC, int Q, int MP>
__global__ __launch_bounds__(NTHREADS) void nf_fused_persistent(
    KArgs a, const SPair *__restrict__ tlist, uint32_t ntiles,
    uint32_t *counter) {
  __shared__ uint32_t cur;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = cur;
    __syncthreads();
    if (t >= ntiles) break;
    regular_tile<KER, BR, BC, Q, MP, true>(a, t, tlist);
  }
}

// Sauter-Schwab points of one tensor tuple (g_i, g_j, g_k, g_l) with
// weights (w_i, w_j, w_k, w_l): the 2 (vertex) or 6 (edge, identical)
// subregion points (tx, sx; ty, sy) and their common weight, exactly as the
// tables of the reference are built (Library/singquad2d.c:80-281).
template <int CLS>
struct SSTuple;

template <>
struct SSTuple<CLS_VERTEX> {
  static constexpr int N = 2;
  template <class F>
  __device__ __forceinline__ static void run(double gi, double gj, double gk,
                                             double gl, double wi, double wj,
                                             double wk, double wl, F &&f) {
    const double e1 = gi, e2 = gj, e3 = gk, xi = gl;
    const double w = wi * wj * wk * wl * xi * xi * xi * e1;
    const double a = e2 * xi, b = e1 * xi, c = e1 * e3 * xi;
    f(w, xi, a, b, c);
    f(w, b, c, xi, a);
  }
};

template <>
struct SSTuple<CLS_EDGE> {
  static constexpr int N = 6;
  template <class F>
  __device__ __forceinline__ static void run(double gi, double gj, double gk,
                                             double gl, double wi, double wj,
                                             double wk, double wl, F &&f) {
    const double e1 = gi, e2 = gj, x1 = gk;
    const double x2 = gl * (1.0 - x1);
    const double w = wi * wj * wk * (wl * (1.0 - x1)) * x1 * x1 * e1;
    const double s = x1 + x2;
    const double p = x1 * e1, pq = x1 * e1 * e2;
    f(w, s, pq, x1 * (1.0 - e1) + x2, x1 * (1.0 - e1));
    f(w, s, p, x1 * (1.0 - e1 * e2) + x2, x1 * (1.0 - e1 * e2));
    f(w, s, x1, p + x2, pq);
    f(w, x1 * (1.0 - e1 + e1 * e2) + x2, pq, s, x1);
    f(w, pq + x2, pq, s, p);
    f(w, p + x2, p, s, pq);
  }
};

template <>
struct SSTuple<CLS_IDENTICAL> {
  static constexpr int N = 6;
  template <class F>
  __device__ __forceinline__ static void run(double gi, double gj, double gk,
                                             double gl, double wi, double wj,
                                             double wk, double wl, F &&f) {
    const double eta = gi, x1 = gj;
    const double x2 = gk * (1.0 - x1);
    const double x3 = gl * (1.0 - x1 - x2);
    const double w = wi * wj * (wk * (1.0 - x1)) * (wl * (1.0 - x1 - x2)) * x1;
    const double a = x1 + x2 + x3;
    const double b = x1 * (1.0 - eta);
    f(w, a, x1 + x2, b + x2 + x3, x2);
    f(w, b + x2 + x3, b + x2, a, x2);
    f(w, a, x1 * eta + x2, x2 + x3, x2);
    f(w, b + x2 + x3, x2, a, x1 + x2);
    f(w, a, x2, b + x2 + x3, b + x2);
    f(w, x2 + x3, x2, a, x1 * eta + x2);
  }
};

// sum over the 16 lanes of a row
__device__ __forceinline__ double row_sum(double v) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
  return v;
}

// element i (0..2) of a three-vector by selects (a dynamic index would put
// the vector in private memory)
template <class V>
__device__ __forceinline__ V sel3(const V (&v)[3], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// One pair with common vertices per row of 16 lanes.  Lane s of the row
// takes the tensor index pairs (k, l) = (s / q2, s % q2), (s + 16) / q2, ...
// and loops over (i, j); the quadrature points are formed on the fly from
// the Gauss nodes (a.gauss = [g | w], q2 each), so no rule tables are read.
// The row's lanes sum their moments; lanes e < NR NC then hand the entry
// (lr, lc) of the tile (tile-local row and column) to sink(lr, lc, re, im):
// global atomics for the launches per rule class, the tile's LDS
// accumulator (or a plain store, constant basis) when the tile's own
// workgroup evaluates its singular pairs.  The value is not conjugated.
template <int KER, int BR, int BC, int CLS, int Q2, bool MP, class Sink>
__device__ __forceinline__ void singular_pair(const KArgs &a, const Tile &T,
                                              const SPair &P, int sub,
                                              Sink &&sink) {
  constexpr bool CPLX = KER >= K_HELM_SLP;
  constexpr bool DLP = (KER == K_LAP_DLP) || (KER == K_HELM_DLP);
  constexpr bool CC = (BR == 0) && (BC == 0);
  constexpr int NR = BR ? 3 : 1;
  constexpr int NC = BC ? 3 : 1;

  const Supp sr = a.supp[T.rsup + P.r], sc = a.supp[T.csup + P.c];
  int tp[3], sp[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    tp[i] = (P.perm >> (2 * i)) & 3;
    sp[i] = (P.perm >> (6 + 2 * i)) & 3;
  }
  uint32_t tv[3], sv[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    tv[i] = sel3(sr.v, tp[i]);
    sv[i] = sel3(sc.v, sp[i]);
  }
  double D0[3], E1[3], E2[3], F1[3], F2[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double At = a.x[3 * tv[0] + d], Bt = a.x[3 * tv[1] + d],
                 Ct = a.x[3 * tv[2] + d];
    const double As = a.x[3 * sv[0] + d], Bs = a.x[3 * sv[1] + d],
                 Cs = a.x[3 * sv[2] + d];
    D0[d] = At - As;
    E1[d] = Bt - At;
    E2[d] = Ct - Bt;
    F1[d] = Bs - As;
    F2[d] = Cs - Bs;
  }
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if constexpr (DLP) {
    double n64[3];
    layer_normal(a, sr.tri, sc.tri, n64);
    nx = n64[0];
    ny = n64[1];
    nz = n64[2];
  }
  double Sre[NR][NC], Sim[NR][NC];
#pragma unroll
  for (int l = 0; l < NR; l++)
#pragma unroll
    for (int k = 0; k < NC; k++) Sre[l][k] = Sim[l][k] = 0.0;

  // Moments M[a][b] = sum w v m_a(x) m_b(y) with m = (1, tx, sx) resp.
  // (1, ty, sy); the linear basis functions are phi = (1 - tx, tx - sx, sx)
  // = A m, so the element matrix A M A^T is formed once after the sum.
  // kernel value at the difference d (FP64; mixed precision: d in FP64,
  // the value in FP32)
  auto value = [&](const double *d, double &vre, double &vim) {
    if constexpr (MP) {
      float fre, fim;
      kval32<KER>((float)d[0], (float)d[1], (float)d[2], (float)nx,
                  (float)ny, (float)nz, a.krf, a.krev, a.kil2, fre, fim);
      vre = fre;
      vim = fim;
    } else {
      kval<KER>(d[0], d[1], d[2], nx, ny, nz, a.kr, a.ki, vre, vim);
    }
  };
  auto point = [&](double wq, double tx, double sx, double ty, double sy) {
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
      d[k] = fma(-F2[k], sy,
                 fma(-F1[k], ty, fma(E2[k], sx, fma(E1[k], tx, D0[k]))));
    double vre, vim;
    value(d, vre, vim);
    double ure[NR], uim[NR];
    ure[0] = wq * vre;
    uim[0] = wq * vim;
    if constexpr (BR) {
      ure[1] = ure[0] * tx;
      ure[2] = ure[0] * sx;
      uim[1] = uim[0] * tx;
      uim[2] = uim[0] * sx;
    }
#pragma unroll
    for (int l = 0; l < NR; l++) {
      Sre[l][0] += ure[l];
      if constexpr (CPLX) Sim[l][0] += uim[l];
      if constexpr (BC) {
        Sre[l][1] = fma(ure[l], ty, Sre[l][1]);
        Sre[l][2] = fma(ure[l], sy, Sre[l][2]);
        if constexpr (CPLX) {
          Sim[l][1] = fma(uim[l], ty, Sim[l][1]);
          Sim[l][2] = fma(uim[l], sy, Sim[l][2]);
        }
      }
    }
  };

  if constexpr (Q2 > 0) {
    // order known at compile time (Q2^2 <= 16): one (k, l) per lane, the
    // nodes in registers, the (i, j) loops unrolled
    static_assert(Q2 * Q2 <= 16, "one tensor index pair per lane");
    double gn[Q2], gw[Q2];
#pragma unroll
    for (int i = 0; i < Q2; i++) {
      gn[i] = a.gauss[i];
      gw[i] = a.gauss[Q2 + i];
    }
    if (sub < Q2 * Q2) {
      const int k = sub / Q2, l = sub % Q2;
      const double gk = a.gauss[k], gl = a.gauss[l];
      const double wk = a.gauss[Q2 + k], wl = a.gauss[Q2 + l];
      if constexpr (CLS == CLS_VERTEX) {
        // The two points of a tuple (SSTuple<CLS_VERTEX>) with e1 = g_i,
        // e2 = g_j, e3 = g_k, xi = g_l:
        //   P1 (tx, sx, ty, sy) = (xi, e2 xi, e1 xi, e1 e3 xi)
        //   P2 (tx, sx, ty, sy) = (e1 xi, e1 e3 xi, xi, e2 xi)
        // so d = D0 + xi (E1 + E2 e2 - (F1 + F2 e3) e1) for P1 and
        // d = D0 + xi ((E1 + E2 e3) e1 - F1 - F2 e2) for P2: affine in e2
        // for fixed i.  In P1 the column moments (1, ty, sy) do not depend
        // on j and the row moments (1, tx, sx) are affine in e2 (P2 the
        // other way round), so the sums over j need only sum u and sum u e2
        // per point, and the moment matrix is updated once per i.
        const double e3 = gk, xi = gl;
        const double wkl = wk * wl * xi * xi * xi;
        double a1[3], b1[3], c1[3], a2[3], b2[3], c2[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          a1[d] = fma(xi, E1[d], D0[d]);
          b1[d] = xi * E2[d];
          c1[d] = xi * fma(F2[d], e3, F1[d]);
          a2[d] = fma(-xi, F1[d], D0[d]);
          b2[d] = xi * fma(E2[d], e3, E1[d]);
          c2[d] = -xi * F2[d];
        }
#pragma unroll
        for (int i = 0; i < Q2; i++) {
          const double e1 = gn[i];
          const double wi1 = wkl * gw[i] * e1;
          double t1[3], t2[3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            t1[d] = fma(-c1[d], e1, a1[d]);
            t2[d] = fma(b2[d], e1, a2[d]);
          }
          double U0r = 0.0, U0i = 0.0, U2r = 0.0, U2i = 0.0;
          double V0r = 0.0, V0i = 0.0, V2r = 0.0, V2i = 0.0;
#pragma unroll
          for (int j = 0; j < Q2; j++) {
            const double e2 = gn[j], w = wi1 * gw[j];
            double d1[3], d2[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
              d1[d] = fma(b1[d], e2, t1[d]);
              d2[d] = fma(c2[d], e2, t2[d]);
            }
            double v1r, v1i, v2r, v2i;
            value(d1, v1r, v1i);
            value(d2, v2r, v2i);
            const double u1r = w * v1r, u2r = w * v2r;
            U0r += u1r;
            U2r = fma(u1r, e2, U2r);
            V0r += u2r;
            V2r = fma(u2r, e2, V2r);
            if constexpr (CPLX) {
              const double u1i = w * v1i, u2i = w * v2i;
              U0i += u1i;
              U2i = fma(u1i, e2, U2i);
              V0i += u2i;
              V2i = fma(u2i, e2, V2i);
            }
          }
          // P1: row moments (U0, xi U0, xi U2), column moments
          // (1, e1 xi, e1 e3 xi); P2: row moments (1, e1 xi, e1 e3 xi),
          // column moments (V0, xi V0, xi V2)
          const double f1 = e1 * xi, f2 = e1 * e3 * xi;
          const double ms[3] = {1.0, f1, f2};
          const double p1r[3] = {U0r, xi * U0r, xi * U2r};
          const double p1i[3] = {U0i, xi * U0i, xi * U2i};
          const double p2r[3] = {V0r, xi * V0r, xi * V2r};
          const double p2i[3] = {V0i, xi * V0i, xi * V2i};
#pragma unroll
          for (int l = 0; l < NR; l++)
#pragma unroll
            for (int kk = 0; kk < NC; kk++) {
              Sre[l][kk] = fma(p1r[l], ms[kk], Sre[l][kk]);
              Sre[l][kk] = fma(ms[l], p2r[kk], Sre[l][kk]);
              if constexpr (CPLX) {
                Sim[l][kk] = fma(p1i[l], ms[kk], Sim[l][kk]);
                Sim[l][kk] = fma(ms[l], p2i[kk], Sim[l][kk]);
              }
            }
        }
      } else if constexpr (CLS == CLS_EDGE) {
        // The six points of SSTuple<CLS_EDGE> with e1 = g_i, e2 = g_j,
        // x1 = g_k, x2 = g_l (1 - x1), s = x1 + x2, p = x1 e1, pq = p e2:
        //   (tx, sx, ty, sy) = (s, pq, s - p, x1 - p), (s, p, s - pq, x1 - pq),
        //   (s, x1, p + x2, pq), (s - p + pq, pq, s, x1), (pq + x2, pq, s, p),
        //   (p + x2, p, s, pq).
        // The pair's first two (permuted) vertices are the shared ones, so
        // D0 = 0 and E1 = F1: d = E1 (tx - ty) + E2 sx - F2 sy is affine in
        // pq for fixed i.  In every point one side's moments are fixed for
        // fixed i and the other side's affine in pq, so the sums over j need
        // sum u and sum u pq per point and the moment matrix is updated once
        // per i.
        const double x1 = gk, x2 = gl * (1.0 - x1), s_ = x1 + x2;
        const double wkl = wk * (wl * (1.0 - x1)) * x1 * x1;
        // d = base + cp p + cq pq per point m
        double cA[3], cB[3], cC[3], Gv[3], Hv[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          cA[d] = -F2[d] * x1;
          cB[d] = (E1[d] + E2[d]) * x1;
          cC[d] = -E1[d] * x1;
          Gv[d] = E1[d] + F2[d];
          Hv[d] = E1[d] + E2[d];
        }
#pragma unroll 1
        for (int i = 0; i < Q2; i++) {
          const double e1 = gn[i];
          const double wi1 = wkl * gw[i] * e1;
          const double p = x1 * e1;
          double bs[6][3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            bs[0][d] = fma(Gv[d], p, cA[d]);
            bs[1][d] = fma(E2[d], p, cA[d]);
            bs[2][d] = fma(-E1[d], p, cB[d]);
            bs[3][d] = fma(-E1[d], p, cA[d]);
            bs[4][d] = fma(-F2[d], p, cC[d]);
            bs[5][d] = fma(Hv[d], p, cC[d]);
          }
          double W0r[6], W1r[6], W0i[6], W1i[6];
#pragma unroll
          for (int m = 0; m < 6; m++) W0r[m] = W1r[m] = W0i[m] = W1i[m] = 0.0;
#pragma unroll
          for (int j = 0; j < Q2; j++) {
            const double pq = p * gn[j], w = wi1 * gw[j];
#pragma unroll
            for (int m = 0; m < 6; m++) {
              double dd[3];
#pragma unroll
              for (int d = 0; d < 3; d++) {
                const double cq = m == 0 ? E2[d]
                                : m == 1 ? Gv[d]
                                : m == 2 ? -F2[d]
                                : m == 5 ? -F2[d]
                                         : Hv[d];
                dd[d] = fma(cq, pq, bs[m][d]);
              }
              double vr, vi;
              value(dd, vr, vi);
              const double ur = w * vr;
              W0r[m] += ur;
              W1r[m] = fma(ur, pq, W1r[m]);
              if constexpr (CPLX) {
                const double ui = w * vi;
                W0i[m] += ui;
                W1i[m] = fma(ui, pq, W1i[m]);
              }
            }
          }
          // row-affine points (1, 4, 5): M += T (x) ms; column-affine
          // points (2, 3, 6): M += mt (x) S
          const double ms1[3] = {1.0, s_ - p, x1 - p};
          const double ms4[3] = {1.0, s_, x1};
          const double ms5[3] = {1.0, s_, p};
          const double mt2[3] = {1.0, s_, p};
          const double mt3[3] = {1.0, s_, x1};
          const double mt6[3] = {1.0, p + x2, p};
          auto accum = [&](const double *W0, const double *W1, double (&Sx)[NR][NC]) {
            const double T1[3] = {W0[0], s_ * W0[0], W1[0]};
            const double S2[3] = {W0[1], s_ * W0[1] - W1[1],
                                  x1 * W0[1] - W1[1]};
            const double S3[3] = {W0[2], (p + x2) * W0[2], W1[2]};
            const double T4[3] = {W0[3], (s_ - p) * W0[3] + W1[3], W1[3]};
            const double T5[3] = {W0[4], x2 * W0[4] + W1[4], W1[4]};
            const double S6[3] = {W0[5], s_ * W0[5], W1[5]};
#pragma unroll
            for (int l = 0; l < NR; l++)
#pragma unroll
              for (int kk = 0; kk < NC; kk++) {
                double t = Sx[l][kk];
                t = fma(T1[l], ms1[kk], t);
                t = fma(mt2[l], S2[kk], t);
                t = fma(mt3[l], S3[kk], t);
                t = fma(T4[l], ms4[kk], t);
                t = fma(T5[l], ms5[kk], t);
                t = fma(mt6[l], S6[kk], t);
                Sx[l][kk] = t;
              }
          };
          accum(W0r, W1r, Sre);
          if constexpr (CPLX) accum(W0i, W1i, Sim);
        }
      } else {
        // The six points of SSTuple<CLS_IDENTICAL> with eta = g_i,
        // x1 = g_j, x2 = g_k (1 - x1), x3 = g_l (1 - x1 - x2),
        // a = x1 + x2 + x3, b = x1 (1 - eta):
        //   (tx, sx, ty, sy) = (a, x1 + x2, b + x2 + x3, x2),
        //   (b + x2 + x3, b + x2, a, x2), (a, x1 eta + x2, x2 + x3, x2),
        //   (b + x2 + x3, x2, a, x1 + x2), (a, x2, b + x2 + x3, b + x2),
        //   (x2 + x3, x2, a, x1 eta + x2).
        // Both triangles are the same (D0 = 0, E1 = F1, E2 = F2), so
        // d = E1 (tx - ty) + E2 (sx - sy) = x1 u with
        // u = eta E1 + E2, -eta E1 + (1 - eta) E2, E1 + eta E2 and their
        // negatives: the difference is formed without cancellation.  For
        // fixed j one side's moments are fixed and the other's affine in b,
        // so the sums over i need sum u and sum u b per point.
#pragma unroll 1
        for (int j = 0; j < Q2; j++) {
          const double x1 = gn[j];
          const double x2 = gk * (1.0 - x1);
          const double x3 = gl * (1.0 - x1 - x2);
          const double aa = x1 + x2 + x3;
          const double wj1 = gw[j] * (wk * (1.0 - x1)) *
                             (wl * (1.0 - x1 - x2)) * x1;
          double W0r[6], W1r[6], W0i[6], W1i[6];
#pragma unroll
          for (int m = 0; m < 6; m++) W0r[m] = W1r[m] = W0i[m] = W1i[m] = 0.0;
#pragma unroll
          for (int i = 0; i < Q2; i++) {
            const double eta = gn[i], w = gw[i] * wj1;
            const double bb = x1 * (1.0 - eta);
            double dv[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
              dv[0][d] = x1 * fma(eta, E1[d], E2[d]);
              dv[1][d] = x1 * fma(1.0 - eta, E2[d], -eta * E1[d]);
              dv[2][d] = x1 * fma(eta, E2[d], E1[d]);
            }
#pragma unroll
            for (int m = 0; m < 6; m++) {
              double dd[3];
#pragma unroll
              for (int d = 0; d < 3; d++)
                dd[d] = m < 3 ? dv[m][d] : -dv[m - 3][d];
              double vr, vi;
              value(dd, vr, vi);
              // point order: P1 = +u1, P2 = +u2, P3 = +u3, P4 = -u1,
              // P5 = -u2, P6 = -u3
              const double ur = w * vr;
              W0r[m] += ur;
              W1r[m] = fma(ur, bb, W1r[m]);
              if constexpr (CPLX) {
                const double ui = w * vi;
                W0i[m] += ui;
                W1i[m] = fma(ui, bb, W1i[m]);
              }
            }
          }
          // m = 0: P1, 1: P2, 2: P3, 3: P4, 4: P5, 5: P6
          const double x23 = x2 + x3, x12 = x1 + x2;
          const double mt1[3] = {1.0, aa, x12};
          const double ms2[3] = {1.0, aa, x2};
          const double ms3[3] = {1.0, x23, x2};
          const double ms4[3] = {1.0, aa, x12};
          const double mt5[3] = {1.0, aa, x2};
          const double mt6[3] = {1.0, x23, x2};
          auto accum = [&](const double *W0, const double *W1,
                           double (&Sx)[NR][NC]) {
            const double S1[3] = {W0[0], fma(x23, W0[0], W1[0]), x2 * W0[0]};
            const double T2[3] = {W0[1], fma(x23, W0[1], W1[1]),
                                  fma(x2, W0[1], W1[1])};
            const double T3[3] = {W0[2], aa * W0[2], fma(x12, W0[2], -W1[2])};
            const double T4[3] = {W0[3], fma(x23, W0[3], W1[3]), x2 * W0[3]};
            const double S5[3] = {W0[4], fma(x23, W0[4], W1[4]),
                                  fma(x2, W0[4], W1[4])};
            const double S6[3] = {W0[5], aa * W0[5], fma(x12, W0[5], -W1[5])};
#pragma unroll
            for (int l = 0; l < NR; l++)
#pragma unroll
              for (int kk = 0; kk < NC; kk++) {
                double t = Sx[l][kk];
                t = fma(mt1[l], S1[kk], t);
                t = fma(T2[l], ms2[kk], t);
                t = fma(T3[l], ms3[kk], t);
                t = fma(T4[l], ms4[kk], t);
                t = fma(mt5[l], S5[kk], t);
                t = fma(mt6[l], S6[kk], t);
                Sx[l][kk] = t;
              }
          };
          accum(W0r, W1r, Sre);
          if constexpr (CPLX) accum(W0i, W1i, Sim);
        }
      }
    }
  } else {
    const uint32_t q2 = a.q2;
    const double *__restrict__ gn = a.gauss;
    const double *__restrict__ gw = a.gauss + q2;
    for (uint32_t kl = sub; kl < q2 * q2; kl += 16) {
      const uint32_t k = kl / q2, l = kl - k * q2;
      const double gk = gn[k], gl = gn[l], wk = gw[k], wl = gw[l];
      for (uint32_t i = 0; i < q2; i++) {
        const double gi = gn[i], wi = gw[i];
        for (uint32_t j = 0; j < q2; j++)
          SSTuple<CLS>::run(gi, gn[j], gk, gl, wi, gw[j], wk, wl, point);
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NR; l++)
#pragma unroll
    for (int k = 0; k < NC; k++) {
      Sre[l][k] = row_sum(Sre[l][k]);
      if constexpr (CPLX) Sim[l][k] = row_sum(Sim[l][k]);
    }
  // moments -> element matrix: (1, t, s) -> (1 - t, t - s, s) on each side
  if constexpr (BR) {
#pragma unroll
    for (int k = 0; k < NC; k++) {
      Sre[0][k] -= Sre[1][k];
      Sre[1][k] -= Sre[2][k];
      if constexpr (CPLX) {
        Sim[0][k] -= Sim[1][k];
        Sim[1][k] -= Sim[2][k];
      }
    }
  }
  if constexpr (BC) {
#pragma unroll
    for (int l = 0; l < NR; l++) {
      Sre[l][0] -= Sre[l][1];
      Sre[l][1] -= Sre[l][2];
      if constexpr (CPLX) {
        Sim[l][0] -= Sim[l][1];
        Sim[l][1] -= Sim[l][2];
      }
    }
  }
  if (sub >= NR * NC) return;
  // lane e hands over the entry of permuted local functions (l, k) =
  // (e % NR, e / NR)
  const int l = sub % NR, k = sub / NR;
  double vre = 0.0, vim = 0.0;
#pragma unroll
  for (int ll = 0; ll < NR; ll++)
#pragma unroll
    for (int kk = 0; kk < NC; kk++)
      if (ll == l && kk == k) {
        vre = Sre[ll][kk];
        vim = Sim[ll][kk];
      }
  const uint32_t lr = loc_of(sr.loc, BR ? sel3(tp, l) : 0);
  const uint32_t lc = loc_of(sc.loc, BC ? sel3(sp, k) : 0);
  if (lr == 0xffu || lc == 0xffu) return;
  const double gt = a.g[sr.tri];
  const double f2 = a.kc * gt * a.g[sc.tri];
  vre *= f2;
  vim *= f2;
  if constexpr (CLS == CLS_IDENTICAL) {
    // identity term of the double layer on identical triangles (tp, sp are
    // the identity here): cc at (t, t) of the block
    // (Library/bem3d.c:1234-1241), otherwise Re(alpha) g mass[l, k]
    // (Library/bem3d.c:3205-3242, 1934-1966, 2530-2562)
    if constexpr (CC) {
      if (T.row0 + lr == T.col0 + lc) {
        vre += 0.5 * a.are * gt;
        vim += 0.5 * a.aim * gt;
      }
    } else {
      // the linear-basis integrators hold alpha g in a real variable
      // (Library/bem3d.c:1822-1823, 2418-2419, 3061-3062): only Re(alpha)
      // reaches the mass term
      const bool allowed = (BR && !BC) ? (T.flags & 1) : true;
      if (allowed) {
        const double m = (BR && BC) ? a.mass[l + 3 * k] : a.mass[BR ? l : k];
        vre += m * a.are * gt;
      }
    }
  }
  sink(lr, lc, vre, vim);
}

// The pairs with common vertices of one rule class, one pair per row of 16
// lanes (16 pairs per workgroup), added into the output with FP64 atomics
// (the regular kernel has stored the tiles before).
template <int KER, int BR, int BC, int CLS, int Q2, bool MP>
__global__ __launch_bounds__(NTHREADS) void nf_singular_kernel(
    KArgs a, const SPair *__restrict__ list, uint32_t npairs) {
  constexpr bool CPLX = KER >= K_HELM_SLP;
  const uint32_t w = blockIdx.x * (NTHREADS / 16) + (threadIdx.x >> 4);
  const int sub = threadIdx.x & 15;
  if (w >= npairs) return;  // uniform per row of 16 lanes
  const SPair P = list[w];
  const Tile T = a.tiles[P.tile];
  const bool conj = T.ntrans && (BR == BC);
  singular_pair<KER, BR, BC, CLS, Q2, MP>(
      a, T, P, sub, [&](uint32_t lr, uint32_t lc, double vre, double vim) {
        const size_t pos =
            T.out_off + (T.ntrans ? (size_t)lc + (size_t)lr * T.ld
                                  : (size_t)lr + (size_t)lc * T.ld);
        if constexpr (CPLX) {
          atomicAdd(&a.out[2 * pos], vre);
          atomicAdd(&a.out[2 * pos + 1], conj ? -vim : vim);
        } else {
          atomicAdd(&a.out[pos], vre);
        }
      });
}

// The pairs with common vertices of one tile and one rule class, by the
// tile's workgroup: 16 pairs per pass, the entries added into the tile's
// LDS accumulator (constant basis: stored, each entry belongs to one pair).
template <int KER, int BR, int BC, int CLS, int Q2, bool MP>
__device__ __forceinline__ void tile_singular(const KArgs &a, const Tile &T,
                                              const SPair *list, uint32_t n,
                                              double *acc_re,
                                              double *acc_im) {
  constexpr bool CPLX = KER >= K_HELM_SLP;
  constexpr bool CC = (BR == 0) && (BC == 0);
  const uint32_t row = threadIdx.x >> 4;
  const int sub = threadIdx.x & 15;
  const bool conj = T.ntrans && (BR == BC);
  for (uint32_t p0 = 0; p0 < n; p0 += NTHREADS / 16) {
    const uint32_t w = p0 + row;
    if (w >= n) continue;  // uniform per row of 16 lanes
    singular_pair<KER, BR, BC, CLS, Q2, MP>(
        a, T, list[w], sub,
        [&](uint32_t lr, uint32_t lc, double vre, double vim) {
          if constexpr (CC) {
            const size_t pos =
                T.out_off + (T.ntrans ? (size_t)lc + (size_t)lr * T.ld
                                      : (size_t)lr + (size_t)lc * T.ld);
            if constexpr (CPLX) {
              a.out[2 * pos] = vre;
              a.out[2 * pos + 1] = conj ? -vim : vim;
            } else {
              a.out[pos] = vre;
            }
          } else {
            const uint32_t e = lr + TILE * lc;
            atomicAdd(&acc_re[e], vre);
            if constexpr (CPLX) atomicAdd(&acc_im[e], vim);
          }
        });
  }
}

// The launches of one precision (MP: 0 FP64, 1 mixed, 2 mixed with
// compensated FP32 outer sums; the singular kernels are the same for 1 and
// 2).
template <int KER, int BR, int BC, int MP>
int launch_bases_p(uint32_t q, const KArgs &a, uint32_t ntiles,
                   const SPair *sp, const uint32_t *ncls, hipStream_t stream) {
  if (a.tlist && !a.far && q == 2 && a.q2 == 4) {
    // fused tiles: one launch for the regular and the singular pairs
    if (!ntiles) return H2HIP_OK;
    if (a.counter) {
      auto fk = nf_fused_persistent<KER, BR, BC, 2, MP>;
      int per_cu = 0, dev = 0, ncu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fk, NTHREADS,
                                                       0) != hipSuccess ||
          hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                dev) != hipSuccess)
        return H2HIP_EDEVICE;
      const uint32_t resident = (uint32_t)std::max(1, per_cu) * (uint32_t)ncu;
      const uint32_t grid = std::min(ntiles, std::max(1u, resident));
      hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, stream, a.counter);
      hipLaunchKernelGGL(fk, dim3(grid), dim3(NTHREADS), 0, stream, a,
                         a.tlist, ntiles, a.counter);
    } else {
      hipLaunchKernelGGL((nf_fused_kernel<KER, BR, BC, 2, MP>), dim3(ntiles),
                         dim3(NTHREADS), 0, stream, a, a.tlist);
    }
    return H2HIP_OK;
  }
  if (ntiles && a.counter) {
    void (*reg)(KArgs, uint32_t, uint32_t *) = nullptr;
    switch (q) {
      case 1: reg = nf_regular_persistent<KER, BR, BC, 1, MP>; break;
      case 2: reg = nf_regular_persistent<KER, BR, BC, 2, MP>; break;
      case 3: reg = nf_regular_persistent<KER, BR, BC, 3, MP>; break;
      case 4: reg = nf_regular_persistent<KER, BR, BC, 4, MP>; break;
      default: return H2HIP_EINVAL;
    }
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reg, NTHREADS,
                                                     0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess)
      return H2HIP_EDEVICE;
    const uint32_t resident = (uint32_t)std::max(1, per_cu) * (uint32_t)ncu;
    const uint32_t grid = std::min(ntiles, std::max(1u, resident));
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, stream, a.counter);
    hipLaunchKernelGGL(reg, dim3(grid), dim3(NTHREADS), 0, stream, a, ntiles,
                       a.counter);
  } else if (ntiles) {
    void (*reg)(KArgs) = nullptr;
    switch (q) {
      case 1: reg = nf_regular_kernel<KER, BR, BC, 1, MP>; break;
      case 2: reg = nf_regular_kernel<KER, BR, BC, 2, MP>; break;
      case 3: reg = nf_regular_kernel<KER, BR, BC, 3, MP>; break;
      case 4: reg = nf_regular_kernel<KER, BR, BC, 4, MP>; break;
      default: return H2HIP_EINVAL;
    }
    hipLaunchKernelGGL(reg, dim3(ntiles), dim3(NTHREADS), 0, stream, a);
  }
  if (a.far) return H2HIP_OK;
  constexpr uint32_t PPB = NTHREADS / 16;  // pairs per workgroup
  const uint32_t n3 = ncls[CLS_IDENTICAL], n2 = ncls[CLS_EDGE],
                 n1 = ncls[CLS_VERTEX];
  using SK = void (*)(KArgs, const SPair *, uint32_t);
  const bool q4 = a.q2 == 4;
  constexpr bool SMP = MP != 0;
  const SK k3 = q4 ? nf_singular_kernel<KER, BR, BC, CLS_IDENTICAL, 4, SMP>
                   : nf_singular_kernel<KER, BR, BC, CLS_IDENTICAL, 0, SMP>;
  const SK k2 = q4 ? nf_singular_kernel<KER, BR, BC, CLS_EDGE, 4, SMP>
                   : nf_singular_kernel<KER, BR, BC, CLS_EDGE, 0, SMP>;
  const SK k1 = q4 ? nf_singular_kernel<KER, BR, BC, CLS_VERTEX, 4, SMP>
                   : nf_singular_kernel<KER, BR, BC, CLS_VERTEX, 0, SMP>;
  if (n3)
    hipLaunchKernelGGL(k3, dim3((n3 + PPB - 1) / PPB), dim3(NTHREADS), 0,
                       stream, a, sp, n3);
  if (n2)
    hipLaunchKernelGGL(k2, dim3((n2 + PPB - 1) / PPB), dim3(NTHREADS), 0,
                       stream, a, sp + n3, n2);
  if (n1)
    hipLaunchKernelGGL(k1, dim3((n1 + PPB - 1) / PPB), dim3(NTHREADS), 0,
                       stream, a, sp + n3 + n2, n1);
  return H2HIP_OK;
}

// Launches the regular tile kernel over `ntiles` tiles and then (unless
// far) one singular launch per rule class; sp holds the pairs sorted
// identical, edge, vertex with ncls[c] pairs of class c.  a.mp selects the
// mixed-precision kernels.
template <int KER, int BR, int BC>
int launch_bases(uint32_t q, const KArgs &a, uint32_t ntiles,
                 const SPair *sp, const uint32_t *ncls, hipStream_t stream) {
  if (a.mp == 2)
    return launch_bases_p<KER, BR, BC, 2>(q, a, ntiles, sp, ncls, stream);
  if (a.mp == 1)
    return launch_bases_p<KER, BR, BC, 1>(q, a, ntiles, sp, ncls, stream);
  return launch_bases_p<KER, BR, BC, 0>(q, a, ntiles, sp, ncls, stream);
}

template <int KER>
int launch_family(int rb, int cb, uint32_t q, const KArgs &a,
                  uint32_t ntiles, const SPair *sp, const uint32_t *ncls,
                  hipStream_t stream) {
  if (!rb && !cb) return launch_bases<KER, 0, 0>(q, a, ntiles, sp, ncls, stream);
  if (rb && cb) return launch_bases<KER, 1, 1>(q, a, ntiles, sp, ncls, stream);
  if (cb) return launch_bases<KER, 0, 1>(q, a, ntiles, sp, ncls, stream);
  return launch_bases<KER, 1, 0>(q, a, ntiles, sp, ncls, stream);
}

}  // namespace nf
}  // namespace h2hip
