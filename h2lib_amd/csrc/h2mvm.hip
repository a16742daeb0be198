// H2-matrix products y += alpha H x and y += alpha H^* x on gfx950.
//
// The reference evaluates the product by three recursions over its pointer
// trees (addeval_h2matrix_avector, Library/h2matrix.c:1268-1286): forward
// transform of x into the column basis coefficients
// (Library/clusterbasis.c:523-573), coupling and near-field blocks
// (fastaddeval_h2matrix_avector, Library/h2matrix.c:1210-1266), backward
// transform into y (Library/clusterbasis.c:748-797).  Here one product is
// three launches over flattened trees:
//
//  1. fwd_up: one workgroup per leaf computes xt_t = V_t^T x|_t, then climbs
//     the tree: the last son to finish (an atomic counter per cluster, reset
//     by that son, agent-scope fences around it) computes
//     xt_f = sum_s E_s^T xt_s for its father and goes on upwards.  No
//     workgroup ever waits for another one.  The same launch clears the
//     coupling accumulator yt.
//  2. couple_near: one wavefront per admissible block, yt_r += S_rc xt_c
//     (FP64 atomics), and one per near-field block, y|_r += alpha N_rc x|_c.
//  3. down_back: one workgroup per leaf of the output tree walks its path
//     from the root, v = yt_s + E_s v, and adds y|_t += alpha V_t v.
//
// Every small product uses lane = output entry and loops over the input
// entries (in 1. and 3. the four waves of the workgroup take a quarter of
// the inputs each, so that the dependent chain up or down the tree is
// short), with each matrix stored in the orientation that makes that loop
// read 64 consecutive entries per step (leaf and transfer matrices are kept
// in both orientations, couplings and near blocks column-major).  The
// products are bound by HBM bandwidth: two flops per stored entry.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "h2hip_internal.h"
#include "dmath.h"
#include "wave_sum.h"

namespace {

constexpr int NT = 256;
constexpr int WPB = NT / 64;
constexpr uint32_t NONE = 0xffffffffu;
constexpr int MAXRANK = 128;
constexpr int MAXPATH = 48;

struct NodeD {
  uint32_t off, size, rank, father;
  uint64_t coef;  // offset of the node's coefficients
  uint64_t voff;  // offset of the leaf matrix
  uint32_t sonbeg, nsons;
};

struct FarD {
  uint32_t r, c;
  uint64_t soff;
};

struct NearD {
  uint32_t roff, rows, coff, cols;
  uint64_t noff;
};

template <int C>
__device__ __forceinline__ void ld(const double *p, uint64_t i, double &re,
                                   double &im) {
  if constexpr (C) {
    re = p[2 * i];
    im = p[2 * i + 1];
  } else {
    re = p[i];
    im = 0.0;
  }
}

template <int C>
__device__ __forceinline__ void st(double *p, uint64_t i, double re,
                                   double im) {
  if constexpr (C) {
    p[2 * i] = re;
    p[2 * i + 1] = im;
  } else {
    p[i] = re;
  }
}

template <int C>
__device__ __forceinline__ void atomic_add(double *p, uint64_t i, double re,
                                           double im) {
  if constexpr (C) {
    atomicAdd(&p[2 * i], re);
    atomicAdd(&p[2 * i + 1], im);
  } else {
    atomicAdd(&p[i], re);
  }
}

// Vectors of one product call: NV right-hand sides at most, nv in use.
// Coefficient vectors xt / yt hold the vectors one after the other,
// cstride doubles apart.
constexpr int MAXNV = 8;

struct VecSet {
  const double *x[MAXNV];
  double *y[MAXNV];
  double are[MAXNV], aim[MAXNV];
  uint32_t nv;
};

// rank (and leaf size) bound of the LDS vectors: 128 for one vector, 64
// when several vectors share a workgroup
template <int NV>
constexpr int rank_cap() {
  return NV == 1 ? MAXRANK : 64;
}

struct FwdArgs {
  const NodeD *nodes;
  // one entry per workgroup: a leaf, or (bit 31 set) a cluster whose sons
  // are all leaves
  const uint32_t *work;
  uint32_t nwork;
  const uint32_t *idx;
  const uint32_t *sons;
  const double *VT;  // leaf: rank x size column-major
  const double *ET;  // per son s: rank_father x rank_s column-major
  uint64_t eld;
  double *xt;
  uint64_t cstride;  // doubles between the coefficient vectors
  uint32_t *cnt;
  double *yt;  // cleared here (nyt doubles per vector)
  uint64_t nyt;
  VecSet v;
  int climb;  // 0: the clusters above the work items are left to fwd_level
};

// Product of one small column-major matrix M (nout x nin, leading
// dimension ldm, nout <= RK) with the NV LDS vectors sv[v * 2RK + ...],
// split over the four waves of a workgroup: output o = lane + 64 c reads
// M[o + i * ldm] for the wave's quarter of the inputs i.  Wave w leaves its
// partial sums in part[(w NV + v) 2RK + 2o (+1)]; the caller adds the four
// after the closing barrier.
template <int C, int NV, int RK>
__device__ __forceinline__ void wg_matvec(const double *__restrict__ M,
                                          uint32_t ldm, uint32_t nout,
                                          uint32_t nin, const double *sv,
                                          double *part) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t q = (nin + WPB - 1) / WPB;
  const uint32_t i0 = wave * q, i1 = min(nin, i0 + q);
  constexpr int UNR = NV == 1 ? 8 : 2;
#pragma unroll
  for (int c = 0; c < RK / 64; c++) {
    const uint32_t o = lane + 64 * c;
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
    if (o < nout) {
#pragma unroll UNR
      for (uint32_t i = i0; i < i1; i++) {
        const double m = M[o + (uint64_t)i * ldm];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          ar[v] = fma(m, sv[v * 2 * RK + 2 * i], ar[v]);
          if constexpr (C) ai[v] = fma(m, sv[v * 2 * RK + 2 * i + 1], ai[v]);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      part[(wave * NV + v) * 2 * RK + 2 * o] = ar[v];
      part[(wave * NV + v) * 2 * RK + 2 * o + 1] = ai[v];
    }
  }
  __syncthreads();
}

template <int NV, int RK>
__device__ __forceinline__ double part_sum(const double *part, int v,
                                           uint32_t e) {
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < WPB; w++) s += part[(w * NV + v) * 2 * RK + e];
  return s;
}

// xt_t = V_t^T x|_t of leaf t into global memory; with keep != nullptr the
// result also goes to the LDS vectors keep (for a product that follows)
template <int C, int K, int NV>
__device__ __forceinline__ void leaf_forward(const FwdArgs &a,
                                             const NodeD &t, double *sv,
                                             double *part, double *keep) {
  constexpr int RK = rank_cap<NV>();
  const int tid = threadIdx.x;
  const uint32_t rank = K ? K : t.rank;
  // x|_t into LDS (leaf size <= RK is checked before the launch)
  for (uint32_t e = tid; e < NV * t.size; e += NT) {
    const uint32_t v = e / t.size, i = e - v * t.size;
    double xr = 0.0, xi = 0.0;
    if (v < a.v.nv) ld<C>(a.v.x[v], a.idx[t.off + i], xr, xi);
    sv[v * 2 * RK + 2 * i] = xr;
    sv[v * 2 * RK + 2 * i + 1] = xi;
  }
  __syncthreads();
  wg_matvec<C, NV, RK>(a.VT + t.voff, rank, rank, t.size, sv, part);
  for (uint32_t e = tid; e < NV * rank; e += NT) {
    const uint32_t v = e / rank, j = e - v * rank;
    const double sr = part_sum<NV, RK>(part, v, 2 * j);
    const double si = part_sum<NV, RK>(part, v, 2 * j + 1);
    if (v < a.v.nv) st<C>(a.xt + v * a.cstride, t.coef + j, sr, si);
    if (keep) {
      keep[v * 2 * RK + 2 * j] = sr;
      keep[v * 2 * RK + 2 * j + 1] = si;
    }
  }
  __syncthreads();
}

template <int C, int K, int NV>
__global__ __launch_bounds__(NT) void fwd_up(FwdArgs a) {
  constexpr int RK = rank_cap<NV>();
  __shared__ double sv[NV * 2 * RK], sv2[NV * 2 * RK];
  __shared__ double part[WPB * NV * 2 * RK];
  __shared__ uint32_t s_old;
  const uint32_t w = blockIdx.x;
  const int tid = threadIdx.x;
  {  // clear a slice of the coupling accumulators
    const uint64_t chunk = (a.nyt + a.nwork - 1) / a.nwork;
    const uint64_t e0 = (uint64_t)w * chunk, e1 = min(a.nyt, e0 + chunk);
    for (uint32_t v = 0; v < a.v.nv; v++)
      for (uint64_t e = e0 + tid; e < e1; e += NT)
        a.yt[v * a.cstride + e] = 0.0;
  }
  const uint32_t item = a.work[w];
  uint32_t node = item & 0x7fffffffu;
  // sums of the current cluster's coefficients, kept by wave 0: output
  // j = lane + 64 c of vector v
  double tot_re[RK / 64][NV], tot_im[RK / 64][NV];
  auto clear_tot = [&]() {
#pragma unroll
    for (int c = 0; c < RK / 64; c++)
#pragma unroll
      for (int v = 0; v < NV; v++) tot_re[c][v] = tot_im[c][v] = 0.0;
  };
  auto add_tot = [&](uint32_t rank) {
    if (tid < 64) {
#pragma unroll
      for (int c = 0; c < RK / 64; c++) {
        const uint32_t j = tid + 64 * c;
        if (j < rank) {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            tot_re[c][v] += part_sum<NV, RK>(part, v, 2 * j);
            tot_im[c][v] += part_sum<NV, RK>(part, v, 2 * j + 1);
          }
        }
      }
    }
  };
  auto store_tot = [&](const NodeD &F, uint32_t rank) {
    if (tid < 64) {
#pragma unroll
      for (int c = 0; c < RK / 64; c++) {
        const uint32_t j = tid + 64 * c;
#pragma unroll
        for (int v = 0; v < NV; v++)
          if (j < rank && v < (int)a.v.nv)
            st<C>(a.xt + v * a.cstride, F.coef + j, tot_re[c][v],
                  tot_im[c][v]);
      }
    }
  };
  if (item >> 31) {
    // a cluster whose sons are all leaves: the sons and the cluster in this
    // workgroup, without counters
    const NodeD F = a.nodes[node];
    const uint32_t rank = K ? K : F.rank;
    clear_tot();
    for (uint32_t so = 0; so < F.nsons; so++) {
      const uint32_t sn = a.sons[F.sonbeg + so];
      const NodeD t = a.nodes[sn];
      leaf_forward<C, K, NV>(a, t, sv, part, sv2);
      wg_matvec<C, NV, RK>(a.ET + (uint64_t)sn * a.eld, rank, rank,
                           K ? K : t.rank, sv2, part);
      add_tot(rank);
      __syncthreads();
    }
    store_tot(F, rank);
  } else {
    leaf_forward<C, K, NV>(a, a.nodes[node], sv, part, nullptr);
  }
  if (!a.climb) return;
  // climb: the last son of a cluster computes the cluster
  while (true) {
    const uint32_t f = a.nodes[node].father;
    if (f == NONE) break;
    // release this workgroup's coefficients: the barrier completes every
    // wave's stores, one agent-scope fence then publishes them before the
    // counter increment
    __syncthreads();
    if (tid == 0) {
      __threadfence();
      s_old = atomicAdd(&a.cnt[f], 1u);
    }
    __syncthreads();
    const NodeD F = a.nodes[f];
    if (s_old + 1 < F.nsons) break;  // uniform
    // acquire the other sons' coefficients (one fence for the workgroup,
    // the barrier keeps every wave's loads behind it)
    if (tid == 0) {
      __threadfence();
      a.cnt[f] = 0;
    }
    __syncthreads();
    const uint32_t rank = K ? K : F.rank;
    clear_tot();
    for (uint32_t so = 0; so < F.nsons; so++) {
      const uint32_t sn = a.sons[F.sonbeg + so];
      const NodeD s = a.nodes[sn];
      const uint32_t srank = K ? K : s.rank;
      for (uint32_t e = tid; e < NV * srank; e += NT) {
        const uint32_t v = e / srank, i = e - v * srank;
        double xr = 0.0, xi = 0.0;
        if (v < a.v.nv) ld<C>(a.xt + v * a.cstride, s.coef + i, xr, xi);
        sv[v * 2 * RK + 2 * i] = xr;
        sv[v * 2 * RK + 2 * i + 1] = xi;
      }
      __syncthreads();
      wg_matvec<C, NV, RK>(a.ET + (uint64_t)sn * a.eld, rank, rank, srank,
                           sv, part);
      add_tot(rank);
      __syncthreads();
    }
    store_tot(F, rank);
    node = f;
  }
}

// Level-synchronous form of the climb: one workgroup per cluster of one
// height (all its sons lower, computed by earlier launches), xt_f =
// sum_s E_s^T xt_s.  Launched once per height above the work items of
// fwd_up (run with climb = 0); kernel boundaries order the levels, so no
// counters or fences are needed.
template <int C, int K, int NV>
__global__ __launch_bounds__(NT) void fwd_level(FwdArgs a,
                                                const uint32_t *list) {
  constexpr int RK = rank_cap<NV>();
  __shared__ double sv[NV * 2 * RK];
  __shared__ double part[WPB * NV * 2 * RK];
  const int tid = threadIdx.x;
  const NodeD F = a.nodes[list[blockIdx.x]];
  const uint32_t rank = K ? K : F.rank;
  double tot_re[RK / 64][NV], tot_im[RK / 64][NV];
#pragma unroll
  for (int c = 0; c < RK / 64; c++)
#pragma unroll
    for (int v = 0; v < NV; v++) tot_re[c][v] = tot_im[c][v] = 0.0;
  for (uint32_t so = 0; so < F.nsons; so++) {
    const uint32_t sn = a.sons[F.sonbeg + so];
    const NodeD s = a.nodes[sn];
    const uint32_t srank = K ? K : s.rank;
    for (uint32_t e = tid; e < NV * srank; e += NT) {
      const uint32_t v = e / srank, i = e - v * srank;
      double xr = 0.0, xi = 0.0;
      if (v < a.v.nv) ld<C>(a.xt + v * a.cstride, s.coef + i, xr, xi);
      sv[v * 2 * RK + 2 * i] = xr;
      sv[v * 2 * RK + 2 * i + 1] = xi;
    }
    __syncthreads();
    wg_matvec<C, NV, RK>(a.ET + (uint64_t)sn * a.eld, rank, rank, srank, sv,
                         part);
    if (tid < 64) {
#pragma unroll
      for (int c = 0; c < RK / 64; c++) {
        const uint32_t j = tid + 64 * c;
        if (j < rank) {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            tot_re[c][v] += part_sum<NV, RK>(part, v, 2 * j);
            tot_im[c][v] += part_sum<NV, RK>(part, v, 2 * j + 1);
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid < 64) {
#pragma unroll
    for (int c = 0; c < RK / 64; c++) {
      const uint32_t j = tid + 64 * c;
#pragma unroll
      for (int v = 0; v < NV; v++)
        if (j < rank && v < (int)a.v.nv)
          st<C>(a.xt + v * a.cstride, F.coef + j, tot_re[c][v],
                tot_im[c][v]);
    }
  }
}

// One coupling entry of an interpolation H2-matrix from the coordinate
// differences of the two tensor Chebyshev points: kernel_const / r
// (Laplace) or kernel_const e^{i kr r} e^{kdamp r} / r (Helmholtz), as
// assemble_bem3d_inter_uniform fills them (Library/bem3d.c:6079-6108).
// Shared by the kernel that stores the couplings and the product that
// recomputes them, so that both produce the same values.
// 1/sqrt(x) for positive finite x: hardware estimate and one third-order
// correction y (1 + e/2 + 3 e^2/8), e = 1 - x y^2, as in the near-field
// kernels (a fraction of the instructions of a square root and a divide)
__device__ __forceinline__ double cpl_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// The coupling value from the squared distance r2 (cpl_value)
__device__ __forceinline__ void cpl_from_r2(double r2, int helmholtz,
                                            double kr, double kdamp,
                                            double &sr, double &si) {
  constexpr double KCONST = 0.0795774715459476679;
  double rs = cpl_rsqrt(r2);
  if (!helmholtz) {
    sr = KCONST * rs;
    si = 0.0;
  } else {
    const double r = r2 * rs;
    if (kdamp != 0.0) rs *= exp(kdamp * r);
    double sn, cs;
    h2hip::sincos_cw(kr * r, sn, cs);
    sr = KCONST * (rs * cs);
    si = KCONST * (sn * rs);
  }
}

__device__ __forceinline__ void cpl_value(double d0, double d1, double d2,
                                          int helmholtz, double kr,
                                          double kdamp, double &sr,
                                          double &si) {
  // (the order of the products of r2 is fixed: the unrolled product paths
  // form the same sums with the first two hoisted)
  cpl_from_r2(fma(d2, d2, fma(d1, d1, d0 * d0)), helmholtz, kr, kdamp, sr,
              si);
}

// Interpolation data for couplings computed in the product: Chebyshev
// points (3 m per node, x then y then z) of both trees and the kernel.
struct CplSrc {
  const double *rpts, *cpts;
  uint32_t m;
  int helmholtz;
  double kr, kdamp;
};

// A mirrored pair of far blocks (r, c) / (c, r), r > c, for
// couple_mf_pairs: the nodes and the four coefficient offsets, so that a
// wave needs one load before its inputs
struct PairD {
  uint32_t r, c;
  uint64_t out_r, in_c;  // block (r, c): yt of row node r, xt of column node c
  uint64_t out_c, in_r;  // mirror (c, r): yt of row node c, xt of column node r
};

struct CNArgs {
  const FarD *far;
  uint32_t nfar;
  const NearD *near;
  uint32_t nnear;
  const NodeD *rn, *cn;
  const uint32_t *ridx, *cidx;
  const double *S, *N;
  const double *xt;
  double *yt;
  uint64_t cstride;
  VecSet v;
  CplSrc cs;  // MF products: couplings computed here
  // reproducible products: block b's results go to slot pslot[b] of the
  // partial-sum buffer P (far blocks first, then near blocks from pnear0;
  // pstride doubles per vector) instead of being added with atomics
  double *P;
  const uint64_t *pslot;
  uint64_t pstride;
  uint32_t pnear0;
};

// acc += sum_j M_j xs_j over n inputs for a column-major block read along
// a row: M points at the row's first entry, entry j lies stride entries
// further (complex entries as pairs); 16 (complex: 8) entries are loaded
// before their products, so that each wave keeps enough loads in flight.
// NTL: the entries are loaded as streaming (non-temporal) data, which the
// caches need not keep (the couplings: read once per product and more
// bytes than the last-level cache holds)
template <int C>
__device__ __forceinline__ void ld_nt(const double *p, uint64_t i, double &re,
                                      double &im) {
  if constexpr (C) {
    re = __builtin_nontemporal_load(p + 2 * i);
    im = __builtin_nontemporal_load(p + 2 * i + 1);
  } else {
    re = __builtin_nontemporal_load(p + i);
    im = 0.0;
  }
}

template <int C, bool NTL = false>
__device__ __forceinline__ void dot_row(const double *__restrict__ M,
                                        uint64_t stride, uint32_t n,
                                        const double *xs, double &ar,
                                        double &ai, uint32_t j0 = 0) {
  constexpr int B = C ? 8 : 16;
  uint32_t j = j0;
  for (; j + B <= n; j += B) {
    double mr[B], mi[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      if constexpr (NTL)
        ld_nt<C>(M, (uint64_t)(j + u) * stride, mr[u], mi[u]);
      else
        ld<C>(M, (uint64_t)(j + u) * stride, mr[u], mi[u]);
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      const double xr = xs[2 * (j + u)], xi = xs[2 * (j + u) + 1];
      ar = fma(mr[u], xr, ar);
      if constexpr (C) {
        ar = fma(-mi[u], xi, ar);
        ai = fma(mr[u], xi, fma(mi[u], xr, ai));
      }
    }
  }
  if (j < n) {
    // the last n - j < B columns, loaded together (indices past n repeat
    // the last column), then the products in column order
    const uint32_t rem = n - j;
    double mr[B], mi[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const uint64_t e = (uint64_t)(j + min((uint32_t)u, rem - 1)) * stride;
      if constexpr (NTL)
        ld_nt<C>(M, e, mr[u], mi[u]);
      else
        ld<C>(M, e, mr[u], mi[u]);
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      if ((uint32_t)u < rem) {
        const double xr = xs[2 * (j + u)], xi = xs[2 * (j + u) + 1];
        ar = fma(mr[u], xr, ar);
        if constexpr (C) {
          ar = fma(-mi[u], xi, ar);
          ai = fma(mr[u], xi, fma(mi[u], xr, ai));
        }
      }
    }
  }
}

// ar/ai[v] += sum_j M[base + j sj] x_v[j] over n columns for NV vectors in
// LDS (complex pairs when C; CONJ: the conjugate of M; NTL: non-temporal
// loads): B columns' entries are loaded together, the last batch with its
// indices past n repeating the last column, before their products, which
// run in column order.
template <int C, int NV, int XR, bool NTL, bool CONJ>
__device__ __forceinline__ void dot_cols_nv(const double *__restrict__ M,
                                            uint64_t base, uint64_t sj,
                                            uint32_t n, const double *xs,
                                            double (&ar)[NV],
                                            double (&ai)[NV]) {
  constexpr int B = C ? 4 : 8;
  for (uint32_t j0 = 0; j0 < n; j0 += B) {
    const uint32_t rem = min((uint32_t)B, n - j0);
    double mr[B], mi[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const uint64_t e = base + (uint64_t)(j0 + min((uint32_t)u, rem - 1)) * sj;
      if constexpr (NTL)
        ld_nt<C>(M, e, mr[u], mi[u]);
      else
        ld<C>(M, e, mr[u], mi[u]);
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      if ((uint32_t)u < rem) {
        const uint32_t j = j0 + u;
        const double sr = mr[u], si = CONJ ? -mi[u] : mi[u];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double xr = xs[v * 2 * XR + 2 * j];
          const double xi = xs[v * 2 * XR + 2 * j + 1];
          ar[v] = fma(sr, xr, ar[v]);
          if constexpr (C) {
            ar[v] = fma(-si, xi, ar[v]);
            ai[v] = fma(sr, xi, fma(si, xr, ai[v]));
          }
        }
      }
    }
  }
}

// The first columns of a row for dot_row_pf, loaded before the wave's input
// gather so that the two round trips overlap (columns past n repeat the
// last one and are not used)
template <int C>
__device__ __forceinline__ void row_prefetch(const double *__restrict__ M,
                                             uint64_t stride, uint32_t n,
                                             double (&mr)[C ? 8 : 16],
                                             double (&mi)[C ? 8 : 16]) {
  constexpr int B = C ? 8 : 16;
#pragma unroll
  for (int u = 0; u < B; u++)
    ld<C>(M, (uint64_t)min((uint32_t)u, n - 1) * stride, mr[u], mi[u]);
}

// dot_row with columns 0 .. B - 1 from row_prefetch: the same products in
// the same order
template <int C>
__device__ __forceinline__ void dot_row_pf(const double *__restrict__ M,
                                           uint64_t stride, uint32_t n,
                                           const double *xs,
                                           const double (&mr)[C ? 8 : 16],
                                           const double (&mi)[C ? 8 : 16],
                                           double &ar, double &ai) {
  constexpr int B = C ? 8 : 16;
#pragma unroll
  for (int u = 0; u < B; u++) {
    if (u < (int)n) {
      const double xr = xs[2 * u], xi = xs[2 * u + 1];
      ar = fma(mr[u], xr, ar);
      if constexpr (C) {
        ar = fma(-mi[u], xi, ar);
        ai = fma(mr[u], xi, fma(mi[u], xr, ai));
      }
    }
  }
  if (n > (uint32_t)B) dot_row<C>(M, stride, n, xs, ar, ai, B);
}

// Laplace couplings computed in the product at interpolation order 4, one
// output point per lane against the 64 input points, unrolled: the input
// points in registers, the x and y differences and their squares formed
// once per line of points, the same arithmetic as cpl_value (so the values
// are those of the stored couplings) and the same order of the sums as the
// general loop.  The output point is p (3 coordinates), the inputs q
// (x, y, z lines of 4); sg = +1 forms p - q, -1 forms q - p.
template <int C, int NV, int XR>
__device__ __forceinline__ void mf_laplace4(const double *p3,
                                            const double *q, double sg,
                                            const double *xs, double (&ar)[NV],
                                            double (&ai)[NV]) {
  constexpr double KCONST = 0.0795774715459476679;
  double qx[4], qy[4], qz[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    qx[t] = q[t];
    qy[t] = q[4 + t];
    qz[t] = q[8 + t];
  }
#pragma unroll
  for (int ja = 0; ja < 4; ja++) {
    const double d0 = sg * (p3[0] - qx[ja]);
    const double s0 = d0 * d0;
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
      const double d1 = sg * (p3[1] - qy[jb]);
      const double s1 = fma(d1, d1, s0);
#pragma unroll
      for (int jc = 0; jc < 4; jc++) {
        const int j = 16 * ja + 4 * jb + jc;
        const double d2 = sg * (p3[2] - qz[jc]);
        const double sr = KCONST * cpl_rsqrt(fma(d2, d2, s1));
#pragma unroll
        for (int v = 0; v < NV; v++) {
          ar[v] = fma(sr, xs[v * 2 * XR + 2 * j], ar[v]);
          if constexpr (C) ai[v] = fma(sr, xs[v * 2 * XR + 2 * j + 1], ai[v]);
        }
      }
    }
  }
}

// yt_r += S_rc xt_c and y|_r += alpha N_rc x|_c; TRANS: yt_c += S^* xt_r and
// y|_c += alpha N^* x|_r.  One wave per block; the wave first gathers its
// block's input entries into its own LDS slice.
// MF (not transposed): the couplings are computed from the interpolation
// points instead of read (cpl_value), lane i keeping row point i and
// stepping through the column points j = (ja m + jb) m + jc; each value
// serves all NV vectors.
// Item w of the wave (w < nfar: far block w, else near block w - nfar);
// xs: the wave's LDS slice of NV * 2 * XR doubles.  PF: near blocks of one
// vector load their first columns before the input gather (more registers;
// used where the launch holds only near blocks).
template <int C, bool TRANS, int NV, bool MF = false, bool ORD = false,
          bool PF = false>
__device__ __forceinline__ void couple_near_dev(const CNArgs &a, uint32_t w,
                                                double *xs) {
  constexpr int XR = NV == 1 ? MAXRANK : 64;  // inputs per LDS vector
  const int lane = threadIdx.x & 63;
  if (w < a.nfar) {
    const FarD b = a.far[w];
    const NodeD r = a.rn[b.r], c = a.cn[b.c];
    const NodeD out = TRANS ? c : r, in = TRANS ? r : c;
    // ranks <= 128 (one vector) or <= 64 (several)
    for (uint32_t e = lane; e < NV * in.rank; e += 64) {
      const uint32_t v = e / in.rank, j = e - v * in.rank;
      double xr = 0.0, xi = 0.0;
      if (v < a.v.nv) ld<C>(a.xt + v * a.cstride, in.coef + j, xr, xi);
      xs[v * 2 * XR + 2 * j] = xr;
      xs[v * 2 * XR + 2 * j + 1] = xi;
    }
    // the wave reads what its own lanes wrote: LDS keeps the order of
    // one wave's accesses, the compiler must not move them
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < out.rank; i += 64) {
      double ar[NV], ai[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
      bool m4 = false;  // done by the unrolled order-4 Laplace path
      (void)m4;
      if constexpr (MF && TRANS) {
        // out = column point i, inputs = row points j; the entry is the
        // coupling S(j, i) (the same differences row - column as the stored
        // one), conjugated
        const uint32_t m = a.cs.m;
        const double *pr = a.cs.rpts + 3 * (uint64_t)m * b.r;
        const double *pc = a.cs.cpts + 3 * (uint64_t)m * b.c;
        const uint32_t ia = i / (m * m), ib = (i / m) % m, ic = i % m;
        const double qx = pc[ia], qy = pc[m + ib], qz = pc[2 * m + ic];
        uint32_t j = 0;
        if (m == 4 && !a.cs.helmholtz) {  // uniform over the grid
          const double p3[3] = {qx, qy, qz};
          mf_laplace4<C, NV, XR>(p3, pr, -1.0, xs, ar, ai);
          m4 = true;
        }
        for (uint32_t ja = 0; ja < (m4 ? 0u : m); ja++) {
          const double d0 = pr[ja] - qx;
          for (uint32_t jb = 0; jb < m; jb++) {
            const double d1 = pr[m + jb] - qy;
            for (uint32_t jc = 0; jc < m; jc++, j++) {
              double sr, si;
              cpl_value(d0, d1, pr[2 * m + jc] - qz, a.cs.helmholtz,
                        a.cs.kr, a.cs.kdamp, sr, si);
              si = -si;
#pragma unroll
              for (int v = 0; v < NV; v++) {
                const double xr = xs[v * 2 * XR + 2 * j];
                const double xi = xs[v * 2 * XR + 2 * j + 1];
                ar[v] = fma(sr, xr, ar[v]);
                if constexpr (C) {
                  ar[v] = fma(-si, xi, ar[v]);
                  ai[v] = fma(sr, xi, fma(si, xr, ai[v]));
                }
              }
            }
          }
        }
      } else if constexpr (MF) {
        const uint32_t m = a.cs.m;
        const double *pr = a.cs.rpts + 3 * (uint64_t)m * b.r;
        const double *pc = a.cs.cpts + 3 * (uint64_t)m * b.c;
        const uint32_t ia = i / (m * m), ib = (i / m) % m, ic = i % m;
        const double px = pr[ia], py = pr[m + ib], pz = pr[2 * m + ic];
        uint32_t j = 0;
        if (m == 4 && !a.cs.helmholtz) {  // uniform over the grid
          const double p3[3] = {px, py, pz};
          mf_laplace4<C, NV, XR>(p3, pc, 1.0, xs, ar, ai);
          m4 = true;
        }
        for (uint32_t ja = 0; ja < (m4 ? 0u : m); ja++) {
          const double d0 = px - pc[ja];
          for (uint32_t jb = 0; jb < m; jb++) {
            const double d1 = py - pc[m + jb];
            for (uint32_t jc = 0; jc < m; jc++, j++) {
              double sr, si;
              cpl_value(d0, d1, pz - pc[2 * m + jc], a.cs.helmholtz, a.cs.kr,
                        a.cs.kdamp, sr, si);
#pragma unroll
              for (int v = 0; v < NV; v++) {
                const double xr = xs[v * 2 * XR + 2 * j];
                const double xi = xs[v * 2 * XR + 2 * j + 1];
                ar[v] = fma(sr, xr, ar[v]);
                if constexpr (C) {
                  ar[v] = fma(-si, xi, ar[v]);
                  ai[v] = fma(sr, xi, fma(si, xr, ai[v]));
                }
              }
            }
          }
        }
      } else if constexpr (NV == 1 && !TRANS) {
        dot_row<C, true>(a.S + (C ? 2 : 1) * (b.soff + i), r.rank, in.rank,
                         xs, ar[0], ai[0]);
      } else {
        dot_cols_nv<C, NV, XR, true, TRANS>(
            a.S, b.soff + (TRANS ? (uint64_t)i * r.rank : (uint64_t)i),
            TRANS ? 1 : r.rank, in.rank, xs, ar, ai);
      }
#pragma unroll
      for (int v = 0; v < NV; v++)
        if (v < (int)a.v.nv) {
          if constexpr (ORD)
            st<C>(a.P + v * a.pstride, a.pslot[w] + i, ar[v], ai[v]);
          else
            atomic_add<C>(a.yt + v * a.cstride, out.coef + i, ar[v], ai[v]);
        }
    }
    return;
  }
  const uint32_t nb = w - a.nfar;
  if (nb >= a.nnear) return;
  const NearD b = a.near[nb];
  const uint32_t nout = TRANS ? b.cols : b.rows, nin = TRANS ? b.rows : b.cols;
  const uint32_t *oidx = TRANS ? a.cidx + b.coff : a.ridx + b.roff;
  const uint32_t *iidx = TRANS ? a.ridx + b.roff : a.cidx + b.coff;
  // one vector: the first columns of the lane's first row are loaded
  // before the input gather
  constexpr int PB = C ? 8 : 16;
  double pr[PB], pim[PB];
  if constexpr (PF && NV == 1 && !TRANS)
    if (lane < nout && nin > 0)
      row_prefetch<C>(a.N + (C ? 2 : 1) * (b.noff + lane), b.rows, nin, pr,
                      pim);
  // leaf sizes <= 128 (one vector) or <= 64 (several)
  for (uint32_t e = lane; e < NV * nin; e += 64) {
    const uint32_t v = e / nin, j = e - v * nin;
    double xr = 0.0, xi = 0.0;
    if (v < a.v.nv) ld<C>(a.v.x[v], iidx[j], xr, xi);
    xs[v * 2 * XR + 2 * j] = xr;
    xs[v * 2 * XR + 2 * j + 1] = xi;
  }
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = lane; i < nout; i += 64) {
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
    if constexpr (NV == 1 && !TRANS) {
      if (PF && i == (uint32_t)lane)
        dot_row_pf<C>(a.N + (C ? 2 : 1) * (b.noff + i), b.rows, nin, xs, pr,
                      pim, ar[0], ai[0]);
      else
        dot_row<C>(a.N + (C ? 2 : 1) * (b.noff + i), b.rows, nin, xs, ar[0],
                   ai[0]);
    } else {
      dot_cols_nv<C, NV, XR, false, TRANS>(
          a.N, b.noff + (TRANS ? (uint64_t)i * b.rows : (uint64_t)i),
          TRANS ? 1 : b.rows, nin, xs, ar, ai);
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (v < (int)a.v.nv) {
        if constexpr (ORD)
          st<C>(a.P + v * a.pstride, a.pslot[a.pnear0 + nb] + i, ar[v],
                ai[v]);
        else
          atomic_add<C>(a.v.y[v], oidx[i],
                        a.v.are[v] * ar[v] - a.v.aim[v] * ai[v],
                        a.v.are[v] * ai[v] + a.v.aim[v] * ar[v]);
      }
  }
}

template <int C, bool TRANS, int NV, bool MF = false, bool ORD = false>
__global__ __launch_bounds__(NT) void couple_near(CNArgs a) {
  constexpr int XS = NV * 2 * (NV == 1 ? MAXRANK : 64);  // per wave
  __shared__ double xs_all[WPB * XS];
  couple_near_dev<C, TRANS, NV, MF, ORD>(
      a, blockIdx.x * WPB + (threadIdx.x >> 6),
      xs_all + (threadIdx.x >> 6) * XS);
}

// Coupling blocks of a real matrix with all ranks 64 for NV (2..8) vectors
// on FP64 matrix cores: yt_r += S xt_c as 64 v_mfma_f64_16x16x4_f64 per
// block (four 16-row tiles of S times sixteen 4-row steps of the 64 x NV
// inputs, the 16 columns of an MFMA padded beyond NV with zeros).  One wave
// per far block; A operands straight from S (lane l: row l % 16, column
// l / 16 of the 16 x 4 tile), B operands from the wave's LDS copy of the
// inputs (lane l: input row l / 16, vector l % 16), results in lane l
// (register r: row 4 r + l / 16 of the tile, vector l % 16) added to the
// output coefficients.
typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int NV>
__global__ __launch_bounds__(NT) void couple_mfma(CNArgs a) {
  __shared__ double xs_all[WPB][64][16];
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.nfar) return;  // uniform over the wave
  double(*xs)[16] = xs_all[threadIdx.x >> 6];
  const FarD b = a.far[w];
  const NodeD r = a.rn[b.r], c = a.cn[b.c];
  // inputs: xs[k][v] = xt_v[c.coef + k], zero for v >= nv
  for (uint32_t e = lane; e < 64 * 16; e += 64) {
    const uint32_t k = e >> 4, v = e & 15;
    xs[k][v] = (v < a.v.nv && v < (uint32_t)NV)
                   ? a.xt[v * a.cstride + c.coef + k]
                   : 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  const double *S = a.S + b.soff;
  const int li = lane & 15, lk = lane >> 4;
  dbl4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int k0 = 0; k0 < 64; k0 += 8) {
    double av[2][4], bv[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int kk = k0 + 4 * h + lk;
      bv[h] = xs[kk][li];
#pragma unroll
      for (int t = 0; t < 4; t++)
        av[h][t] = __builtin_nontemporal_load(S + (16 * t + li) + 64 * kk);
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h][t], bv[h],
                                                       acc[t], 0, 0, 0);
  }
  if (li >= NV || li >= (int)a.v.nv) return;
  double *yt = a.yt + li * a.cstride + r.coef;
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int q = 0; q < 4; q++)
      atomicAdd(&yt[16 * t + 4 * q + lk], acc[t][q]);
}

// Adjoint blocks (yt_c += S^* xt_r, y|_c += alpha N^* x|_r) with one
// workgroup per block for blocks of at most 64 x 64: the block is staged in
// LDS with contiguous reads (a padded leading dimension keeps the column
// reads of the lanes free of bank conflicts), each wave sums a quarter of
// the inputs for every output (lane = output), and the four partial sums
// are added in wave order.  Far blocks w < nfar, then near blocks.  The
// per-block wave path above reads the adjoint's entries with a stride of a
// column per lane.
template <int C, int NV, bool ORD>
__global__ __launch_bounds__(NT) void couple_trans_wg(CNArgs a) {
  constexpr int LDM = 65;
  __shared__ double Mr[64 * LDM];
  __shared__ double Mi[C ? 64 * LDM : 1];
  __shared__ double xs[NV][2][64];
  __shared__ double part[WPB][NV][2][64];
  const uint32_t w = blockIdx.x;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const bool isfar = w < a.nfar;
  uint32_t nin, nout;
  const double *M;
  FarD fb = {};
  NearD nb = {};
  if (isfar) {
    fb = a.far[w];
    nin = a.rn[fb.r].rank;
    nout = a.cn[fb.c].rank;
    M = a.S + (C ? 2 : 1) * fb.soff;
  } else {
    nb = a.near[w - a.nfar];
    nin = nb.rows;
    nout = nb.cols;
    M = a.N + (C ? 2 : 1) * nb.noff;
  }
  // the block, column-major with leading dimension nin, into LDS
  for (uint32_t e = tid; e < nin * nout; e += NT) {
    const uint32_t i = e % nin, j = e / nin;
    double mr, mi;
    if (isfar)
      ld_nt<C>(M, e, mr, mi);
    else
      ld<C>(M, e, mr, mi);
    Mr[i + j * LDM] = mr;
    if constexpr (C) Mi[i + j * LDM] = mi;
  }
  for (uint32_t e = tid; e < NV * nin; e += NT) {
    const uint32_t v = e / nin, i = e - v * nin;
    double xr = 0.0, xi = 0.0;
    if (v < a.v.nv) {
      if (isfar)
        ld<C>(a.xt + v * a.cstride, a.rn[fb.r].coef + i, xr, xi);
      else
        ld<C>(a.v.x[v], a.ridx[nb.roff + i], xr, xi);
    }
    xs[v][0][i] = xr;
    xs[v][1][i] = xi;
  }
  __syncthreads();
  const uint32_t q = (nin + WPB - 1) / WPB;
  const uint32_t i0 = wv * q, i1 = min(nin, i0 + q);
  if ((uint32_t)lane < nout) {
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
    for (uint32_t i = i0; i < i1; i++) {
      const double mr = Mr[i + lane * LDM];
      const double mi = C ? -Mi[i + lane * LDM] : 0.0;  // conjugate
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double xr = xs[v][0][i], xi = xs[v][1][i];
        ar[v] = fma(mr, xr, ar[v]);
        if constexpr (C) {
          ar[v] = fma(-mi, xi, ar[v]);
          ai[v] = fma(mr, xi, fma(mi, xr, ai[v]));
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      part[wv][v][0][lane] = ar[v];
      part[wv][v][1][lane] = ai[v];
    }
  }
  __syncthreads();
  for (uint32_t e = tid; e < NV * nout; e += NT) {
    const uint32_t v = e / nout, j = e - v * nout;
    if (v >= a.v.nv) continue;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < WPB; k++) {
      sr += part[k][v][0][j];
      si += part[k][v][1][j];
    }
    if constexpr (ORD) {
      st<C>(a.P + v * a.pstride,
            a.pslot[isfar ? w : a.pnear0 + (w - a.nfar)] + j, sr, si);
    } else if (isfar) {
      atomic_add<C>(a.yt + v * a.cstride, a.cn[fb.c].coef + j, sr, si);
    } else {
      atomic_add<C>(a.v.y[v], a.cidx[nb.coff + j],
                    a.v.are[v] * sr - a.v.aim[v] * si,
                    a.v.are[v] * si + a.v.aim[v] * sr);
    }
  }
}

// Couplings computed in the product at interpolation order M = 4 or 5, one
// vector, for a symmetric block structure (row and column trees with the
// same points, every admissible block (r, c) with its mirror (c, r), whose
// coupling matrix is S_rc^T: the squared distances, and so the values of
// 1/r or e^{ikr}/r, are the same bit for bit).  One wave per pair r > c
// evaluates each value once for both blocks: lane l keeps the sums of row
// points l and l + 64 (rank M^3 <= 128) of block (r, c), yt_r[i] += sum_j
// S_ij xt_c[j], with the arithmetic and order of the per-block kernel
// (mf_laplace4, the general loop otherwise), and the products S_ij xt_r[i]
// of W columns at a time, whose sums over the rows, yt_c[j] += sum_i S_ij
// xt_r[i] for the mirror block, a butterfly over the lanes forms.  Half the
// evaluations of the two blocks' waves.  C: Helmholtz (complex).
template <int C, int M>
__global__ __launch_bounds__(NT) void couple_mf_pairs(CNArgs a,
                                                      const PairD *pairs,
                                                      uint32_t npair) {
  constexpr double KCONST = 0.0795774715459476679;
  constexpr int R = M * M * M, RH = (R + 63) / 64;
  // columns per butterfly: the products of a pass stay in registers until
  // then (fewer for the complex values, whose sine and cosine take the
  // registers)
  constexpr int W = C ? (RH > 1 ? 2 : 4) : 16;
  constexpr int NP = (R + W - 1) / W;
  __shared__ double xs_all[WPB][2 * R];
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= npair) return;  // uniform over the wave; no block barrier
  const PairD b = pairs[w];
  double *xs = xs_all[threadIdx.x >> 6];
  for (int e = lane; e < R; e += 64) {
    if constexpr (C) {
      xs[2 * e] = a.xt[2 * (b.in_c + e)];
      xs[2 * e + 1] = a.xt[2 * (b.in_c + e) + 1];
    } else {
      xs[e] = a.xt[b.in_c + e];
    }
  }
  // the lane's row points and their inputs for the mirror block
  double xrr[RH], xri[RH], px[RH], py[RH], pz[RH];
  const double *pr = a.cs.rpts + 3 * M * (uint64_t)b.r;
  const double *pc = a.cs.cpts + 3 * M * (uint64_t)b.c;
#pragma unroll
  for (int h = 0; h < RH; h++) {
    const int i = lane + 64 * h;
    const int ii = i < R ? i : R - 1;  // rows past the rank: never stored
    xrr[h] = xri[h] = 0.0;
    if (i < R) {
      if constexpr (C) {
        xrr[h] = a.xt[2 * (b.in_r + i)];
        xri[h] = a.xt[2 * (b.in_r + i) + 1];
      } else {
        xrr[h] = a.xt[b.in_r + i];
      }
    }
    px[h] = pr[ii / (M * M)];
    py[h] = pr[M + (ii / M) % M];
    pz[h] = pr[2 * M + ii % M];
  }
  __builtin_amdgcn_wave_barrier();
  double ar[RH], ai[RH];
#pragma unroll
  for (int h = 0; h < RH; h++) ar[h] = ai[h] = 0.0;
#pragma unroll 1
  for (int p = 0; p < NP; p++) {
    double pt[W], pti[C ? W : 1];
#pragma unroll
    for (int u = 0; u < W; u++) {
      pt[u] = 0.0;
      if constexpr (C) pti[u] = 0.0;
      const int j = W * p + u;
      if (j < R) {  // uniform
        const double qx = pc[j / (M * M)], qy = pc[M + (j / M) % M],
                     qz = pc[2 * M + j % M];
#pragma unroll
        for (int h = 0; h < RH; h++) {
          const double d0 = px[h] - qx, d1 = py[h] - qy, d2 = pz[h] - qz;
          const double r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
          if constexpr (C) {
            double sr, si;
            cpl_from_r2(r2, 1, a.cs.kr, a.cs.kdamp, sr, si);
            const double xr = xs[2 * j], xi = xs[2 * j + 1];
            ar[h] = fma(sr, xr, ar[h]);
            ar[h] = fma(-si, xi, ar[h]);
            ai[h] = fma(sr, xi, fma(si, xr, ai[h]));
            pt[u] += fma(sr, xrr[h], -si * xri[h]);
            pti[u] += fma(sr, xri[h], si * xrr[h]);
          } else {
            const double sr = KCONST * cpl_rsqrt(r2);
            ar[h] = fma(sr, xs[j], ar[h]);
            pt[u] = fma(sr, xrr[h], pt[u]);
          }
        }
      }
    }
    // column W p + (lane & (W - 1)), written by the W lanes of that index
    const double cs = h2hip::wave_sum_n<W>(pt, lane);
    const int col = W * p + (lane & (W - 1));
    const bool writer = lane / W == p % (64 / W) && col < R;
    if constexpr (C) {
      const double csi = h2hip::wave_sum_n<W>(pti, lane);
      if (writer) {
        atomicAdd(a.yt + 2 * (b.out_c + col), cs);
        atomicAdd(a.yt + 2 * (b.out_c + col) + 1, csi);
      }
    } else {
      if (writer) atomicAdd(a.yt + b.out_c + col, cs);
    }
  }
#pragma unroll
  for (int h = 0; h < RH; h++) {
    const int i = lane + 64 * h;
    if (i < R) {
      if constexpr (C) {
        atomicAdd(a.yt + 2 * (b.out_r + i), ar[h]);
        atomicAdd(a.yt + 2 * (b.out_r + i) + 1, ai[h]);
      } else {
        atomicAdd(a.yt + b.out_r + i, ar[h]);
      }
    }
  }
}

// Self-adjoint product from the lower block triangle (row and column
// bases equal): one wave per block.  A far block (r, c) with r > c adds
// S xt_c to yt_r and S^* xt_r to yt_c, reading S once: lane i keeps row
// i's sum, and the column sums are reduced across the wave, one per
// column, and kept by lane j mod 64.  A near block below the diagonal adds
// alpha N x|_c to y|_r and alpha N^* x|_r to y|_c; a diagonal near block
// its lower triangle and the conjugate of the part below the diagonal.
// Blocks above the diagonal are skipped.  (addevalsymm_diag /
// addevalsymm_offdiag, Library/h2matrix.c:1366-1567.)
__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int C>
__global__ __launch_bounds__(NT) void couple_symm(CNArgs a) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  double a1r[2] = {0.0, 0.0}, a1i[2] = {0.0, 0.0};  // rows lane, lane + 64
  double a2r[2] = {0.0, 0.0}, a2i[2] = {0.0, 0.0};  // cols lane, lane + 64
  if (w < a.nfar) {
    const FarD b = a.far[w];
    if (b.r <= b.c) return;  // uniform per wave
    const NodeD r = a.rn[b.r], c = a.cn[b.c];
    double xr[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int h = 0; h < 2; h++)
      if (lane + 64 * h < r.rank)
        ld<C>(a.xt, r.coef + lane + 64 * h, xr[h][0], xr[h][1]);
    for (uint32_t j = 0; j < c.rank; j++) {
      double xcr, xci;
      ld<C>(a.xt, c.coef + j, xcr, xci);
      double pr = 0.0, pi = 0.0;
      for (int h = 0; h < 2; h++) {
        const uint32_t i = lane + 64 * h;
        if (i < r.rank) {
          double sr, si;
          ld_nt<C>(a.S, b.soff + i + (uint64_t)j * r.rank, sr, si);
          a1r[h] = fma(sr, xcr, a1r[h]);
          pr = fma(sr, xr[h][0], pr);
          if constexpr (C) {
            a1r[h] = fma(-si, xci, a1r[h]);
            a1i[h] = fma(sr, xci, fma(si, xcr, a1i[h]));
            // conj(s) x_r
            pr = fma(si, xr[h][1], pr);
            pi = fma(sr, xr[h][1], fma(-si, xr[h][0], pi));
          }
        }
      }
      pr = wave_sum64(pr);
      if constexpr (C) pi = wave_sum64(pi);
      if (lane == (j & 63)) {
        a2r[j >> 6] = pr;
        a2i[j >> 6] = pi;
      }
    }
    for (int h = 0; h < 2; h++) {
      if (lane + 64 * h < r.rank)
        atomic_add<C>(a.yt, r.coef + lane + 64 * h, a1r[h], a1i[h]);
      if (lane + 64 * h < c.rank)
        atomic_add<C>(a.yt, c.coef + lane + 64 * h, a2r[h], a2i[h]);
    }
    return;
  }
  const uint32_t nb = w - a.nfar;
  if (nb >= a.nnear) return;
  const NearD b = a.near[nb];
  if (b.roff < b.coff) return;  // above the diagonal: skipped
  const bool diag = b.roff == b.coff;
  const uint32_t *ridx = a.ridx + b.roff, *cidx = a.cidx + b.coff;
  double xr[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int h = 0; h < 2; h++)
    if (lane + 64 * h < b.rows)
      ld<C>(a.v.x[0], ridx[lane + 64 * h], xr[h][0], xr[h][1]);
  for (uint32_t j = 0; j < b.cols; j++) {
    double xcr, xci;
    ld<C>(a.v.x[0], cidx[j], xcr, xci);
    double pr = 0.0, pi = 0.0;
    for (int h = 0; h < 2; h++) {
      const uint32_t i = lane + 64 * h;
      if (i < b.rows && (!diag || i >= j)) {
        double nr, ni;
        ld<C>(a.N, b.noff + i + (uint64_t)j * b.rows, nr, ni);
        a1r[h] = fma(nr, xcr, a1r[h]);
        if constexpr (C) {
          a1r[h] = fma(-ni, xci, a1r[h]);
          a1i[h] = fma(nr, xci, fma(ni, xcr, a1i[h]));
        }
        if (!diag || i > j) {
          pr = fma(nr, xr[h][0], pr);
          if constexpr (C) {
            pr = fma(ni, xr[h][1], pr);
            pi = fma(nr, xr[h][1], fma(-ni, xr[h][0], pi));
          }
        }
      }
    }
    pr = wave_sum64(pr);
    if constexpr (C) pi = wave_sum64(pi);
    if (lane == (j & 63)) {
      a2r[j >> 6] = pr;
      a2i[j >> 6] = pi;
    }
  }
  const double ar = a.v.are[0], ai = a.v.aim[0];
  for (int h = 0; h < 2; h++) {
    if (lane + 64 * h < b.rows)
      atomic_add<C>(a.v.y[0], ridx[lane + 64 * h], ar * a1r[h] - ai * a1i[h],
                    ar * a1i[h] + ai * a1r[h]);
    if (lane + 64 * h < b.cols)
      atomic_add<C>(a.v.y[0], cidx[lane + 64 * h], ar * a2r[h] - ai * a2i[h],
                    ar * a2i[h] + ai * a2r[h]);
  }
}

struct DBArgs {
  const NodeD *nodes;
  const uint32_t *leaves;
  uint32_t nleaves;
  const uint32_t *paths;  // per leaf: MAXPATH node numbers, root first
  const uint32_t *plen;
  const uint32_t *idx;
  const double *V;  // leaf: size x rank column-major
  const double *E;  // per node s: rank_s x rank_father column-major
  uint64_t eld;
  const double *yt;
  uint64_t cstride;
  VecSet v;
};

// v = yt_root; for each node s on the path below the root v = yt_s + E_s v;
// at the leaf y|_t += alpha V_t v.  Four waves split the inputs of every
// product.
template <int C, int K, int NV>
__global__ __launch_bounds__(NT) void down_back(DBArgs a) {
  constexpr int RK = rank_cap<NV>();
  __shared__ double sv[NV * 2 * RK];
  __shared__ double part[WPB * NV * 2 * RK];
  const uint32_t w = blockIdx.x;
  const int tid = threadIdx.x;
  const uint32_t *path = a.paths + (uint64_t)w * MAXPATH;
  const uint32_t len = a.plen[w];
  {
    const NodeD r = a.nodes[path[0]];
    for (uint32_t e = tid; e < NV * r.rank; e += NT) {
      const uint32_t v = e / r.rank, i = e - v * r.rank;
      double vr = 0.0, vi = 0.0;
      if (v < a.v.nv) ld<C>(a.yt + v * a.cstride, r.coef + i, vr, vi);
      sv[v * 2 * RK + 2 * i] = vr;
      sv[v * 2 * RK + 2 * i + 1] = vi;
    }
  }
  __syncthreads();
  for (uint32_t p = 1; p < len; p++) {
    const uint32_t sn = path[p];
    const NodeD s = a.nodes[sn];
    const uint32_t kf = a.nodes[s.father].rank;
    const uint32_t rank = K ? K : s.rank;
    wg_matvec<C, NV, RK>(a.E + (uint64_t)sn * a.eld, rank, rank, kf, sv,
                         part);
    // sv is free again (all waves passed the barrier in wg_matvec)
    for (uint32_t e = tid; e < NV * rank; e += NT) {
      const uint32_t v = e / rank, i = e - v * rank;
      double vr = 0.0, vi = 0.0;
      if (v < a.v.nv) ld<C>(a.yt + v * a.cstride, s.coef + i, vr, vi);
      sv[v * 2 * RK + 2 * i] = vr + part_sum<NV, RK>(part, v, 2 * i);
      sv[v * 2 * RK + 2 * i + 1] = vi + part_sum<NV, RK>(part, v, 2 * i + 1);
    }
    __syncthreads();
  }
  const NodeD t = a.nodes[a.leaves[w]];
  const uint32_t rank = K ? K : t.rank;
  wg_matvec<C, NV, RK>(a.V + t.voff, t.size, t.size, rank, sv, part);
  for (uint32_t e = tid; e < NV * t.size; e += NT) {
    const uint32_t v = e / t.size, i = e - v * t.size;
    if (v >= a.v.nv) continue;
    const double ar = part_sum<NV, RK>(part, v, 2 * i);
    const double ai = part_sum<NV, RK>(part, v, 2 * i + 1);
    const uint32_t d = a.idx[t.off + i];
    double *y = a.v.y[v];
    if constexpr (C) {
      y[2 * d] += a.v.are[v] * ar - a.v.aim[v] * ai;
      y[2 * d + 1] += a.v.are[v] * ai + a.v.aim[v] * ar;
    } else {
      y[d] += a.v.are[v] * ar;
    }
  }
}

// ---- flat transforms ----------------------------------------------------------
//
// With the explicit cluster bases restricted to the leaves, W_a|_l = V_l
// E_l E_{father(l)} ... (the rows of leaf l of the basis of its ancestor a;
// nested bases make them equal to the bases built directly), the two
// transforms lose their dependence chains:
//   xt_a = sum over the leaves l below a of W_a|_l^T x|_l
//   y|_l += alpha sum over the ancestors a of l of W_a|_l yt_a
// One wavefront per (leaf, ancestor) pair in both launches; the upward one
// adds into the coefficients with FP64 atomics (cleared by a launch of its
// own), the downward one into y.  The blocks W_a|_l (|l| x k_a) are built
// once on the device from V and E, row-major for the upward sweep and
// column-major for the downward one, so that both read 64 consecutive
// doubles per lane step.

struct FlatItem {
  uint32_t leaf, node;  // node: the ancestor a (the leaf itself included)
  uint64_t woff;        // offset of the block W_a|_l
};

struct FlatArgs {
  const NodeD *nodes;
  const FlatItem *items;
  uint32_t nitems;
  const uint32_t *idx;
  const double *W;
  double *xt;        // upward: coefficients (atomics)
  const double *yt;  // downward: coefficients
  uint64_t cstride;  // doubles between the vectors' coefficients
  double *zero;      // cleared by this launch (the buffer the next launch
  uint64_t nzero;    // accumulates into / the next product starts from),
                     // nzero doubles per vector in use, cstride apart
  VecSet v;
  // reproducible products: upward item w writes its k coefficients to
  // P[v pstride + (w pk + j)] (pk: the largest rank) instead of adding
  double *P;
  uint64_t pstride;
  uint32_t pk;
};

// a slice of a.zero per workgroup, for every vector in use
__device__ __forceinline__ void flat_zero(const FlatArgs &a, uint32_t blk,
                                          uint32_t nblk) {
  const uint64_t chunk = (a.nzero + nblk - 1) / nblk;
  const uint64_t e0 = (uint64_t)blk * chunk,
                 e1 = min(a.nzero, e0 + chunk);
  for (uint32_t v = 0; v < a.v.nv; v++)
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += NT)
      a.zero[v * a.cstride + e] = 0.0;
}

// acc_v += sum_i M[i * stride] xs_v[i] for NV LDS vectors (XR entries
// apart, complex pairs), 16 matrix entries loaded before their products
template <int C, int NV, int XR>
__device__ __forceinline__ void dot_strided_nv(const double *__restrict__ M,
                                               uint64_t stride, uint32_t n,
                                               const double *xs, double *ar,
                                               double *ai, uint32_t i0 = 0) {
  constexpr int B = 16;
  uint32_t i = i0;
  for (; i + B <= n; i += B) {
    double m[B];
#pragma unroll
    for (int u = 0; u < B; u++) m[u] = M[(uint64_t)(i + u) * stride];
#pragma unroll
    for (int u = 0; u < B; u++)
#pragma unroll
      for (int v = 0; v < NV; v++) {
        ar[v] = fma(m[u], xs[v * 2 * XR + 2 * (i + u)], ar[v]);
        if constexpr (C) ai[v] = fma(m[u], xs[v * 2 * XR + 2 * (i + u) + 1], ai[v]);
      }
  }
  if (i < n) {
    // the last n - i < B rows, loaded together, then the products in row
    // order
    const uint32_t rem = n - i;
    double m[B];
#pragma unroll
    for (int u = 0; u < B; u++)
      m[u] = M[(uint64_t)(i + min((uint32_t)u, rem - 1)) * stride];
#pragma unroll
    for (int u = 0; u < B; u++)
      if ((uint32_t)u < rem) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          ar[v] = fma(m[u], xs[v * 2 * XR + 2 * (i + u)], ar[v]);
          if constexpr (C)
            ai[v] = fma(m[u], xs[v * 2 * XR + 2 * (i + u) + 1], ai[v]);
        }
      }
  }
}

// dot_strided_nv for one vector with rows 0 .. 15 of M loaded beforehand
// (m0; rows past n repeat the last one and are not used): the same products
// in the same order
template <int C, int XR>
__device__ __forceinline__ void dot_strided_pf(const double *__restrict__ M,
                                               uint64_t stride, uint32_t n,
                                               const double *xs,
                                               const double (&m0)[16],
                                               double &ar, double &ai) {
#pragma unroll
  for (int u = 0; u < 16; u++)
    if (u < (int)n) {
      ar = fma(m0[u], xs[2 * u], ar);
      if constexpr (C) ai = fma(m0[u], xs[2 * u + 1], ai);
    }
  if (n > 16u) dot_strided_nv<C, 1, XR>(M, stride, n, xs, &ar, &ai, 16);
}

// Upward sweep, one wave per (leaf, ancestor) block, NV vectors at once
// (leaves and ranks <= 64 when NV > 1): x of the leaf gathered into LDS,
// W read once for all vectors, atomics into the coefficients.
// block blk of nblk; xs: the wave's LDS slice of NV * 2 * XR doubles
template <int C, int K, int NV, bool ORD = false>
__device__ __forceinline__ void fwd_flat_dev(const FlatArgs &a, uint32_t blk,
                                             uint32_t nblk, double *xs) {
  constexpr int XR = NV == 1 ? MAXRANK : 64;
  if constexpr (!ORD) flat_zero(a, blk, nblk);  // the coupling accumulators
  const uint32_t w = blk * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.nitems) return;  // uniform over the wave; no block barrier
  const FlatItem it = a.items[w];
  const NodeD L = a.nodes[it.leaf];
  const NodeD A = a.nodes[it.node];
  const uint32_t sz = L.size, k = K ? K : A.rank;
  // one vector: the first rows of the lane's first column of W are loaded
  // before the input gather, so that the two round trips overlap
  double pw[16];
  if constexpr (NV == 1)
    if ((uint32_t)lane < k && sz > 0) {
#pragma unroll
      for (int u = 0; u < 16; u++)
        pw[u] = a.W[it.woff + lane + (uint64_t)min((uint32_t)u, sz - 1) * k];
    }
  for (uint32_t e = lane; e < NV * sz; e += 64) {
    const uint32_t v = e / sz, i = e - v * sz;
    double xr = 0.0, xi = 0.0;
    if (v < a.v.nv) ld<C>(a.v.x[v], a.idx[L.off + i], xr, xi);
    xs[v * 2 * XR + 2 * i] = xr;
    xs[v * 2 * XR + 2 * i + 1] = xi;
  }
  // the wave reads what its own lanes wrote: LDS keeps the order of one
  // wave's accesses, the compiler must not move them
  __builtin_amdgcn_wave_barrier();
  for (uint32_t j = lane; j < k; j += 64) {
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
    if constexpr (NV == 1) {
      if (j == (uint32_t)lane)
        dot_strided_pf<C, XR>(a.W + it.woff + j, k, sz, xs, pw, ar[0], ai[0]);
      else
        dot_strided_nv<C, NV, XR>(a.W + it.woff + j, k, sz, xs, ar, ai);
    } else {
      dot_strided_nv<C, NV, XR>(a.W + it.woff + j, k, sz, xs, ar, ai);
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (v < (int)a.v.nv) {
        if constexpr (ORD)
          st<C>(a.P + v * a.pstride, (uint64_t)w * a.pk + j, ar[v], ai[v]);
        else
          atomic_add<C>(a.xt + v * a.cstride, A.coef + j, ar[v], ai[v]);
      }
  }
}

template <int C, int K, int NV, bool ORD = false>
__global__ __launch_bounds__(NT) void fwd_flat(FlatArgs a) {
  constexpr int XS = NV * 2 * (NV == 1 ? MAXRANK : 64);
  __shared__ double xs_all[WPB * XS];
  fwd_flat_dev<C, K, NV, ORD>(a, blockIdx.x, gridDim.x,
                              xs_all + (threadIdx.x >> 6) * XS);
}

// The upward sweep in blocks [0, fb) and the near-field blocks (c.nfar = 0)
// after them, in one launch: the near blocks do not wait for the
// coefficients, so they fill the device while the short upward items run.
// TRANS: the adjoint product (near blocks N^*).
template <int C, int K, int NV, bool ORD = false, bool TRANS = false>
__global__ __launch_bounds__(NT) void fwd_flat_near(FlatArgs a, CNArgs c,
                                                    uint32_t fb) {
  constexpr int XS = NV * 2 * (NV == 1 ? MAXRANK : 64);
  __shared__ double xs_all[WPB * XS];
  double *xs = xs_all + (threadIdx.x >> 6) * XS;
  if (blockIdx.x < fb)  // uniform over the block
    fwd_flat_dev<C, K, NV, ORD>(a, blockIdx.x, fb, xs);
  else
    couple_near_dev<C, TRANS, NV, false, ORD, true>(
        c, (blockIdx.x - fb) * WPB + (threadIdx.x >> 6), xs);
}

// ---- reproducible products ----------------------------------------------------
//
// With h2hip_set_reproducible(1) no product adds with atomics: every upward
// item and every coupling or near-field block writes its result to a slot of
// a partial-sum buffer, and the launches below add the slots of each output
// in a fixed order (items of an ancestor in leaf order, blocks of an output
// cluster in block order, ancestors of a leaf root first).  Every vector's
// arithmetic is the same whatever the number of vectors of the product.

// xt_a = sum of the upward items of cluster a, one workgroup per cluster of
// the input side
template <int C, int NV>
__global__ __launch_bounds__(NT) void ord_fwd_reduce(FlatArgs a,
                                                     const uint32_t *ptr,
                                                     const uint32_t *ids,
                                                     uint32_t nnodes) {
  const uint32_t node = blockIdx.x;
  if (node >= nnodes) return;
  const NodeD A = a.nodes[node];
  const uint32_t b0 = ptr[node], b1 = ptr[node + 1];
  for (uint32_t e = threadIdx.x; e < a.v.nv * A.rank; e += NT) {
    const uint32_t v = e / A.rank, j = e - v * A.rank;
    double sr = 0.0, si = 0.0;
    for (uint32_t t = b0; t < b1; t++) {
      double pr, pi;
      ld<C>(a.P + v * a.pstride, (uint64_t)ids[t] * a.pk + j, pr, pi);
      sr += pr;
      si += pi;
    }
    st<C>(a.xt + v * a.cstride, A.coef + j, sr, si);
  }
}

struct OrdReduce {
  const NodeD *onodes;  // nodes of the output side
  const uint32_t *oidx;
  uint32_t nnodes;
  const uint32_t *fptr, *fids;  // far blocks per output cluster
  const uint32_t *nptr, *nids;  // near blocks per output cluster
  double *yt;
};

// One workgroup per cluster of the output side: yt of the cluster = sum of
// its far blocks' slots (zero without far blocks); for a leaf with near
// blocks, y|_leaf += alpha sum of their slots.
template <int C, int NV>
__global__ __launch_bounds__(NT) void ord_reduce(CNArgs a, OrdReduce r) {
  const uint32_t node = blockIdx.x;
  if (node >= r.nnodes) return;
  const NodeD O = r.onodes[node];
  const uint32_t f0 = r.fptr[node], f1 = r.fptr[node + 1];
  for (uint32_t e = threadIdx.x; e < a.v.nv * O.rank; e += NT) {
    const uint32_t v = e / O.rank, i = e - v * O.rank;
    double sr = 0.0, si = 0.0;
    for (uint32_t t = f0; t < f1; t++) {
      double pr, pi;
      ld<C>(a.P + v * a.pstride, a.pslot[r.fids[t]] + i, pr, pi);
      sr += pr;
      si += pi;
    }
    st<C>(r.yt + v * a.cstride, O.coef + i, sr, si);
  }
  const uint32_t n0 = r.nptr[node], n1 = r.nptr[node + 1];
  if (n0 == n1) return;
  for (uint32_t e = threadIdx.x; e < a.v.nv * O.size; e += NT) {
    const uint32_t v = e / O.size, i = e - v * O.size;
    double sr = 0.0, si = 0.0;
    for (uint32_t t = n0; t < n1; t++) {
      double pr, pi;
      ld<C>(a.P + v * a.pstride, a.pslot[a.pnear0 + r.nids[t]] + i, pr, pi);
      sr += pr;
      si += pi;
    }
    const uint32_t d = r.oidx[O.off + i];
    double *y = a.v.y[v];
    if constexpr (C) {
      y[2 * d] += a.v.are[v] * sr - a.v.aim[v] * si;
      y[2 * d + 1] += a.v.are[v] * si + a.v.aim[v] * sr;
    } else {
      y[d] += a.v.are[v] * sr;
    }
  }
}

// Downward sweep, one wave per leaf: y|_l += alpha sum over the ancestors a
// of l (root first) of W_a|_l yt_a, one row per lane, no atomics.
template <int C, int K, int NV>
__global__ __launch_bounds__(NT) void back_flat_ord(FlatArgs a,
                                                    const uint32_t *ibeg,
                                                    uint32_t nleaves) {
  constexpr int XR = NV == 1 ? MAXRANK : 64;
  __shared__ double ys_all[WPB][NV * 2 * XR];
  // the input coefficients xt, zero for the next product whatever its mode
  flat_zero(a, blockIdx.x, gridDim.x);
  const uint32_t l = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (l >= nleaves) return;
  double *ys = ys_all[threadIdx.x >> 6];
  const uint32_t t0 = ibeg[l], t1 = ibeg[l + 1];
  if (t0 == t1) return;  // no ancestor of the leaf in an admissible block
  const NodeD L = a.nodes[a.items[t0].leaf];
  const uint32_t sz = L.size;
  double ar[2][NV], ai[2][NV];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int v = 0; v < NV; v++) ar[h][v] = ai[h][v] = 0.0;
  for (uint32_t t = t0; t < t1; t++) {
    const FlatItem it = a.items[t];
    const NodeD A = a.nodes[it.node];
    const uint32_t k = K ? K : A.rank;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e = lane; e < NV * k; e += 64) {
      const uint32_t v = e / k, j = e - v * k;
      double yr = 0.0, yi = 0.0;
      if (v < a.v.nv) ld<C>(a.yt + v * a.cstride, A.coef + j, yr, yi);
      ys[v * 2 * XR + 2 * j] = yr;
      ys[v * 2 * XR + 2 * j + 1] = yi;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t i = lane + 64 * h;
      if (i < sz)
        dot_strided_nv<C, NV, XR>(a.W + it.woff + i, sz, k, ys, ar[h],
                                  ai[h]);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t i = lane + 64 * h;
    if (i >= sz) continue;
    const uint32_t d = a.idx[L.off + i];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (v >= (int)a.v.nv) continue;
      double *y = a.v.y[v];
      if constexpr (C) {
        y[2 * d] += a.v.are[v] * ar[h][v] - a.v.aim[v] * ai[h][v];
        y[2 * d + 1] += a.v.are[v] * ai[h][v] + a.v.aim[v] * ar[h][v];
      } else {
        y[d] += a.v.are[v] * ar[h][v];
      }
    }
  }
}

// Downward sweep, one wave per (leaf, ancestor) block, NV vectors at once:
// yt of the ancestor in LDS, one row per lane, atomics into y.
template <int C, int K, int NV>
__global__ __launch_bounds__(NT) void back_flat(FlatArgs a) {
  constexpr int XR = NV == 1 ? MAXRANK : 64;
  __shared__ double ys_all[WPB][NV * 2 * XR];
  flat_zero(a, blockIdx.x, gridDim.x);  // the input coefficients xt, for
                                        // the next product
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.nitems) return;
  double *ys = ys_all[threadIdx.x >> 6];
  const FlatItem it = a.items[w];
  const NodeD L = a.nodes[it.leaf];
  const NodeD A = a.nodes[it.node];
  const uint32_t sz = L.size, k = K ? K : A.rank;
  for (uint32_t e = lane; e < NV * k; e += 64) {
    const uint32_t v = e / k, j = e - v * k;
    double yr = 0.0, yi = 0.0;
    if (v < a.v.nv) ld<C>(a.yt + v * a.cstride, A.coef + j, yr, yi);
    ys[v * 2 * XR + 2 * j] = yr;
    ys[v * 2 * XR + 2 * j + 1] = yi;
  }
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = lane; i < sz; i += 64) {
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ar[v] = ai[v] = 0.0;
    dot_strided_nv<C, NV, XR>(a.W + it.woff + i, sz, k, ys, ar, ai);
    const uint32_t d = a.idx[L.off + i];
#pragma unroll
    for (int v = 0; v < NV; v++)
      if (v < (int)a.v.nv)
        atomic_add<C>(a.v.y[v], d, a.v.are[v] * ar[v] - a.v.aim[v] * ai[v],
                      a.v.are[v] * ai[v] + a.v.aim[v] * ar[v]);
  }
}

struct WBuildArgs {
  const NodeD *nodes;
  const FlatItem *items;
  const uint32_t *ibeg;     // items of leaf l: ibeg[l] .. ibeg[l + 1]
  const uint32_t *leaves;   // node numbers of the leaves
  const uint32_t *rowleaf;  // index position -> leaf number
  uint32_t ndof;
  const double *VT;  // leaf: rank x size column-major
  const double *ET;  // per son: rank_father x rank_son column-major
  uint64_t eld;
  double *WF, *WB;  // row-major / column-major blocks (either may be null)
};

// One wavefront per row i of a leaf l: the row of V_l, then of V_l E_l,
// V_l E_l E_f, ... up to the root, written to every block of the leaf.
__global__ __launch_bounds__(NT) void w_build(WBuildArgs a) {
  __shared__ double row_all[WPB][MAXRANK];
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.ndof) return;
  double *row = row_all[threadIdx.x >> 6];
  const uint32_t l = a.rowleaf[w];
  uint32_t cur = a.leaves[l];
  const NodeD L = a.nodes[cur];
  const uint32_t i = w - L.off, b = a.ibeg[l];
  uint32_t k = L.rank;
  for (uint32_t j = lane; j < k; j += 64)
    row[j] = a.VT[L.voff + j + (uint64_t)i * k];
  __builtin_amdgcn_wave_barrier();
  // the leaf's items (root first) name only the ancestors that take part
  // in the products: walk every ancestor, write the blocks of those
  for (uint32_t p = a.ibeg[l + 1] - b; p > 0;) {
    const FlatItem it = a.items[b + p - 1];
    if (it.node == cur) {
      for (uint32_t j = lane; j < k; j += 64) {
        if (a.WF) a.WF[it.woff + (uint64_t)i * k + j] = row[j];
        if (a.WB) a.WB[it.woff + (uint64_t)j * L.size + i] = row[j];
      }
      if (--p == 0) break;
    }
    const uint32_t f = a.nodes[cur].father;
    if (f == NONE) break;
    const uint32_t kf = a.nodes[f].rank;
    const double *ET = a.ET + (uint64_t)cur * a.eld;
    double nx[MAXRANK / 64];
#pragma unroll
    for (int h = 0; h < MAXRANK / 64; h++) {
      const uint32_t jf = lane + 64 * h;
      double acc = 0.0;
      if (jf < kf)
        for (uint32_t t = 0; t < k; t++)
          acc = fma(row[t], ET[jf + (uint64_t)t * kf], acc);
      nx[h] = acc;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < MAXRANK / 64; h++) {
      const uint32_t jf = lane + 64 * h;
      if (jf < kf) row[jf] = nx[h];
    }
    __builtin_amdgcn_wave_barrier();
    k = kf;
    cur = f;
  }
}

// Couplings of interpolation H2-matrices on the device: S_b(i, j) =
// kernel_const G(xi_i, eta_j) at the tensor Chebyshev points of the row
// and column boxes of admissible block b (i = (a m + b) m + c over the
// three coordinates), column-major at the block's coupling offset.
// G = 1/r (Laplace) or e^{i kr r} e^{-ki r}/r (Helmholtz; complex pairs),
// as assemble_bem3d_inter_uniform fills them (Library/bem3d.c:6079-6108).
struct CplArgs {
  const FarD *far;
  uint32_t nfar;
  const double *rpts, *cpts;  // 3 m per node (node records keep the order
                              // of the tree arrays)
  uint32_t m;
  int helmholtz;
  double kr, kdamp;  // phase factor kr, damping exponent -ki
  double *S;
};

__global__ __launch_bounds__(NT) void interp_couplings(CplArgs a) {
  const uint32_t b = blockIdx.x;
  if (b >= a.nfar) return;
  const FarD f = a.far[b];
  const uint32_t m = a.m, k = m * m * m, kk = k * k;
  const double *pr = a.rpts + 3 * (uint64_t)m * f.r;
  const double *pc = a.cpts + 3 * (uint64_t)m * f.c;
  for (uint32_t e = threadIdx.x; e < kk; e += NT) {
    const uint32_t i = e % k, j = e / k;
    const uint32_t ia = i / (m * m), ib = (i / m) % m, ic = i % m;
    const uint32_t ja = j / (m * m), jb = (j / m) % m, jc = j % m;
    double sr, si;
    cpl_value(pr[ia] - pc[ja], pr[m + ib] - pc[m + jb],
              pr[2 * m + ic] - pc[2 * m + jc], a.helmholtz, a.kr, a.kdamp,
              sr, si);
    const uint64_t o = f.soff + e;
    if (!a.helmholtz) {
      a.S[o] = sr;
    } else {
      a.S[2 * o] = sr;
      a.S[2 * o + 1] = si;
    }
  }
}

// ---- cluster bases by interpolation on the device ---------------------------
//
// Leaf matrices V_t[i, j] = int phi_i L_j (or int phi_i dL_j/dn with the
// triangle normal for the column basis of the double layer) with the
// single-triangle Gauss rule of order q, and transfer matrices E_s[i, j] =
// L^f_j(xi^s_i) for a son s of f, over the tensor Lagrange polynomials of
// the clusters' Chebyshev points, j = (ja m + jb) m + jc
// (Library/bem3d.c:5937-6005, 7555-7884, 8024-8336).  The arithmetic is
// that of the host construction: products of one-dimensional Lagrange
// polynomials in the same order, sums over the rule points (and over the
// support triangles of a vertex in increasing triangle order).

struct IBArgs {
  const NodeD *nodes;
  const uint32_t *idx, *leaves;
  uint32_t nnodes;
  const double *x, *n, *g;
  const uint32_t *t;
  const uint32_t *vt_ptr, *vt;  // vertex -> 4 triangle + local vertex
  const double *pts;            // 3 m per node
  uint32_t m;
  int basis, dn;
  uint32_t nq;
  double rx0[64], rx1[64], rw[64];
  double *V, *VT, *E, *ET;
  uint64_t eld;
};

__device__ __forceinline__ double lagr1d(double x, const double *xi,
                                         uint32_t mu, uint32_t m) {
  double v = 1.0;
  for (uint32_t l = 0; l < mu; l++) v *= (x - xi[l]) / (xi[mu] - xi[l]);
  for (uint32_t l = mu + 1; l < m; l++) v *= (x - xi[l]) / (xi[mu] - xi[l]);
  return v;
}

__device__ __forceinline__ double dlagr1d(double x, const double *xi,
                                          uint32_t mu, uint32_t m) {
  double sum = 0.0;
  for (uint32_t l = 0; l < m; l++) {
    if (l == mu) continue;
    double v = 1.0;
    for (uint32_t l2 = 0; l2 < m; l2++) {
      if (l2 == mu || l2 == l) continue;
      v *= (x - xi[l2]) / (xi[mu] - xi[l2]);
    }
    sum += v / (xi[mu] - xi[l]);
  }
  return sum;
}

// L_j (dn: <n, grad L_j>) at x for the points p of one cluster
__device__ __forceinline__ double lagr3d_one(const double *x, const double *p,
                                             uint32_t m, uint32_t ja,
                                             uint32_t jb, uint32_t jc,
                                             const double *nrm) {
  const double lx = lagr1d(x[0], p, ja, m), ly = lagr1d(x[1], p + m, jb, m),
               lz = lagr1d(x[2], p + 2 * m, jc, m);
  if (!nrm) return lx * ly * lz;
  const double dx = dlagr1d(x[0], p, ja, m), dy = dlagr1d(x[1], p + m, jb, m),
               dz = dlagr1d(x[2], p + 2 * m, jc, m);
  return nrm[0] * (dx * ly * lz) + nrm[1] * (lx * dy * lz) +
         nrm[2] * (lx * ly * dz);
}

__global__ __launch_bounds__(NT) void interp_leaf(IBArgs a) {
  const uint32_t node = a.leaves[blockIdx.x];
  const NodeD L = a.nodes[node];
  const uint32_t rows = L.size, k = L.rank, m = a.m;
  const double *p = a.pts + 3 * (uint64_t)m * node;
  for (uint32_t e = threadIdx.x; e < rows * k; e += NT) {
    const uint32_t i = e % rows, j = e / rows;
    const uint32_t ja = j / (m * m), jb = (j / m) % m, jc = j % m;
    double acc = 0.0, xq[3];
    if (!a.basis) {
      const uint32_t tri = a.idx[L.off + i];
      const double *A = a.x + 3 * a.t[3 * tri], *B = a.x + 3 * a.t[3 * tri + 1],
                   *C = a.x + 3 * a.t[3 * tri + 2];
      const double *nrm = a.dn ? a.n + 3 * tri : nullptr;
      for (uint32_t q = 0; q < a.nq; q++) {
        const double u = a.rx0[q], v = a.rx1[q];
        for (int d = 0; d < 3; d++)
          xq[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
        acc += a.rw[q] * lagr3d_one(xq, p, m, ja, jb, jc, nrm);
      }
      acc *= a.g[tri];
    } else {
      const uint32_t vtx = a.idx[L.off + i];
      for (uint32_t pp = a.vt_ptr[vtx]; pp < a.vt_ptr[vtx + 1]; pp++) {
        const uint32_t tri = a.vt[pp] >> 2, l = a.vt[pp] & 3;
        const double *A = a.x + 3 * a.t[3 * tri],
                     *B = a.x + 3 * a.t[3 * tri + 1],
                     *C = a.x + 3 * a.t[3 * tri + 2];
        const double *nrm = a.dn ? a.n + 3 * tri : nullptr;
        for (uint32_t q = 0; q < a.nq; q++) {
          const double u = a.rx0[q], w = a.rx1[q];
          for (int d = 0; d < 3; d++)
            xq[d] = A[d] * (1.0 - u) + B[d] * (u - w) + C[d] * w;
          const double phi = l == 0 ? 1.0 - u : (l == 1 ? u - w : w);
          const double wq = a.rw[q] * phi * a.g[tri];
          acc += wq * lagr3d_one(xq, p, m, ja, jb, jc, nrm);
        }
      }
    }
    a.V[L.voff + i + (uint64_t)j * rows] = acc;
    a.VT[L.voff + j + (uint64_t)i * k] = acc;
  }
}

__global__ __launch_bounds__(NT) void interp_transfer(IBArgs a) {
  const uint32_t s = blockIdx.x;
  const NodeD S = a.nodes[s];
  if (S.father == NONE) return;
  const uint32_t m = a.m, ks = S.rank, kf = a.nodes[S.father].rank;
  const double *ps = a.pts + 3 * (uint64_t)m * s;
  const double *pf = a.pts + 3 * (uint64_t)m * S.father;
  for (uint32_t e = threadIdx.x; e < ks * kf; e += NT) {
    const uint32_t i = e % ks, j = e / ks;
    const uint32_t ia = i / (m * m), ib = (i / m) % m, ic = i % m;
    const uint32_t ja = j / (m * m), jb = (j / m) % m, jc = j % m;
    double v = lagr1d(ps[ia], pf, ja, m);
    v *= lagr1d(ps[m + ib], pf + m, jb, m);
    v *= lagr1d(ps[2 * m + ic], pf + 2 * m, jc, m);
    a.E[(uint64_t)s * a.eld + i + (uint64_t)j * ks] = v;
    a.ET[(uint64_t)s * a.eld + j + (uint64_t)i * kf] = v;
  }
}

}  // namespace

struct h2hip_h2matrix {
  int cplx = 0;
  uint32_t urank = 0;  // common rank of all clusters, 0 if they differ
  struct Side {
    uint32_t nnodes = 0, ndof = 0, depth = 0;
    std::vector<NodeD> nodes;
    std::vector<uint32_t> leaves, sons, paths, plen, fwork, idx;
    // clusters above the fwd_up work items by height (2, 3, ...): nodes of
    // height h at hnodes[hoff[h - 2] .. hoff[h - 1])
    std::vector<uint32_t> hnodes, hoff;
    uint64_t ncoef = 0, vlen = 0, eld = 0;
    // flat transforms: one item per (leaf, ancestor), leaf by leaf
    // (ancestors root first); W blocks of wlen doubles per orientation
    std::vector<FlatItem> items;
    std::vector<uint32_t> ibeg, rowleaf;
    uint64_t wlen = 0;
    bool flat_ok = true;  // the leaves partition the index positions
    bool near_leaf = true;  // near blocks have leaves of this side as
                            // output clusters
    h2hip::DevBuf<FlatItem> d_items;
    h2hip::DevBuf<double> d_WF, d_WB;
    // reproducible products: items of each cluster (in leaf order), far and
    // near blocks of each cluster as output of the product (block order;
    // side 0 for y += H x, side 1 for the adjoint), the item ranges of the
    // leaves
    std::vector<uint32_t> it_ptr, it_ids, far_ptr, far_ids, near_ptr,
        near_ids;
    h2hip::DevBuf<uint32_t> d_it_ptr, d_it_ids, d_far_ptr, d_far_ids,
        d_near_ptr, d_near_ids, d_ibeg;
    // the flat blocks of the other orientation, for the adjoint product
    // (built on its first use; shared with the other side when both sides
    // hold the same bases)
    h2hip::DevBuf<double> d_WFt, d_WBt;
    const double *WF_adj = nullptr, *WB_adj = nullptr;
    h2hip::DevBuf<NodeD> d_nodes;
    h2hip::DevBuf<uint32_t> d_idx, d_leaves, d_sons, d_paths, d_plen, d_cnt,
        d_fwork, d_hnodes;
    h2hip::DevBuf<double> d_V, d_VT, d_E, d_ET;
  } side[2];
  uint32_t nnear = 0, nfar = 0;
  std::vector<FarD> far;
  std::vector<NearD> near;
  uint64_t nlen = 0, slen = 0;
  h2hip::DevBuf<FarD> d_far;
  h2hip::DevBuf<NearD> d_near;
  h2hip::DevBuf<double> d_S, d_N, d_xt, d_yt;
  h2hip::DevBuf<double> d_xtf, d_ytf;  // coefficients of the flat products
  // interpolation couplings (h2hip_h2matrix_interp_couplings): the points
  // and the kernel, kept for products that compute the couplings
  h2hip::DevBuf<double> d_rpts, d_cpts;
  CplSrc cpl = {};
  // symmetric block structure (same points on both sides, every far block
  // with its mirror): the far blocks r > c, whose waves compute each
  // coupling value once for both blocks (couple_mf_pairs)
  h2hip::DevBuf<PairD> d_pairs;
  uint32_t npair = 0;
  int pairs_on = 1;
  bool cpl_known = false;
  int cpl_recompute = 0;
  uint32_t maxleaf = 0, maxrank = 0;  // over both sides
  // schedule of the transforms: 0 one upward launch with counters (the
  // last son computes its father), 1 one upward launch per height above
  // the leaves and twigs, 2 flat transforms (products y += alpha H x of
  // one to eight vectors; the adjoint uses 1)
  int level_sync = 0;
  // checksums of the bases as set, per side (the self-adjoint product
  // requires equal row and column bases)
  double basis_sum[2] = {0.0, 0.0};
  int basis_set[2] = {0, 0};
  uint64_t cstride = 0;               // doubles per coefficient vector
  // coefficient vectors of the batched products (MAXNV vectors), allocated
  // on first use
  mutable std::mutex batch_lock;
  mutable h2hip::DevBuf<double> d_xtb, d_ytb;
  // reproducible products: slot of every far block, then every near block
  // (the larger of its two orientations' output sizes), partial sums of the
  // blocks and of the flat upward items for up to ord_nv vectors
  std::vector<uint64_t> pslot;
  uint64_t pslot_len = 0;
  h2hip::DevBuf<uint64_t> d_pslot;
  mutable h2hip::DevBuf<double> d_P, d_Pf;
  mutable uint32_t ord_nv = 0;
  // the adjoint product's flat blocks, built on first use
  mutable std::mutex adj_lock;
  // transposed (conjugated) copies of the couplings and near blocks for the
  // adjoint product (h2hip_h2matrix_set_adjoint_storage): made by the first
  // adjoint product after the blocks were set, on its stream
  int adj_keep = 0;
  mutable bool adjT_ok = false;
  // products of several vectors multiply the stored couplings on FP64
  // matrix cores (real matrices of rank 64)
  int cpl_mfma = 0;
  mutable h2hip::DevBuf<FarD> d_farT;
  mutable h2hip::DevBuf<NearD> d_nearT;
  mutable h2hip::DevBuf<double> d_ST, d_NT;
};

using h2hip::set_error;

namespace {

// live: the clusters whose coefficients the products use (those in an
// admissible block); the flat transforms keep items for them only (nullptr:
// all clusters)
int build_side(h2hip_h2matrix::Side &S, uint32_t nnodes, const uint32_t *tree,
               const uint32_t *rank, uint32_t ndof, const uint32_t *idx,
               const std::vector<char> *live = nullptr) {
  S.nnodes = nnodes;
  S.ndof = ndof;
  S.idx.assign(idx, idx + ndof);
  S.nodes.resize(nnodes);
  std::vector<uint32_t> nsons(nnodes, 0), level(nnodes, 0);
  uint32_t root = NONE;
  for (uint32_t i = 0; i < nnodes; i++) {
    NodeD &d = S.nodes[i];
    d.off = tree[3 * i];
    d.size = tree[3 * i + 1];
    d.father = tree[3 * i + 2];
    d.rank = rank[i];
    if ((uint64_t)d.off + d.size > ndof) {
      set_error("h2matrix: cluster exceeds the index array");
      return H2HIP_EINVAL;
    }
    if (d.rank > (uint32_t)MAXRANK) {
      set_error("h2matrix: rank above 128");
      return H2HIP_EINVAL;
    }
    if (d.father == NONE) {
      if (root != NONE) {
        set_error("h2matrix: more than one root");
        return H2HIP_EINVAL;
      }
      root = i;
    } else if (d.father >= nnodes) {
      set_error("h2matrix: father out of range");
      return H2HIP_EINVAL;
    } else {
      nsons[d.father]++;
    }
  }
  if (root == NONE) {
    set_error("h2matrix: no root");
    return H2HIP_EINVAL;
  }
  for (uint32_t i = 0; i < nnodes; i++) {
    uint32_t l = 0, f = S.nodes[i].father;
    while (f != NONE) {
      l++;
      f = S.nodes[f].father;
      if (l >= (uint32_t)MAXPATH) {
        set_error("h2matrix: tree deeper than 47 levels or cyclic");
        return H2HIP_EINVAL;
      }
    }
    level[i] = l;
    S.depth = std::max(S.depth, l);
  }
  std::vector<uint32_t> beg(nnodes + 1, 0);
  for (uint32_t i = 0; i < nnodes; i++) beg[i + 1] = beg[i] + nsons[i];
  S.sons.assign(beg[nnodes], 0);
  std::vector<uint32_t> fill(nnodes, 0);
  for (uint32_t i = 0; i < nnodes; i++) {
    const uint32_t f = S.nodes[i].father;
    if (f != NONE) S.sons[beg[f] + fill[f]++] = i;
  }
  S.leaves.clear();
  uint64_t coef = 0, v = 0, maxr = 0;
  for (uint32_t i = 0; i < nnodes; i++) {
    NodeD &d = S.nodes[i];
    d.coef = coef;
    d.sonbeg = beg[i];
    d.nsons = nsons[i];
    coef += d.rank;
    maxr = std::max<uint64_t>(maxr, d.rank);
    if (nsons[i] == 0) {
      if (d.size > (uint32_t)MAXRANK) {
        set_error("h2matrix: leaf with more than 128 DOFs");
        return H2HIP_EINVAL;
      }
      d.voff = v;
      v += (uint64_t)d.size * d.rank;
      S.leaves.push_back(i);
    } else {
      d.voff = 0;
    }
  }
  // root-to-leaf paths
  S.paths.assign((uint64_t)S.leaves.size() * MAXPATH, 0);
  S.plen.assign(S.leaves.size(), 0);
  for (size_t l = 0; l < S.leaves.size(); l++) {
    uint32_t chain[MAXPATH], n = 0;
    for (uint32_t t = S.leaves[l]; t != NONE; t = S.nodes[t].father)
      chain[n++] = t;
    for (uint32_t k = 0; k < n; k++)
      S.paths[l * MAXPATH + k] = chain[n - 1 - k];
    S.plen[l] = n;
  }
  // upward-sweep work items: clusters whose sons are all leaves, then the
  // leaves not covered by such a cluster
  S.fwork.clear();
  std::vector<char> covered(nnodes, 0);
  for (uint32_t i = 0; i < nnodes; i++) {
    if (nsons[i] == 0) continue;
    bool twig = true;
    for (uint32_t k = beg[i]; k < beg[i + 1]; k++)
      twig = twig && nsons[S.sons[k]] == 0;
    if (!twig) continue;
    S.fwork.push_back(i | 0x80000000u);
    for (uint32_t k = beg[i]; k < beg[i + 1]; k++) covered[S.sons[k]] = 1;
  }
  for (uint32_t l : S.leaves)
    if (!covered[l]) S.fwork.push_back(l);
  if (nnodes > 0x7fffffffu) {
    set_error("h2matrix: too many clusters");
    return H2HIP_EINVAL;
  }
  // heights (leaves 0); sons follow their father in the depth-first arrays
  std::vector<uint32_t> height(nnodes, 0);
  for (uint32_t i = nnodes; i-- > 0;) {
    const uint32_t f = S.nodes[i].father;
    if (f != NONE) height[f] = std::max(height[f], height[i] + 1);
  }
  const uint32_t hmax = height[root];
  S.hnodes.clear();
  S.hoff.assign(1, 0);
  for (uint32_t h = 2; h <= hmax; h++) {
    for (uint32_t i = 0; i < nnodes; i++)
      if (height[i] == h) S.hnodes.push_back(i);
    S.hoff.push_back((uint32_t)S.hnodes.size());
  }
  // flat-transform items: the ancestors of every leaf that take part in
  // admissible blocks, root first (the coefficients of the other clusters
  // are never read: their blocks W_a|_l are neither stored nor streamed)
  S.items.clear();
  S.ibeg.assign(1, 0);
  S.rowleaf.assign(ndof, NONE);
  uint64_t w = 0;
  for (size_t l = 0; l < S.leaves.size(); l++) {
    const NodeD &L = S.nodes[S.leaves[l]];
    for (uint32_t k = 0; k < S.plen[l]; k++) {
      const uint32_t a = S.paths[l * MAXPATH + k];
      if (live && !(*live)[a]) continue;
      S.items.push_back(FlatItem{S.leaves[l], a, w});
      w += (uint64_t)L.size * S.nodes[a].rank;
    }
    S.ibeg.push_back((uint32_t)S.items.size());
    for (uint32_t i = 0; i < L.size; i++) {
      if (S.rowleaf[L.off + i] != NONE) S.flat_ok = false;  // overlap
      S.rowleaf[L.off + i] = (uint32_t)l;
    }
  }
  for (uint32_t i = 0; i < ndof; i++)
    if (S.rowleaf[i] == NONE) S.flat_ok = false;  // not covered
  S.wlen = w;
  S.ncoef = coef;
  S.vlen = v;
  S.eld = maxr * maxr;
  // reproducible products: the items of every cluster, in item order
  S.it_ptr.assign(nnodes + 1, 0);
  for (const FlatItem &it : S.items) S.it_ptr[it.node + 1]++;
  for (uint32_t i = 0; i < nnodes; i++) S.it_ptr[i + 1] += S.it_ptr[i];
  S.it_ids.assign(S.items.size(), 0);
  {
    std::vector<uint32_t> pos(S.it_ptr.begin(), S.it_ptr.end() - 1);
    for (uint32_t t = 0; t < (uint32_t)S.items.size(); t++)
      S.it_ids[pos[S.items[t].node]++] = t;
  }
  std::vector<uint32_t> zero(nnodes, 0);
  if (S.d_nodes.upload(S.nodes.data(), nnodes) != hipSuccess ||
      S.d_idx.upload(idx, ndof) != hipSuccess ||
      S.d_leaves.upload(S.leaves.data(), S.leaves.size()) != hipSuccess ||
      S.d_sons.upload(S.sons.data(), S.sons.size()) != hipSuccess ||
      S.d_paths.upload(S.paths.data(), S.paths.size()) != hipSuccess ||
      S.d_plen.upload(S.plen.data(), S.plen.size()) != hipSuccess ||
      S.d_cnt.upload(zero.data(), nnodes) != hipSuccess ||
      S.d_fwork.upload(S.fwork.data(), S.fwork.size()) != hipSuccess ||
      S.d_hnodes.upload(S.hnodes.data(), S.hnodes.size()) != hipSuccess ||
      S.d_items.upload(S.items.data(), S.items.size()) != hipSuccess ||
      S.d_it_ptr.upload(S.it_ptr.data(), S.it_ptr.size()) != hipSuccess ||
      S.d_it_ids.upload(S.it_ids.data(), S.it_ids.size()) != hipSuccess ||
      S.d_ibeg.upload(S.ibeg.data(), S.ibeg.size()) != hipSuccess) {
    set_error("h2matrix: device allocation failed");
    return H2HIP_ENOMEM;
  }
  return H2HIP_OK;
}

template <int C, int K, int NV>
void launch(const VecSet &vs, const h2hip_h2matrix *H, double *xt, double *yt,
            hipStream_t st, bool trans, bool symm = false) {
  const auto &in = H->side[trans ? 0 : 1];
  const auto &out = H->side[trans ? 1 : 0];
  // flat blocks: row-major of the input side, column-major of the output
  // side (the adjoint's are built on its first use)
  const double *WFin = trans ? in.WF_adj : in.d_WF.p;
  const double *WBout = trans ? out.WB_adj : out.d_WB.p;
  const bool flat = H->level_sync == 2 && WFin && WBout && H->d_xtf.p;
  // fixed-order sums (not for the self-adjoint product)
  const bool ord = h2hip::reproducible() && !symm && H->d_P.p &&
                   H->ord_nv >= vs.nv;
  // the adjoint from the transposed copies of the blocks when they are
  // kept: the forward kernels on S^* and N^* (the couplings computed in the
  // product still take the adjoint path, from the interpolation points)
  const bool tv = trans && !symm && H->adjT_ok;
  // other adjoint blocks of at most 64 x 64 go through LDS
  // (couple_trans_wg)
  const bool twg = trans && !symm && !tv && H->maxrank <= 64 &&
                   H->maxleaf <= 64;
  // flat, not symmetric: the near blocks run in the upward launch (not for
  // the workgroup adjoint blocks)
  const bool fuse = flat && !symm && !twg && H->nnear > 0 &&
                    !in.items.empty();
  const int cm = C ? 2 : 1;
  const bool mf = H->cpl_recompute && !symm;
  // view: the transposed blocks, whose rows are the column side's
  auto coupling_args = [&](bool view = false) {
    CNArgs c;
    const int a = view ? 1 : 0;
    c.far = view ? H->d_farT.p : H->d_far.p;
    c.nfar = H->nfar;
    c.near = view ? H->d_nearT.p : H->d_near.p;
    c.nnear = H->nnear;
    c.rn = H->side[a].d_nodes.p;
    c.cn = H->side[1 - a].d_nodes.p;
    c.ridx = H->side[a].d_idx.p;
    c.cidx = H->side[1 - a].d_idx.p;
    c.S = view ? H->d_ST.p : H->d_S.p;
    c.N = view ? H->d_NT.p : H->d_N.p;
    c.xt = xt;
    c.yt = yt;
    c.cstride = H->cstride;
    c.v = vs;
    c.cs = H->cpl;
    c.P = H->d_P.p;
    c.pslot = H->d_pslot.p;
    c.pstride = (uint64_t)cm * H->pslot_len;
    c.pnear0 = H->nfar;
    return c;
  };
  FlatArgs fa;
  if (flat) {
    // coefficient buffers of the flat products only: xt is zero on entry
    // (cleared by the previous flat product's downward launch, or at
    // allocation); the upward launch clears yt before the couplings add
    // into it
    xt = H->d_xtf.p;
    yt = H->d_ytf.p;
    fa.cstride = H->cstride;
    fa.nodes = in.d_nodes.p;
    fa.items = in.d_items.p;
    fa.nitems = (uint32_t)in.items.size();
    fa.idx = in.d_idx.p;
    fa.W = WFin;
    fa.xt = xt;
    fa.yt = nullptr;
    fa.zero = yt;
    fa.nzero = (uint64_t)cm * out.ncoef;
    fa.v = vs;
    fa.P = H->d_Pf.p;
    fa.pk = H->maxrank;
    fa.pstride = (uint64_t)cm * H->maxrank *
                 std::max(H->side[0].items.size(), H->side[1].items.size());
    const uint32_t fb = (fa.nitems + WPB - 1) / WPB;
    const dim3 gf(fb + (fuse ? (H->nnear + WPB - 1) / WPB : 0));
    CNArgs cn = coupling_args(tv);
    cn.xt = xt;
    cn.yt = yt;
    cn.nfar = 0;
    // the near blocks of the adjoint: transposed copies (forward kernel) or
    // the adjoint kernel
    const bool ntrans = trans && !tv;
    // no items: no cluster of the input side is in an admissible block,
    // and no coefficient is needed
    if (ord) {
      if (fuse && ntrans)
        hipLaunchKernelGGL((fwd_flat_near<C, K, NV, true, true>), gf,
                           dim3(NT), 0, st, fa, cn, fb);
      else if (fuse)
        hipLaunchKernelGGL((fwd_flat_near<C, K, NV, true, false>), gf,
                           dim3(NT), 0, st, fa, cn, fb);
      else if (fb)
        hipLaunchKernelGGL((fwd_flat<C, K, NV, true>), gf, dim3(NT), 0, st,
                           fa);
      hipLaunchKernelGGL((ord_fwd_reduce<C, NV>), dim3(in.nnodes), dim3(NT),
                         0, st, fa, (const uint32_t *)in.d_it_ptr.p,
                         (const uint32_t *)in.d_it_ids.p, in.nnodes);
    } else if (fuse && ntrans) {
      hipLaunchKernelGGL((fwd_flat_near<C, K, NV, false, true>), gf,
                         dim3(NT), 0, st, fa, cn, fb);
    } else if (fuse) {
      hipLaunchKernelGGL((fwd_flat_near<C, K, NV, false, false>), gf,
                         dim3(NT), 0, st, fa, cn, fb);
    } else if (fb) {
      hipLaunchKernelGGL((fwd_flat<C, K, NV>), gf, dim3(NT), 0, st, fa);
    }
  }
  FwdArgs f;
  f.nodes = in.d_nodes.p;
  f.work = in.d_fwork.p;
  f.nwork = (uint32_t)in.fwork.size();
  f.idx = in.d_idx.p;
  f.sons = in.d_sons.p;
  f.VT = in.d_VT.p;
  f.ET = in.d_ET.p;
  f.eld = in.eld;
  f.xt = xt;
  f.cstride = H->cstride;
  f.cnt = in.d_cnt.p;
  f.yt = yt;
  f.nyt = (uint64_t)cm * out.ncoef;
  f.v = vs;
  f.climb = H->level_sync ? 0 : 1;
  if (!flat)
    hipLaunchKernelGGL((fwd_up<C, K, NV>), dim3(f.nwork), dim3(NT), 0, st, f);
  if (H->level_sync && !flat) {
    for (size_t h = 0; h + 1 < in.hoff.size(); h++) {
      const uint32_t cnt = in.hoff[h + 1] - in.hoff[h];
      if (cnt)
        hipLaunchKernelGGL((fwd_level<C, K, NV>), dim3(cnt), dim3(NT), 0, st,
                           f, (const uint32_t *)(in.d_hnodes.p + in.hoff[h]));
    }
  }

  // computed couplings keep the adjoint kernel for the far blocks; their
  // near blocks then come from the transposed copies in a launch of their
  // own (unless the upward launch did them)
  CNArgs c = coupling_args(tv && !mf);
  const bool ktrans = trans && !(tv && !mf);
  if (fuse) {
    c.nnear = 0;  // done in the upward launch
  } else if (tv && mf && H->nnear) {
    CNArgs cnr = coupling_args(true);
    cnr.nfar = 0;
    const dim3 gn((cnr.nnear + WPB - 1) / WPB);
    if (ord)
      hipLaunchKernelGGL((couple_near<C, false, NV, false, true>), gn,
                         dim3(NT), 0, st, cnr);
    else
      hipLaunchKernelGGL((couple_near<C, false, NV>), gn, dim3(NT), 0, st,
                         cnr);
    c.nnear = 0;
  }
  if (twg) {
    // stored couplings and near blocks: one workgroup each; computed
    // couplings: the wave path for the far blocks, workgroups for the near
    CNArgs cw = c;
    if (mf) cw.nfar = 0;
    const uint32_t nwg = cw.nfar + cw.nnear;
    if (nwg) {
      if (ord)
        hipLaunchKernelGGL((couple_trans_wg<C, NV, true>), dim3(nwg),
                           dim3(NT), 0, st, cw);
      else
        hipLaunchKernelGGL((couple_trans_wg<C, NV, false>), dim3(nwg),
                           dim3(NT), 0, st, cw);
    }
    c.nnear = 0;
    if (!mf) c.nfar = 0;
  }
  // several vectors, stored real couplings of rank 64: the far blocks on
  // matrix cores (h2hip_h2matrix_set_mfma)
  if constexpr (!C && NV > 1) {
    if (H->cpl_mfma && H->urank == 64 && !mf && !ord && !symm && !ktrans &&
        c.nfar) {
      hipLaunchKernelGGL((couple_mfma<NV>), dim3((c.nfar + WPB - 1) / WPB),
                         dim3(NT), 0, st, c);
      c.nfar = 0;  // the near blocks (if any are left) as below
    }
  }
  const uint32_t nw = c.nfar + c.nnear;
  const dim3 gc((nw + WPB - 1) / WPB);
  if (nw && symm) {
    if constexpr (NV == 1)
      hipLaunchKernelGGL((couple_symm<C>), gc, dim3(NT), 0, st, c);
  } else if (nw) {
    if (ktrans && mf) {
      if (ord)
        hipLaunchKernelGGL((couple_near<C, true, NV, true, true>), gc,
                           dim3(NT), 0, st, c);
      else
        hipLaunchKernelGGL((couple_near<C, true, NV, true>), gc, dim3(NT), 0,
                           st, c);
    } else if (ktrans) {
      if (ord)
        hipLaunchKernelGGL((couple_near<C, true, NV, false, true>), gc,
                           dim3(NT), 0, st, c);
      else
        hipLaunchKernelGGL((couple_near<C, true, NV>), gc, dim3(NT), 0, st,
                           c);
    } else if (mf) {
      // symmetric structure, order-4 or -5 couplings, one vector, far
      // blocks only: each value once for a block and its mirror
      bool paired = false;
      if constexpr (NV == 1)
        paired = H->pairs_on && H->npair && c.nnear == 0 && !ord;
      const dim3 gp((H->npair + WPB - 1) / WPB);
      if (paired && H->cpl.m == 4)
        hipLaunchKernelGGL((couple_mf_pairs<C, 4>), gp, dim3(NT), 0, st, c,
                           (const PairD *)H->d_pairs.p, H->npair);
      else if (paired)
        hipLaunchKernelGGL((couple_mf_pairs<C, 5>), gp, dim3(NT), 0, st, c,
                           (const PairD *)H->d_pairs.p, H->npair);
      else if (ord)
        hipLaunchKernelGGL((couple_near<C, false, NV, true, true>), gc,
                           dim3(NT), 0, st, c);
      else
        hipLaunchKernelGGL((couple_near<C, false, NV, true>), gc, dim3(NT),
                           0, st, c);
    } else {
      if (ord)
        hipLaunchKernelGGL((couple_near<C, false, NV, false, true>), gc,
                           dim3(NT), 0, st, c);
      else
        hipLaunchKernelGGL((couple_near<C, false, NV>), gc, dim3(NT), 0, st,
                           c);
    }
  }
  if (ord) {
    // coupling slots into yt, near-field slots into y, in block order
    OrdReduce r;
    r.onodes = out.d_nodes.p;
    r.oidx = out.d_idx.p;
    r.nnodes = out.nnodes;
    r.fptr = out.d_far_ptr.p;
    r.fids = out.d_far_ids.p;
    r.nptr = out.d_near_ptr.p;
    r.nids = out.d_near_ids.p;
    r.yt = yt;
    hipLaunchKernelGGL((ord_reduce<C, NV>), dim3(out.nnodes), dim3(NT), 0,
                       st, coupling_args(), r);
  }

  if (flat) {
    fa.nodes = out.d_nodes.p;
    fa.items = out.d_items.p;
    fa.nitems = (uint32_t)out.items.size();
    fa.idx = out.d_idx.p;
    fa.W = WBout;
    fa.xt = nullptr;
    fa.yt = yt;
    fa.zero = xt;
    fa.nzero = (uint64_t)cm * in.ncoef;
    if (ord) {
      const uint32_t nl = (uint32_t)out.leaves.size();
      hipLaunchKernelGGL((back_flat_ord<C, K, NV>),
                         dim3((nl + WPB - 1) / WPB), dim3(NT), 0, st, fa,
                         (const uint32_t *)out.d_ibeg.p, nl);
    } else if (fa.nitems) {
      hipLaunchKernelGGL((back_flat<C, K, NV>),
                         dim3((fa.nitems + WPB - 1) / WPB), dim3(NT), 0, st,
                         fa);
    }
    return;
  }
  DBArgs d;
  d.nodes = out.d_nodes.p;
  d.leaves = out.d_leaves.p;
  d.nleaves = (uint32_t)out.leaves.size();
  d.paths = out.d_paths.p;
  d.plen = out.d_plen.p;
  d.idx = out.d_idx.p;
  d.V = out.d_V.p;
  d.E = out.d_E.p;
  d.eld = out.eld;
  d.yt = yt;
  d.cstride = H->cstride;
  d.v = vs;
  hipLaunchKernelGGL((down_back<C, K, NV>), dim3(d.nleaves), dim3(NT), 0, st,
                     d);
}

template <int C>
int apply_symm(const double *alpha, const h2hip_h2matrix *H, const double *x,
               double *y, hipStream_t st) {
  VecSet vs = {};
  vs.x[0] = x;
  vs.y[0] = y;
  vs.are[0] = alpha[0];
  vs.aim[0] = alpha[1];
  vs.nv = 1;
  double *xt = H->d_xt.p, *yt = H->d_yt.p;
  switch (H->urank) {
    case 8: launch<C, 8, 1>(vs, H, xt, yt, st, false, true); break;
    case 27: launch<C, 27, 1>(vs, H, xt, yt, st, false, true); break;
    case 64: launch<C, 64, 1>(vs, H, xt, yt, st, false, true); break;
    case 125: launch<C, 125, 1>(vs, H, xt, yt, st, false, true); break;
    default: launch<C, 0, 1>(vs, H, xt, yt, st, false, true); break;
  }
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// One block of a transpose: rows x cols column-major at element offset off
// of the source, cols x rows column-major at the same offset of the
// destination.
struct TJob {
  uint64_t off;
  uint32_t rows, cols;
};

// dst = src^* block by block, one workgroup per block, through 32 x 32 LDS
// tiles (reads and writes along columns)
template <int C>
__global__ __launch_bounds__(NT) void block_transpose(const TJob *jobs,
                                                      const double *src,
                                                      double *dst) {
  __shared__ double tr[32][33];
  __shared__ double ti[C ? 32 : 1][33];
  const TJob J = jobs[blockIdx.x];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (uint32_t r0 = 0; r0 < J.rows; r0 += 32)
    for (uint32_t c0 = 0; c0 < J.cols; c0 += 32) {
      for (uint32_t yy = ty; yy < 32; yy += NT / 32) {
        const uint32_t r = r0 + tx, c = c0 + yy;
        if (r < J.rows && c < J.cols) {
          double re, im;
          ld<C>(src, J.off + r + (uint64_t)c * J.rows, re, im);
          tr[yy][tx] = re;
          if constexpr (C) ti[yy][tx] = im;
        }
      }
      __syncthreads();
      for (uint32_t yy = ty; yy < 32; yy += NT / 32) {
        const uint32_t c = c0 + tx, r = r0 + yy;
        if (r < J.rows && c < J.cols) {
          if constexpr (C)
            st<C>(dst, J.off + c + (uint64_t)r * J.cols, tr[tx][yy],
                  -ti[tx][yy]);
          else
            st<C>(dst, J.off + c + (uint64_t)r * J.cols, tr[tx][yy], 0.0);
        }
      }
      __syncthreads();
    }
}

// Buffers a product needs beyond the matrix: the adjoint's flat blocks and
// transposed blocks, the partial sums of reproducible products (allocated
// on first use; the transposes ordered on st).
int ensure_product(const h2hip_h2matrix *H, uint32_t nv, bool trans,
                   hipStream_t st);

template <int C>
int apply(const double *alpha, const h2hip_h2matrix *H, const double *x,
          double *y, hipStream_t st, bool trans) {
  {
    const int rc = ensure_product(H, 1, trans, st);
    if (rc != H2HIP_OK) return rc;
  }
  VecSet vs = {};
  vs.x[0] = x;
  vs.y[0] = y;
  vs.are[0] = alpha[0];
  vs.aim[0] = alpha[1];
  vs.nv = 1;
  double *xt = H->d_xt.p, *yt = H->d_yt.p;
  switch (H->urank) {
    case 8: launch<C, 8, 1>(vs, H, xt, yt, st, trans); break;
    case 27: launch<C, 27, 1>(vs, H, xt, yt, st, trans); break;
    case 64: launch<C, 64, 1>(vs, H, xt, yt, st, trans); break;
    case 125: launch<C, 125, 1>(vs, H, xt, yt, st, trans); break;
    default: launch<C, 0, 1>(vs, H, xt, yt, st, trans); break;
  }
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

template <int C, int NV>
void launch_batch(const VecSet &vs, const h2hip_h2matrix *H, double *xt,
                  double *yt, hipStream_t st, bool trans) {
  if (H->urank == 64)
    launch<C, 64, NV>(vs, H, xt, yt, st, trans);
  else
    launch<C, 0, NV>(vs, H, xt, yt, st, trans);
}

template <int C>
int apply_batch(uint32_t nvec, const double *alpha, const h2hip_h2matrix *H,
                const void *const *x, void *const *y, hipStream_t st,
                bool trans) {
  if (nvec == 1) return apply<C>(alpha, H, (const double *)x[0],
                                 (double *)y[0], st, trans);
  if (H->maxleaf > 64 || H->maxrank > 64) {
    set_error("batched H2 product: leaves and ranks of at most 64 needed");
    return H2HIP_EINVAL;
  }
  {
    std::lock_guard<std::mutex> guard(H->batch_lock);
    if (!H->d_xtb.p) {
      if (H->d_xtb.alloc(MAXNV * H->cstride) != hipSuccess ||
          H->d_ytb.alloc(MAXNV * H->cstride) != hipSuccess) {
        set_error("batched H2 product: device allocation failed");
        return H2HIP_ENOMEM;
      }
    }
  }
  {
    const int rc = ensure_product(H, nvec, trans, st);
    if (rc != H2HIP_OK) return rc;
  }
  VecSet vs = {};
  for (uint32_t v = 0; v < nvec; v++) {
    vs.x[v] = (const double *)x[v];
    vs.y[v] = (double *)y[v];
    vs.are[v] = alpha[2 * v];
    vs.aim[v] = alpha[2 * v + 1];
  }
  vs.nv = nvec;
  double *xt = H->d_xtb.p, *yt = H->d_ytb.p;
  if (nvec <= 2)
    launch_batch<C, 2>(vs, H, xt, yt, st, trans);
  else if (nvec <= 4)
    launch_batch<C, 4>(vs, H, xt, yt, st, trans);
  else
    launch_batch<C, 8>(vs, H, xt, yt, st, trans);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// Builds the blocks W_a|_l of one side from its bases on the device into WF
// (row-major, upward sweep) and / or WB (column-major, downward sweep);
// either may be null.  Complete on return.
int build_flat_into(const h2hip_h2matrix::Side &S, double *WF, double *WB) {
  if (!S.d_VT.p || !S.d_ET.p) {
    set_error("h2matrix: flat transforms need the bases");
    return H2HIP_EINVAL;
  }
  if (!S.flat_ok) {
    set_error("h2matrix: flat transforms need leaves that partition the "
              "index positions");
    return H2HIP_EINVAL;
  }
  h2hip::DevBuf<uint32_t> d_rowleaf;
  if (d_rowleaf.upload(S.rowleaf.data(), S.rowleaf.size()) != hipSuccess)
    return H2HIP_ENOMEM;
  WBuildArgs a;
  a.nodes = S.d_nodes.p;
  a.items = S.d_items.p;
  a.ibeg = S.d_ibeg.p;
  a.leaves = S.d_leaves.p;
  a.rowleaf = d_rowleaf.p;
  a.ndof = S.ndof;
  a.VT = S.d_VT.p;
  a.ET = S.d_ET.p;
  a.eld = S.eld;
  a.WF = WF;
  a.WB = WB;
  hipLaunchKernelGGL(w_build, dim3((S.ndof + WPB - 1) / WPB), dim3(NT), 0,
                     nullptr, a);
  H2HIP_CHECK(hipGetLastError());
  // the index buffer is freed on return
  H2HIP_CHECK(hipStreamSynchronize(nullptr));
  return H2HIP_OK;
}

int build_flat(h2hip_h2matrix::Side &S, bool rowmajor, bool colmajor) {
  if (rowmajor && S.d_WF.alloc(S.wlen) != hipSuccess) return H2HIP_ENOMEM;
  if (colmajor && S.d_WB.alloc(S.wlen) != hipSuccess) return H2HIP_ENOMEM;
  return build_flat_into(S, rowmajor ? S.d_WF.p : nullptr,
                         colmajor ? S.d_WB.p : nullptr);
}

// the adjoint's blocks are rebuilt on their next use
void drop_adjoint_flat(h2hip_h2matrix *H) {
  std::lock_guard<std::mutex> guard(H->adj_lock);
  for (auto &S : H->side) {
    S.d_WFt.release();
    S.d_WBt.release();
    S.WF_adj = S.WB_adj = nullptr;
  }
}

// flat transforms of the products y += alpha H x: row-major blocks of the
// column side, column-major blocks of the row side
int build_flat_pair(h2hip_h2matrix *H) {
  int rc = build_flat(H->side[1], true, false);
  if (rc == H2HIP_OK) rc = build_flat(H->side[0], false, true);
  if (rc == H2HIP_OK && !H->d_xtf.p) {
    // coefficients of up to MAXNV vectors (batched products)
    const uint64_t len = (uint64_t)MAXNV * H->cstride;
    if (H->d_xtf.alloc(len) != hipSuccess ||
        H->d_ytf.alloc(len) != hipSuccess ||
        hipMemsetAsync(H->d_xtf.p, 0, sizeof(double) * len, nullptr) !=
            hipSuccess ||
        hipMemsetAsync(H->d_ytf.p, 0, sizeof(double) * len, nullptr) !=
            hipSuccess ||
        hipStreamSynchronize(nullptr) != hipSuccess)
      rc = H2HIP_ENOMEM;
  }
  if (rc != H2HIP_OK) {
    H->side[1].d_WF.release();
    H->side[0].d_WB.release();
    H->d_xtf.release();
    H->d_ytf.release();
    if (rc == H2HIP_ENOMEM) set_error("h2matrix: flat transforms: device "
                                      "allocation failed");
  }
  drop_adjoint_flat(H);
  return rc;
}

// Do both sides hold the same bases (same tree layout, same interpolation
// data)?  Then the blocks of one side serve the other.
bool same_bases(const h2hip_h2matrix *H) {
  const auto &r = H->side[0], &c = H->side[1];
  bool same = r.nnodes == c.nnodes && r.ndof == c.ndof &&
              H->basis_set[0] && H->basis_set[1] &&
              H->basis_sum[0] == H->basis_sum[1] && r.idx == c.idx;
  for (uint32_t i = 0; same && i < r.nnodes; i++)
    same = r.nodes[i].off == c.nodes[i].off &&
           r.nodes[i].size == c.nodes[i].size &&
           r.nodes[i].rank == c.nodes[i].rank &&
           r.nodes[i].father == c.nodes[i].father;
  return same;
}

int ensure_product(const h2hip_h2matrix *Hc, uint32_t nv, bool trans,
                   hipStream_t st) {
  h2hip_h2matrix *H = const_cast<h2hip_h2matrix *>(Hc);
  if (trans && H->adj_keep) {
    std::lock_guard<std::mutex> guard(H->adj_lock);
    if (!H->adjT_ok) {
      // far block (r, c) -> (c, r) and near block (roff, rows, coff, cols)
      // -> (coff, cols, roff, rows), each at its own offset
      std::vector<FarD> farT(H->far.size());
      std::vector<TJob> js(H->far.size()), jn(H->near.size());
      for (size_t w = 0; w < H->far.size(); w++) {
        const FarD &b = H->far[w];
        farT[w] = FarD{b.c, b.r, b.soff};
        js[w] = TJob{b.soff, H->side[0].nodes[b.r].rank,
                     H->side[1].nodes[b.c].rank};
      }
      std::vector<NearD> nearT(H->near.size());
      for (size_t w = 0; w < H->near.size(); w++) {
        const NearD &b = H->near[w];
        nearT[w] = NearD{b.coff, b.cols, b.roff, b.rows, b.noff};
        jn[w] = TJob{b.noff, b.rows, b.cols};
      }
      h2hip::DevBuf<TJob> d_js, d_jn;
      const bool ok =
          H->d_farT.upload(farT.data(), farT.size()) == hipSuccess &&
          H->d_nearT.upload(nearT.data(), nearT.size()) == hipSuccess &&
          d_js.upload(js.data(), js.size()) == hipSuccess &&
          d_jn.upload(jn.data(), jn.size()) == hipSuccess &&
          (H->d_ST.n == H->d_S.n || H->d_ST.alloc(H->d_S.n) == hipSuccess) &&
          (H->d_NT.n == H->d_N.n || H->d_NT.alloc(H->d_N.n) == hipSuccess);
      if (!ok) {
        H->d_ST.release();
        H->d_NT.release();
        set_error("h2matrix: transposed blocks of the adjoint: device "
                  "allocation failed");
        return H2HIP_ENOMEM;
      }
      if (!js.empty())
        hipLaunchKernelGGL(H->cplx ? block_transpose<1> : block_transpose<0>,
                           dim3((uint32_t)js.size()), dim3(NT), 0, st,
                           (const TJob *)d_js.p, (const double *)H->d_S.p,
                           H->d_ST.p);
      if (!jn.empty())
        hipLaunchKernelGGL(H->cplx ? block_transpose<1> : block_transpose<0>,
                           dim3((uint32_t)jn.size()), dim3(NT), 0, st,
                           (const TJob *)d_jn.p, (const double *)H->d_N.p,
                           H->d_NT.p);
      // the job lists are freed on return: wait for the transposes
      H2HIP_CHECK(hipStreamSynchronize(st));
      H2HIP_CHECK(hipGetLastError());
      H->adjT_ok = true;
    }
  }
  if (trans && H->level_sync == 2 && H->d_xtf.p) {
    // the adjoint's flat blocks: row-major of the row side, column-major
    // of the column side
    std::lock_guard<std::mutex> guard(H->adj_lock);
    auto &R = H->side[0], &Cs = H->side[1];
    if (!R.WF_adj || !Cs.WB_adj) {
      if (same_bases(H)) {
        R.WF_adj = Cs.d_WF.p;
        Cs.WB_adj = R.d_WB.p;
      } else {
        if (R.d_WFt.alloc(R.wlen) != hipSuccess ||
            Cs.d_WBt.alloc(Cs.wlen) != hipSuccess) {
          R.d_WFt.release();
          Cs.d_WBt.release();
          set_error("h2matrix: adjoint flat blocks: device allocation "
                    "failed");
          return H2HIP_ENOMEM;
        }
        int rc = build_flat_into(R, R.d_WFt.p, nullptr);
        if (rc == H2HIP_OK) rc = build_flat_into(Cs, nullptr, Cs.d_WBt.p);
        if (rc != H2HIP_OK) {
          R.d_WFt.release();
          Cs.d_WBt.release();
          return rc;
        }
        R.WF_adj = R.d_WFt.p;
        Cs.WB_adj = Cs.d_WBt.p;
      }
    }
  }
  if (h2hip::reproducible()) {
    // every output row is written by one workgroup: the leaves partition
    // the index positions and the near blocks have leaves as output
    // clusters (as in strict block trees)
    const auto &O = H->side[trans ? 1 : 0];
    if (!O.flat_ok || !O.near_leaf) {
      set_error("h2matrix: reproducible products need leaves that partition "
                "the index positions and near blocks between leaves");
      return H2HIP_EINVAL;
    }
    // partial sums for the vector count class of the launch (1, 2, 4, 8)
    const uint32_t cap = nv <= 1 ? 1 : nv <= 2 ? 2 : nv <= 4 ? 4 : 8;
    std::lock_guard<std::mutex> guard(H->batch_lock);
    if (H->ord_nv < cap) {
      const uint64_t cm = H->cplx ? 2 : 1;
      const uint64_t items = std::max(H->side[0].items.size(),
                                      H->side[1].items.size());
      if (H->d_P.alloc(cm * H->pslot_len * cap) != hipSuccess ||
          H->d_Pf.alloc(cm * items * H->maxrank * cap) != hipSuccess) {
        H->d_P.release();
        H->d_Pf.release();
        H->ord_nv = 0;
        set_error("h2matrix: reproducible products: device allocation "
                  "failed");
        return H2HIP_ENOMEM;
      }
      H->ord_nv = cap;
    }
  }
  return H2HIP_OK;
}

}  // namespace

extern "C" {

int h2hip_h2matrix_create(h2hip_h2matrix **out, int is_complex,
                          uint32_t rnodes, const uint32_t *rtree,
                          const uint32_t *rrank, uint32_t nrdof,
                          const uint32_t *ridx, uint32_t cnodes,
                          const uint32_t *ctree, const uint32_t *crank,
                          uint32_t ncdof, const uint32_t *cidx,
                          uint32_t nnear, const uint32_t *near, uint32_t nfar,
                          const uint32_t *far) {
  if (!out || !rtree || !rrank || !ridx || !ctree || !crank || !cidx ||
      (nnear && !near) || (nfar && !far) || rnodes == 0 || cnodes == 0) {
    set_error("h2hip_h2matrix_create: invalid arguments");
    return H2HIP_EINVAL;
  }
  h2hip_h2matrix *H = new h2hip_h2matrix;
  H->cplx = is_complex ? 1 : 0;
  // clusters in admissible blocks, per side; with equal trees the union of
  // both sides (the self-adjoint product reads both coefficients of a
  // block from one side)
  std::vector<char> rlive(rnodes, 0), clive(cnodes, 0);
  for (uint32_t b = 0; b < nfar; b++) {
    if (far[2 * b] < rnodes) rlive[far[2 * b]] = 1;
    if (far[2 * b + 1] < cnodes) clive[far[2 * b + 1]] = 1;
  }
  if (rnodes == cnodes && nrdof == ncdof &&
      std::equal(rtree, rtree + 3 * (size_t)rnodes, ctree) &&
      std::equal(ridx, ridx + nrdof, cidx))
    for (uint32_t i = 0; i < rnodes; i++)
      rlive[i] = clive[i] = (char)(rlive[i] | clive[i]);
  int rc = build_side(H->side[0], rnodes, rtree, rrank, nrdof, ridx, &rlive);
  if (rc == H2HIP_OK)
    rc = build_side(H->side[1], cnodes, ctree, crank, ncdof, cidx, &clive);
  if (rc != H2HIP_OK) {
    delete H;
    return rc;
  }
  H->urank = rrank[0];
  for (uint32_t i = 0; i < rnodes; i++)
    if (rrank[i] != H->urank) H->urank = 0;
  for (uint32_t i = 0; i < cnodes; i++)
    if (crank[i] != H->urank) H->urank = 0;
  H->nnear = nnear;
  H->nfar = nfar;
  uint64_t no = 0, so = 0;
  for (uint32_t b = 0; b < nnear; b++) {
    if (near[2 * b] >= rnodes || near[2 * b + 1] >= cnodes) {
      set_error("h2hip_h2matrix_create: near block node out of range");
      delete H;
      return H2HIP_EINVAL;
    }
    const NodeD &r = H->side[0].nodes[near[2 * b]];
    const NodeD &c = H->side[1].nodes[near[2 * b + 1]];
    H->near.push_back(NearD{r.off, r.size, c.off, c.size, no});
    no += (uint64_t)r.size * c.size;
  }
  for (uint32_t b = 0; b < nfar; b++) {
    if (far[2 * b] >= rnodes || far[2 * b + 1] >= cnodes) {
      set_error("h2hip_h2matrix_create: far block node out of range");
      delete H;
      return H2HIP_EINVAL;
    }
    const NodeD &r = H->side[0].nodes[far[2 * b]];
    const NodeD &c = H->side[1].nodes[far[2 * b + 1]];
    H->far.push_back(FarD{far[2 * b], far[2 * b + 1], so});
    so += (uint64_t)r.rank * c.rank;
  }
  H->nlen = no;
  H->slen = so;
  // reproducible products: a slot per block, sized for either orientation,
  // and the blocks of every cluster as output (side 0: y += H x, side 1:
  // the adjoint), in block order
  H->pslot.resize((size_t)nfar + nnear);
  uint64_t ps = 0;
  for (uint32_t b = 0; b < nfar; b++) {
    H->pslot[b] = ps;
    ps += std::max(H->side[0].nodes[far[2 * b]].rank,
                   H->side[1].nodes[far[2 * b + 1]].rank);
  }
  for (uint32_t b = 0; b < nnear; b++) {
    H->pslot[nfar + b] = ps;
    ps += std::max(H->near[b].rows, H->near[b].cols);
  }
  H->pslot_len = ps;
  for (int sd = 0; sd < 2; sd++) {
    auto &S = H->side[sd];
    for (uint32_t b = 0; b < nnear && S.near_leaf; b++)
      S.near_leaf = S.nodes[near[2 * b + sd]].nsons == 0;
    auto csr = [&](const uint32_t *pairs, uint32_t cnt,
                   std::vector<uint32_t> &ptr, std::vector<uint32_t> &ids) {
      ptr.assign(S.nnodes + 1, 0);
      for (uint32_t b = 0; b < cnt; b++) ptr[pairs[2 * b + sd] + 1]++;
      for (uint32_t i = 0; i < S.nnodes; i++) ptr[i + 1] += ptr[i];
      ids.assign(cnt, 0);
      std::vector<uint32_t> pos(ptr.begin(), ptr.end() - 1);
      for (uint32_t b = 0; b < cnt; b++) ids[pos[pairs[2 * b + sd]]++] = b;
    };
    csr(far, nfar, S.far_ptr, S.far_ids);
    csr(near, nnear, S.near_ptr, S.near_ids);
    if (S.d_far_ptr.upload(S.far_ptr.data(), S.far_ptr.size()) !=
            hipSuccess ||
        S.d_far_ids.upload(S.far_ids.data(), S.far_ids.size()) !=
            hipSuccess ||
        S.d_near_ptr.upload(S.near_ptr.data(), S.near_ptr.size()) !=
            hipSuccess ||
        S.d_near_ids.upload(S.near_ids.data(), S.near_ids.size()) !=
            hipSuccess) {
      set_error("h2hip_h2matrix_create: device allocation failed");
      delete H;
      return H2HIP_ENOMEM;
    }
  }
  if (H->d_pslot.upload(H->pslot.data(), H->pslot.size()) != hipSuccess) {
    set_error("h2hip_h2matrix_create: device allocation failed");
    delete H;
    return H2HIP_ENOMEM;
  }
  const int cm = H->cplx ? 2 : 1;
  const uint64_t ncoef =
      std::max(H->side[0].ncoef, H->side[1].ncoef) * (uint64_t)cm;
  H->cstride = ncoef;
  for (int sd = 0; sd < 2; sd++)
    for (const NodeD &d : H->side[sd].nodes) {
      H->maxrank = std::max(H->maxrank, d.rank);
      if (d.nsons == 0) H->maxleaf = std::max(H->maxleaf, d.size);
    }
  if (H->d_near.upload(H->near.data(), H->near.size()) != hipSuccess ||
      H->d_far.upload(H->far.data(), H->far.size()) != hipSuccess ||
      H->d_N.alloc(cm * no) != hipSuccess ||
      H->d_S.alloc(cm * so) != hipSuccess ||
      H->d_xt.alloc(ncoef) != hipSuccess ||
      H->d_yt.alloc(ncoef) != hipSuccess) {
    set_error("h2hip_h2matrix_create: device allocation failed");
    delete H;
    return H2HIP_ENOMEM;
  }
  *out = H;
  return H2HIP_OK;
}

void h2hip_h2matrix_destroy(h2hip_h2matrix *H) { delete H; }

int h2hip_h2matrix_set_basis(h2hip_h2matrix *H, int side, const double *V,
                             const uint64_t *voff, const double *E,
                             uint64_t eld) {
  if (!H || side < 0 || side > 1 || !V || !voff || !E) return H2HIP_EINVAL;
  auto &S = H->side[side];
  std::vector<double> v(S.vlen), vt(S.vlen);
  std::vector<double> e(S.eld * S.nnodes, 0.0), et(S.eld * S.nnodes, 0.0);
  for (uint32_t i = 0; i < S.nnodes; i++) {
    const NodeD &d = S.nodes[i];
    if (d.nsons == 0) {
      const double *src = V + voff[i];
      for (uint32_t c = 0; c < d.rank; c++)
        for (uint32_t r = 0; r < d.size; r++) {
          v[d.voff + r + (uint64_t)c * d.size] = src[r + (uint64_t)c * d.size];
          vt[d.voff + c + (uint64_t)r * d.rank] =
              src[r + (uint64_t)c * d.size];
        }
    }
    if (d.father != NONE) {
      const uint32_t kf = S.nodes[d.father].rank;
      const double *src = E + (uint64_t)i * eld;
      for (uint32_t c = 0; c < kf; c++)
        for (uint32_t r = 0; r < d.rank; r++) {
          e[(uint64_t)i * S.eld + r + (uint64_t)c * d.rank] =
              src[r + (uint64_t)c * d.rank];
          et[(uint64_t)i * S.eld + c + (uint64_t)r * kf] =
              src[r + (uint64_t)c * d.rank];
        }
    }
  }
  // checksum of the side's bases (for the self-adjoint product, which
  // needs both sides equal)
  {
    double cs = 0.0, w = 1.0;
    for (size_t i = 0; i < v.size(); i++, w += 1e-7) cs += w * v[i];
    for (size_t i = 0; i < e.size(); i++, w += 1e-7) cs += w * e[i];
    H->basis_sum[side] = cs;
    H->basis_set[side] = 1;
  }
  H2HIP_CHECK(S.d_V.upload(v.data(), v.size()));
  H2HIP_CHECK(S.d_VT.upload(vt.data(), vt.size()));
  H2HIP_CHECK(S.d_E.upload(e.data(), e.size()));
  H2HIP_CHECK(S.d_ET.upload(et.data(), et.size()));
  // new bases: rebuild the flat transforms' blocks
  if (H->level_sync == 2 && H->side[0].d_VT.p && H->side[1].d_VT.p) {
    const int rc = build_flat_pair(H);
    if (rc != H2HIP_OK) {
      H->level_sync = 1;
      return rc;
    }
  }
  return H2HIP_OK;
}

int h2hip_h2matrix_set_couplings(h2hip_h2matrix *H, const void *S,
                                 uint64_t sld) {
  if (!H || (!S && H->nfar)) return H2HIP_EINVAL;
  H->adjT_ok = false;
  const int cm = H->cplx ? 2 : 1;
  std::vector<double> s(cm * H->slen);
  const double *src = (const double *)S;
  for (uint32_t b = 0; b < H->nfar; b++) {
    const NodeD &r = H->side[0].nodes[H->far[b].r];
    const NodeD &c = H->side[1].nodes[H->far[b].c];
    std::memcpy(&s[cm * H->far[b].soff], src + cm * (uint64_t)b * sld,
                sizeof(double) * cm * r.rank * c.rank);
  }
  H2HIP_CHECK(hipMemcpy(H->d_S.p, s.data(), sizeof(double) * s.size(),
                        hipMemcpyHostToDevice));
  // couplings given as values: products read them
  H->cpl_known = false;
  H->cpl_recompute = 0;
  return H2HIP_OK;
}

int h2hip_h2matrix_interp_bases(h2hip_h2matrix *H, int side,
                                const h2hip_surface *s, int basis, int dn,
                                uint32_t m, uint32_t q, const double *pts,
                                void *stream) {
  if (!H || side < 0 || side > 1 || !s || !pts || m < 1 || m > 5 ||
      (basis != 0 && basis != 1) || q < 1 || q > 8) {
    set_error("h2hip_h2matrix_interp_bases: invalid arguments");
    return H2HIP_EINVAL;
  }
  auto &S = H->side[side];
  const uint32_t k = m * m * m;
  for (const NodeD &d : S.nodes)
    if (d.rank != k) {
      set_error("h2hip_h2matrix_interp_bases: ranks differ from m^3");
      return H2HIP_EINVAL;
    }
  if (S.ndof != (basis ? s->nv : s->nt)) {
    set_error("h2hip_h2matrix_interp_bases: the tree's index range does not "
              "match the surface's degrees of freedom");
    return H2HIP_EINVAL;
  }
  if (basis) {
    const int rc = h2hip::surface_vertex_lists(s);
    if (rc != H2HIP_OK) return rc;
  }
  const h2hip::TriRule tr = h2hip::make_triangle_rule(q);
  IBArgs a;
  if (tr.w.size() > 64) return H2HIP_EINVAL;
  a.nq = (uint32_t)tr.w.size();
  for (uint32_t i = 0; i < a.nq; i++) {
    a.rx0[i] = tr.x0[i];
    a.rx1[i] = tr.x1[i];
    a.rw[i] = tr.w[i];
  }
  h2hip::DevBuf<double> d_pts;
  H2HIP_CHECK(d_pts.upload(pts, 3 * (uint64_t)m * S.nnodes));
  if (S.d_V.n != S.vlen) H2HIP_CHECK(S.d_V.alloc(S.vlen));
  if (S.d_VT.n != S.vlen) H2HIP_CHECK(S.d_VT.alloc(S.vlen));
  const uint64_t elen = S.eld * S.nnodes;
  if (S.d_E.n != elen) H2HIP_CHECK(S.d_E.alloc(elen));
  if (S.d_ET.n != elen) H2HIP_CHECK(S.d_ET.alloc(elen));
  hipStream_t st = (hipStream_t)stream;
  H2HIP_CHECK(hipMemsetAsync(S.d_E.p, 0, sizeof(double) * elen, st));
  H2HIP_CHECK(hipMemsetAsync(S.d_ET.p, 0, sizeof(double) * elen, st));
  a.nodes = S.d_nodes.p;
  a.idx = S.d_idx.p;
  a.leaves = S.d_leaves.p;
  a.nnodes = S.nnodes;
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.t = s->d_t.p;
  a.vt_ptr = basis ? s->d_vt_ptr.p : nullptr;
  a.vt = basis ? s->d_vt.p : nullptr;
  a.pts = d_pts.p;
  a.m = m;
  a.basis = basis;
  a.dn = dn ? 1 : 0;
  a.V = S.d_V.p;
  a.VT = S.d_VT.p;
  a.E = S.d_E.p;
  a.ET = S.d_ET.p;
  a.eld = S.eld;
  if (!S.leaves.empty())
    hipLaunchKernelGGL(interp_leaf, dim3((uint32_t)S.leaves.size()),
                       dim3(NT), 0, st, a);
  hipLaunchKernelGGL(interp_transfer, dim3(S.nnodes), dim3(NT), 0, st, a);
  H2HIP_CHECK(hipGetLastError());
  // the point buffer is freed on return
  H2HIP_CHECK(hipStreamSynchronize(st));
  // identity of the side's bases for the self-adjoint product: the same
  // surface, tree, basis, derivative, order and points
  {
    double cs = (double)(uintptr_t)s * 1e-12 + basis * 3.0 + dn * 7.0 +
                m * 11.0 + q * 13.0;
    double w = 1.0;
    for (uint64_t i = 0; i < 3 * (uint64_t)m * S.nnodes; i++, w += 1e-7)
      cs += w * pts[i];
    H->basis_sum[side] = cs;
    H->basis_set[side] = 1;
  }
  if (H->level_sync == 2 && H->basis_set[0] && H->basis_set[1]) {
    const int rc = build_flat_pair(H);
    if (rc != H2HIP_OK) {
      H->level_sync = 1;
      return rc;
    }
  }
  return H2HIP_OK;
}

int h2hip_h2matrix_set_coupling_mode(h2hip_h2matrix *H, int recompute) {
  if (!H || recompute < 0 || recompute > 1) {
    set_error("h2hip_h2matrix_set_coupling_mode: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (recompute && !H->cpl_known && H->nfar) {
    set_error("h2hip_h2matrix_set_coupling_mode: the couplings were not "
              "set by h2hip_h2matrix_interp_couplings");
    return H2HIP_EINVAL;
  }
  H->cpl_recompute = recompute;
  return H2HIP_OK;
}

int h2hip_h2matrix_set_nearfield(h2hip_h2matrix *H, const void *N) {
  if (!H || (!N && H->nlen)) return H2HIP_EINVAL;
  H->adjT_ok = false;
  const int cm = H->cplx ? 2 : 1;
  H2HIP_CHECK(hipMemcpy(H->d_N.p, N, sizeof(double) * cm * H->nlen,
                        hipMemcpyHostToDevice));
  return H2HIP_OK;
}

int h2hip_addevalsymm_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                               const void *x, void *y, void *stream) {
  if (!alpha || !H || !x || !y) {
    set_error("h2hip_addevalsymm_h2matrix: invalid arguments");
    return H2HIP_EINVAL;
  }
  const auto &r = H->side[0], &c = H->side[1];
  bool same = r.nnodes == c.nnodes && r.ncoef == c.ncoef &&
              H->basis_set[0] && H->basis_set[1] &&
              H->basis_sum[0] == H->basis_sum[1];
  for (uint32_t i = 0; same && i < r.nnodes; i++)
    same = r.nodes[i].off == c.nodes[i].off &&
           r.nodes[i].size == c.nodes[i].size &&
           r.nodes[i].rank == c.nodes[i].rank;
  if (!same) {
    set_error("h2hip_addevalsymm_h2matrix: row and column trees and bases "
              "must be equal");
    return H2HIP_EINVAL;
  }
  if (H->maxleaf > 128 || H->maxrank > 128) {
    set_error("h2hip_addevalsymm_h2matrix: leaves and ranks up to 128");
    return H2HIP_EINVAL;
  }
  return H->cplx ? apply_symm<1>(alpha, H, (const double *)x, (double *)y,
                                 (hipStream_t)stream)
                 : apply_symm<0>(alpha, H, (const double *)x, (double *)y,
                                 (hipStream_t)stream);
}

int h2hip_h2matrix_set_schedule(h2hip_h2matrix *H, int level_sync) {
  if (!H || level_sync < 0 || level_sync > 2) {
    set_error("h2hip_h2matrix_set_schedule: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (level_sync == 2) {
    const int rc = build_flat_pair(H);
    if (rc != H2HIP_OK) return rc;
  } else {
    drop_adjoint_flat(H);
    H->side[1].d_WF.release();
    H->side[0].d_WB.release();
  }
  H->level_sync = level_sync;
  return H2HIP_OK;
}

int h2hip_h2matrix_interp_couplings(h2hip_h2matrix *H, int kernel,
                                    const double *k, uint32_t m,
                                    const double *rpts, const double *cpts,
                                    void *stream) {
  if (H) H->adjT_ok = false;
  if (!H || m < 1 || m > 8 || (!rpts && H->nfar) || (!cpts && H->nfar) ||
      (kernel != H2HIP_LAPLACE && kernel != H2HIP_HELMHOLTZ) ||
      (kernel == H2HIP_HELMHOLTZ && (!k || !H->cplx)) ||
      (kernel == H2HIP_LAPLACE && H->cplx)) {
    set_error("h2hip_h2matrix_interp_couplings: invalid arguments");
    return H2HIP_EINVAL;
  }
  const uint32_t kr = m * m * m;
  for (int sd = 0; sd < 2; sd++)
    for (const NodeD &d : H->side[sd].nodes)
      if (d.rank != kr) {
        set_error("h2hip_h2matrix_interp_couplings: ranks differ from m^3");
        return H2HIP_EINVAL;
      }
  if (H->nfar == 0) return H2HIP_OK;
  const uint32_t rn = H->side[0].nnodes, cn = H->side[1].nnodes;
  // the points stay on the device for products that compute the couplings
  H->cpl_known = false;
  H2HIP_CHECK(H->d_rpts.upload(rpts, 3 * (uint64_t)m * rn));
  H2HIP_CHECK(H->d_cpts.upload(cpts, 3 * (uint64_t)m * cn));
  CplArgs a;
  a.far = H->d_far.p;
  a.nfar = H->nfar;
  a.rpts = H->d_rpts.p;
  a.cpts = H->d_cpts.p;
  a.m = m;
  a.helmholtz = kernel == H2HIP_HELMHOLTZ;
  a.kr = a.helmholtz ? k[0] : 0.0;
  a.kdamp = a.helmholtz ? -k[1] : 0.0;
  a.S = H->d_S.p;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(interp_couplings, dim3(H->nfar), dim3(NT), 0, st, a);
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipStreamSynchronize(st));
  H->cpl = CplSrc{H->d_rpts.p, H->d_cpts.p, m, a.helmholtz, a.kr, a.kdamp};
  H->cpl_known = true;
  // the pairs of mirrored far blocks, when the structure is symmetric
  H->npair = 0;
  if ((m == 4 || m == 5) && rn == cn &&
      std::memcmp(rpts, cpts, sizeof(double) * 3 * (uint64_t)m * rn) == 0) {
    std::vector<std::pair<uint64_t, uint32_t>> key(H->nfar);
    for (uint32_t f = 0; f < H->nfar; f++)
      key[f] = {(uint64_t)H->far[f].r << 32 | H->far[f].c, f};
    std::sort(key.begin(), key.end());
    std::vector<PairD> pairs;
    const std::vector<NodeD> &rnodes = H->side[0].nodes,
                             &cnodes = H->side[1].nodes;
    bool symmetric = true;
    for (uint32_t f = 0; f < H->nfar && symmetric; f++) {
      const FarD &b = H->far[f];
      const uint64_t mirror = (uint64_t)b.c << 32 | b.r;
      const auto it = std::lower_bound(
          key.begin(), key.end(), std::make_pair(mirror, 0u));
      symmetric = b.r != b.c && it != key.end() && it->first == mirror;
      if (b.r > b.c)
        pairs.push_back(PairD{b.r, b.c, rnodes[b.r].coef, cnodes[b.c].coef,
                              rnodes[b.c].coef, cnodes[b.r].coef});
    }
    if (symmetric && !pairs.empty()) {
      H2HIP_CHECK(H->d_pairs.upload(pairs.data(), pairs.size()));
      H->npair = (uint32_t)pairs.size();
    }
  }
  return H2HIP_OK;
}

int h2hip_h2matrix_set_coupling_pairs(h2hip_h2matrix *H, int on) {
  if (!H || on < 0 || on > 1) {
    set_error("h2hip_h2matrix_set_coupling_pairs: invalid arguments");
    return H2HIP_EINVAL;
  }
  H->pairs_on = on;
  return H2HIP_OK;
}

int h2hip_h2matrix_coupling_pairs(const h2hip_h2matrix *H) {
  return H ? (int)H->npair : -1;
}

void *h2hip_h2matrix_nearfield_device(h2hip_h2matrix *H) {
  if (H) H->adjT_ok = false;  // the caller writes the near field
  return H ? (void *)H->d_N.p : nullptr;
}

int h2hip_h2matrix_bytes(const h2hip_h2matrix *H, uint64_t *out3) {
  if (!H || !out3) return H2HIP_EINVAL;
  const int cm = H->cplx ? 2 : 1;
  out3[0] = 8ull * cm * H->nlen;
  // products that compute the couplings read none of them
  out3[1] = H->cpl_recompute ? 0 : 8ull * cm * H->slen;
  uint64_t b = 0;
  if (H->level_sync == 2 && H->side[1].d_WF.p && H->side[0].d_WB.p) {
    // the flat transforms read the blocks W_a|_l of both sides
    out3[2] = 8ull * (H->side[0].wlen + H->side[1].wlen);
    return H2HIP_OK;
  }
  for (int s = 0; s < 2; s++) {
    b += 8ull * H->side[s].vlen;
    for (uint32_t i = 0; i < H->side[s].nnodes; i++) {
      const NodeD &d = H->side[s].nodes[i];
      if (d.father != NONE)
        b += 8ull * d.rank * H->side[s].nodes[d.father].rank;
    }
  }
  out3[2] = b;
  return H2HIP_OK;
}

int h2hip_h2matrix_storage(const h2hip_h2matrix *H, uint64_t *out4) {
  if (!H || !out4) return H2HIP_EINVAL;
  const uint64_t cm = H->cplx ? 2 : 1;
  out4[0] = 8ull * cm * H->nlen;
  // every admissible block's k_r x k_c coupling, stored or computed
  out4[1] = 8ull * cm * H->slen;
  // leaf and transfer matrices of both sides, one copy each (real)
  uint64_t b = 0;
  for (int s = 0; s < 2; s++) {
    b += 8ull * H->side[s].vlen;
    for (const NodeD &d : H->side[s].nodes)
      if (d.father != NONE)
        b += 8ull * d.rank * H->side[s].nodes[d.father].rank;
  }
  out4[2] = b;
  // every device buffer the matrix holds
  uint64_t r = 0;
  auto add = [&r](const auto &buf) { r += sizeof(*buf.p) * buf.n; };
  for (int s = 0; s < 2; s++) {
    const auto &S = H->side[s];
    add(S.d_items);
    add(S.d_WF);
    add(S.d_WB);
    add(S.d_nodes);
    add(S.d_idx);
    add(S.d_leaves);
    add(S.d_sons);
    add(S.d_paths);
    add(S.d_plen);
    add(S.d_cnt);
    add(S.d_fwork);
    add(S.d_hnodes);
    add(S.d_V);
    add(S.d_VT);
    add(S.d_E);
    add(S.d_ET);
  }
  add(H->d_far);
  add(H->d_near);
  add(H->d_S);
  add(H->d_N);
  add(H->d_xt);
  add(H->d_yt);
  add(H->d_xtf);
  add(H->d_ytf);
  add(H->d_rpts);
  add(H->d_cpts);
  add(H->d_xtb);
  add(H->d_ytb);
  add(H->d_farT);
  add(H->d_nearT);
  add(H->d_ST);
  add(H->d_NT);
  for (int s = 0; s < 2; s++) {
    add(H->side[s].d_WFt);
    add(H->side[s].d_WBt);
  }
  out4[3] = r;
  return H2HIP_OK;
}

int h2hip_h2matrix_set_mfma(h2hip_h2matrix *H, int on) {
  if (!H || on < 0 || on > 1) {
    set_error("h2hip_h2matrix_set_mfma: invalid arguments");
    return H2HIP_EINVAL;
  }
  H->cpl_mfma = on;
  return H2HIP_OK;
}

int h2hip_h2matrix_set_adjoint_storage(h2hip_h2matrix *H, int keep) {
  if (!H || keep < 0 || keep > 1) {
    set_error("h2hip_h2matrix_set_adjoint_storage: invalid arguments");
    return H2HIP_EINVAL;
  }
  std::lock_guard<std::mutex> guard(H->adj_lock);
  H->adj_keep = keep;
  H->adjT_ok = false;  // made again by the next adjoint product
  if (!keep) {
    H->d_farT.release();
    H->d_nearT.release();
    H->d_ST.release();
    H->d_NT.release();
  }
  return H2HIP_OK;
}

int h2hip_addeval_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                           const void *d_x, void *d_y, void *stream) {
  if (!alpha || !H || !d_x || !d_y) return H2HIP_EINVAL;
  return H->cplx ? apply<1>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, false)
                 : apply<0>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, false);
}

int h2hip_addeval_h2matrix_batch(const h2hip_h2matrix *H, uint32_t nvec,
                                 const double *alpha, const void *const *d_x,
                                 void *const *d_y, int trans, void *stream) {
  if (!H || !alpha || !d_x || !d_y || nvec == 0 || nvec > MAXNV) {
    set_error("h2hip_addeval_h2matrix_batch: invalid arguments (1..8 "
              "vectors)");
    return H2HIP_EINVAL;
  }
  for (uint32_t v = 0; v < nvec; v++)
    if (!d_x[v] || !d_y[v]) {
      set_error("h2hip_addeval_h2matrix_batch: null vector");
      return H2HIP_EINVAL;
    }
  return H->cplx ? apply_batch<1>(nvec, alpha, H, d_x, d_y,
                                  (hipStream_t)stream, trans != 0)
                 : apply_batch<0>(nvec, alpha, H, d_x, d_y,
                                  (hipStream_t)stream, trans != 0);
}

int h2hip_addevaltrans_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                                const void *d_x, void *d_y, void *stream) {
  if (!alpha || !H || !d_x || !d_y) return H2HIP_EINVAL;
  return H->cplx ? apply<1>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, true)
                 : apply<0>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, true);
}

}  // extern "C"
