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
//  1. fwd_up: one wavefront per leaf computes xt_t = V_t^T x|_t, then climbs
//     the tree: the last son to finish (an atomic counter per cluster, reset
//     by that son) computes xt_f = sum_s E_s^T xt_s for its father and goes
//     on upwards.  No wavefront ever waits for another one.  The same launch
//     clears the coupling accumulator yt.
//  2. couple_near: one wavefront per admissible block, yt_r += S_rc xt_c
//     (FP64 atomics), and one per near-field block, y|_r += alpha N_rc x|_c.
//  3. down_back: one wavefront per leaf of the output tree walks its path
//     from the root, v = yt_s + E_s v, and adds y|_t += alpha V_t v.
//
// Every small product uses lane = output entry and loops over the input
// entries, with each matrix stored in the orientation that makes that loop
// read 64 consecutive entries per step (leaf and transfer matrices are kept
// in both orientations, couplings and near blocks column-major).  The
// products are bound by HBM bandwidth: two flops per stored entry.
#include <algorithm>
#include <cstring>

#include "h2hip_internal.h"

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

// load written by another wavefront of the same launch
__device__ __forceinline__ double ld_agent(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

struct FwdArgs {
  const NodeD *nodes;
  const uint32_t *leaves;
  uint32_t nleaves;
  const uint32_t *idx;
  const uint32_t *sons;
  const double *VT;  // leaf: rank x size column-major
  const double *ET;  // per son s: rank_father x rank_s column-major
  uint64_t eld;
  const double *x;
  double *xt;
  uint32_t *cnt;
  double *yt;  // cleared here (doubles)
  uint64_t nyt;
};

template <int C, int K>
__global__ __launch_bounds__(64) void fwd_up(FwdArgs a) {
  __shared__ double sv[2 * MAXRANK];
  const uint32_t w = blockIdx.x;
  const int lane = threadIdx.x;
  {  // clear a slice of the coupling accumulator
    const uint64_t chunk = (a.nyt + a.nleaves - 1) / a.nleaves;
    const uint64_t e0 = (uint64_t)w * chunk, e1 = min(a.nyt, e0 + chunk);
    for (uint64_t e = e0 + lane; e < e1; e += 64) a.yt[e] = 0.0;
  }
  uint32_t node = a.leaves[w];
  {
    const NodeD t = a.nodes[node];
    const uint32_t rank = K ? K : t.rank;
    // x|_t into LDS (leaf size <= MAXRANK is checked at creation)
    for (uint32_t i = lane; i < t.size; i += 64) {
      double xr, xi;
      ld<C>(a.x, a.idx[t.off + i], xr, xi);
      sv[2 * i] = xr;
      sv[2 * i + 1] = xi;
    }
    __syncthreads();
    const double *Vt = a.VT + t.voff;
    for (uint32_t j = lane; j < rank; j += 64) {
      double ar = 0.0, ai = 0.0;
#pragma unroll 8
      for (uint32_t i = 0; i < t.size; i++) {
        const double v = Vt[j + (uint64_t)i * rank];
        ar = fma(v, sv[2 * i], ar);
        if constexpr (C) ai = fma(v, sv[2 * i + 1], ai);
      }
      st<C>(a.xt, t.coef + j, ar, ai);
    }
  }
  // climb: the last son of a cluster computes the cluster
  while (true) {
    const uint32_t f = a.nodes[node].father;
    if (f == NONE) break;
    __threadfence();  // release this wavefront's coefficients
    uint32_t old = 0;
    if (lane == 0) old = atomicAdd(&a.cnt[f], 1u);
    old = __shfl(old, 0, 64);
    const NodeD F = a.nodes[f];
    if (old + 1 < F.nsons) break;
    __threadfence();  // acquire the other sons' coefficients
    if (lane == 0) a.cnt[f] = 0;
    const uint32_t rank = K ? K : F.rank;
    double accr[MAXRANK / 64], acci[MAXRANK / 64];
#pragma unroll
    for (int c = 0; c < MAXRANK / 64; c++) accr[c] = acci[c] = 0.0;
    for (uint32_t so = 0; so < F.nsons; so++) {
      const uint32_t sn = a.sons[F.sonbeg + so];
      const NodeD s = a.nodes[sn];
      const uint32_t srank = K ? K : s.rank;
      __syncthreads();
      for (uint32_t i = lane; i < srank; i += 64) {
        if constexpr (C) {
          sv[2 * i] = ld_agent(a.xt + 2 * (s.coef + i));
          sv[2 * i + 1] = ld_agent(a.xt + 2 * (s.coef + i) + 1);
        } else {
          sv[2 * i] = ld_agent(a.xt + s.coef + i);
        }
      }
      __syncthreads();
      const double *E = a.ET + (uint64_t)sn * a.eld;
#pragma unroll
      for (int c = 0; c < MAXRANK / 64; c++) {
        const uint32_t j = lane + 64 * c;
        if (j < rank) {
#pragma unroll 8
          for (uint32_t i = 0; i < srank; i++) {
            const double e = E[j + (uint64_t)i * rank];
            accr[c] = fma(e, sv[2 * i], accr[c]);
            if constexpr (C) acci[c] = fma(e, sv[2 * i + 1], acci[c]);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < MAXRANK / 64; c++) {
      const uint32_t j = lane + 64 * c;
      if (j < rank) st<C>(a.xt, F.coef + j, accr[c], acci[c]);
    }
    node = f;
  }
}

struct CNArgs {
  const FarD *far;
  uint32_t nfar;
  const NearD *near;
  uint32_t nnear;
  const NodeD *rn, *cn;
  const uint32_t *ridx, *cidx;
  const double *S, *N;
  const double *x, *xt;
  double *yt, *y;
  double are, aim;
};

// yt_r += S_rc xt_c and y|_r += alpha N_rc x|_c; TRANS: yt_c += S^* xt_r and
// y|_c += alpha N^* x|_r
template <int C, bool TRANS>
__global__ __launch_bounds__(NT) void couple_near(CNArgs a) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w < a.nfar) {
    const FarD b = a.far[w];
    const NodeD r = a.rn[b.r], c = a.cn[b.c];
    const NodeD out = TRANS ? c : r, in = TRANS ? r : c;
    for (uint32_t i = lane; i < out.rank; i += 64) {
      double ar = 0.0, ai = 0.0;
#pragma unroll 8
      for (uint32_t j = 0; j < in.rank; j++) {
        const uint64_t e = TRANS ? (uint64_t)j + (uint64_t)i * r.rank
                                 : (uint64_t)i + (uint64_t)j * r.rank;
        double sr, si, xr, xi;
        ld<C>(a.S, b.soff + e, sr, si);
        if (TRANS) si = -si;
        ld<C>(a.xt, in.coef + j, xr, xi);
        ar = fma(sr, xr, ar);
        if constexpr (C) {
          ar = fma(-si, xi, ar);
          ai = fma(sr, xi, fma(si, xr, ai));
        }
      }
      atomic_add<C>(a.yt, out.coef + i, ar, ai);
    }
    return;
  }
  const uint32_t nb = w - a.nfar;
  if (nb >= a.nnear) return;
  const NearD b = a.near[nb];
  const uint32_t nout = TRANS ? b.cols : b.rows, nin = TRANS ? b.rows : b.cols;
  const uint32_t *oidx = TRANS ? a.cidx + b.coff : a.ridx + b.roff;
  const uint32_t *iidx = TRANS ? a.ridx + b.roff : a.cidx + b.coff;
  for (uint32_t i = lane; i < nout; i += 64) {
    double ar = 0.0, ai = 0.0;
#pragma unroll 8
    for (uint32_t j = 0; j < nin; j++) {
      const uint64_t e = TRANS ? (uint64_t)j + (uint64_t)i * b.rows
                               : (uint64_t)i + (uint64_t)j * b.rows;
      double nr, ni, xr, xi;
      ld<C>(a.N, b.noff + e, nr, ni);
      if (TRANS) ni = -ni;
      ld<C>(a.x, iidx[j], xr, xi);
      ar = fma(nr, xr, ar);
      if constexpr (C) {
        ar = fma(-ni, xi, ar);
        ai = fma(nr, xi, fma(ni, xr, ai));
      }
    }
    atomic_add<C>(a.y, oidx[i], a.are * ar - a.aim * ai,
                  a.are * ai + a.aim * ar);
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
  double *y;
  double are, aim;
};

// v = yt_root; for each node s on the path below the root v = yt_s + E_s v;
// at the leaf y|_t += alpha V_t v
template <int C, int K>
__global__ __launch_bounds__(64) void down_back(DBArgs a) {
  __shared__ double sv[2 * MAXRANK];
  const uint32_t w = blockIdx.x;
  const int lane = threadIdx.x;
  const uint32_t *path = a.paths + (uint64_t)w * MAXPATH;
  const uint32_t len = a.plen[w];
  {
    const NodeD r = a.nodes[path[0]];
    for (uint32_t i = lane; i < r.rank; i += 64) {
      double vr, vi;
      ld<C>(a.yt, r.coef + i, vr, vi);
      sv[2 * i] = vr;
      sv[2 * i + 1] = vi;
    }
  }
  for (uint32_t p = 1; p < len; p++) {
    const uint32_t sn = path[p];
    const NodeD s = a.nodes[sn];
    const uint32_t kf = a.nodes[s.father].rank;
    const uint32_t rank = K ? K : s.rank;
    const double *E = a.E + (uint64_t)sn * a.eld;
    double accr[MAXRANK / 64], acci[MAXRANK / 64];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXRANK / 64; c++) {
      accr[c] = acci[c] = 0.0;
      const uint32_t i = lane + 64 * c;
      if (i < rank) {
        ld<C>(a.yt, s.coef + i, accr[c], acci[c]);
#pragma unroll 8
        for (uint32_t j = 0; j < kf; j++) {
          const double e = E[i + (uint64_t)j * rank];
          accr[c] = fma(e, sv[2 * j], accr[c]);
          if constexpr (C) acci[c] = fma(e, sv[2 * j + 1], acci[c]);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXRANK / 64; c++) {
      const uint32_t i = lane + 64 * c;
      if (i < rank) {
        sv[2 * i] = accr[c];
        sv[2 * i + 1] = acci[c];
      }
    }
  }
  __syncthreads();
  const NodeD t = a.nodes[a.leaves[w]];
  const uint32_t rank = K ? K : t.rank;
  const double *Vt = a.V + t.voff;
  for (uint32_t i = lane; i < t.size; i += 64) {
    double ar = 0.0, ai = 0.0;
#pragma unroll 8
    for (uint32_t j = 0; j < rank; j++) {
      const double v = Vt[i + (uint64_t)j * t.size];
      ar = fma(v, sv[2 * j], ar);
      if constexpr (C) ai = fma(v, sv[2 * j + 1], ai);
    }
    const uint32_t d = a.idx[t.off + i];
    if constexpr (C) {
      a.y[2 * d] += a.are * ar - a.aim * ai;
      a.y[2 * d + 1] += a.are * ai + a.aim * ar;
    } else {
      a.y[d] += a.are * ar;
    }
  }
}

}  // namespace

struct h2hip_h2matrix {
  int cplx = 0;
  uint32_t urank = 0;  // common rank of all clusters, 0 if they differ
  struct Side {
    uint32_t nnodes = 0, ndof = 0, depth = 0;
    std::vector<NodeD> nodes;
    std::vector<uint32_t> leaves, sons, paths, plen;
    uint64_t ncoef = 0, vlen = 0, eld = 0;
    h2hip::DevBuf<NodeD> d_nodes;
    h2hip::DevBuf<uint32_t> d_idx, d_leaves, d_sons, d_paths, d_plen, d_cnt;
    h2hip::DevBuf<double> d_V, d_VT, d_E, d_ET;
  } side[2];
  uint32_t nnear = 0, nfar = 0;
  std::vector<FarD> far;
  std::vector<NearD> near;
  uint64_t nlen = 0, slen = 0;
  h2hip::DevBuf<FarD> d_far;
  h2hip::DevBuf<NearD> d_near;
  h2hip::DevBuf<double> d_S, d_N, d_xt, d_yt;
};

using h2hip::set_error;

namespace {

int build_side(h2hip_h2matrix::Side &S, uint32_t nnodes, const uint32_t *tree,
               const uint32_t *rank, uint32_t ndof, const uint32_t *idx) {
  S.nnodes = nnodes;
  S.ndof = ndof;
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
  S.ncoef = coef;
  S.vlen = v;
  S.eld = maxr * maxr;
  std::vector<uint32_t> zero(nnodes, 0);
  if (S.d_nodes.upload(S.nodes.data(), nnodes) != hipSuccess ||
      S.d_idx.upload(idx, ndof) != hipSuccess ||
      S.d_leaves.upload(S.leaves.data(), S.leaves.size()) != hipSuccess ||
      S.d_sons.upload(S.sons.data(), S.sons.size()) != hipSuccess ||
      S.d_paths.upload(S.paths.data(), S.paths.size()) != hipSuccess ||
      S.d_plen.upload(S.plen.data(), S.plen.size()) != hipSuccess ||
      S.d_cnt.upload(zero.data(), nnodes) != hipSuccess) {
    set_error("h2matrix: device allocation failed");
    return H2HIP_ENOMEM;
  }
  return H2HIP_OK;
}

template <int C, int K>
void launch(const double *alpha, const h2hip_h2matrix *H, const double *x,
            double *y, hipStream_t st, bool trans) {
  const auto &in = H->side[trans ? 0 : 1];
  const auto &out = H->side[trans ? 1 : 0];
  FwdArgs f;
  f.nodes = in.d_nodes.p;
  f.leaves = in.d_leaves.p;
  f.nleaves = (uint32_t)in.leaves.size();
  f.idx = in.d_idx.p;
  f.sons = in.d_sons.p;
  f.VT = in.d_VT.p;
  f.ET = in.d_ET.p;
  f.eld = in.eld;
  f.x = x;
  f.xt = H->d_xt.p;
  f.cnt = in.d_cnt.p;
  f.yt = H->d_yt.p;
  f.nyt = (C ? 2 : 1) * out.ncoef;
  hipLaunchKernelGGL((fwd_up<C, K>), dim3(f.nleaves), dim3(64), 0, st, f);

  CNArgs c;
  c.far = H->d_far.p;
  c.nfar = H->nfar;
  c.near = H->d_near.p;
  c.nnear = H->nnear;
  c.rn = H->side[0].d_nodes.p;
  c.cn = H->side[1].d_nodes.p;
  c.ridx = H->side[0].d_idx.p;
  c.cidx = H->side[1].d_idx.p;
  c.S = H->d_S.p;
  c.N = H->d_N.p;
  c.x = x;
  c.xt = H->d_xt.p;
  c.yt = H->d_yt.p;
  c.y = y;
  c.are = alpha[0];
  c.aim = alpha[1];
  const uint32_t nw = H->nfar + H->nnear;
  if (nw) {
    if (trans)
      hipLaunchKernelGGL((couple_near<C, true>), dim3((nw + WPB - 1) / WPB),
                         dim3(NT), 0, st, c);
    else
      hipLaunchKernelGGL((couple_near<C, false>), dim3((nw + WPB - 1) / WPB),
                         dim3(NT), 0, st, c);
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
  d.yt = H->d_yt.p;
  d.y = y;
  d.are = alpha[0];
  d.aim = alpha[1];
  hipLaunchKernelGGL((down_back<C, K>), dim3(d.nleaves), dim3(64), 0, st, d);
}

template <int C>
int apply(const double *alpha, const h2hip_h2matrix *H, const double *x,
          double *y, hipStream_t st, bool trans) {
  switch (H->urank) {
    case 8: launch<C, 8>(alpha, H, x, y, st, trans); break;
    case 27: launch<C, 27>(alpha, H, x, y, st, trans); break;
    case 64: launch<C, 64>(alpha, H, x, y, st, trans); break;
    case 125: launch<C, 125>(alpha, H, x, y, st, trans); break;
    default: launch<C, 0>(alpha, H, x, y, st, trans); break;
  }
  H2HIP_CHECK(hipGetLastError());
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
  int rc = build_side(H->side[0], rnodes, rtree, rrank, nrdof, ridx);
  if (rc == H2HIP_OK)
    rc = build_side(H->side[1], cnodes, ctree, crank, ncdof, cidx);
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
  const int cm = H->cplx ? 2 : 1;
  const uint64_t ncoef =
      std::max(H->side[0].ncoef, H->side[1].ncoef) * (uint64_t)cm;
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
  H2HIP_CHECK(S.d_V.upload(v.data(), v.size()));
  H2HIP_CHECK(S.d_VT.upload(vt.data(), vt.size()));
  H2HIP_CHECK(S.d_E.upload(e.data(), e.size()));
  H2HIP_CHECK(S.d_ET.upload(et.data(), et.size()));
  return H2HIP_OK;
}

int h2hip_h2matrix_set_couplings(h2hip_h2matrix *H, const void *S,
                                 uint64_t sld) {
  if (!H || (!S && H->nfar)) return H2HIP_EINVAL;
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
  return H2HIP_OK;
}

int h2hip_h2matrix_set_nearfield(h2hip_h2matrix *H, const void *N) {
  if (!H || (!N && H->nlen)) return H2HIP_EINVAL;
  const int cm = H->cplx ? 2 : 1;
  H2HIP_CHECK(hipMemcpy(H->d_N.p, N, sizeof(double) * cm * H->nlen,
                        hipMemcpyHostToDevice));
  return H2HIP_OK;
}

void *h2hip_h2matrix_nearfield_device(h2hip_h2matrix *H) {
  return H ? (void *)H->d_N.p : nullptr;
}

int h2hip_h2matrix_bytes(const h2hip_h2matrix *H, uint64_t *out3) {
  if (!H || !out3) return H2HIP_EINVAL;
  const int cm = H->cplx ? 2 : 1;
  out3[0] = 8ull * cm * H->nlen;
  out3[1] = 8ull * cm * H->slen;
  uint64_t b = 0;
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

int h2hip_addeval_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                           const void *d_x, void *d_y, void *stream) {
  if (!alpha || !H || !d_x || !d_y) return H2HIP_EINVAL;
  return H->cplx ? apply<1>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, false)
                 : apply<0>(alpha, H, (const double *)d_x, (double *)d_y,
                            (hipStream_t)stream, false);
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
