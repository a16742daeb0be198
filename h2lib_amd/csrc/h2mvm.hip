// H2-matrix products y += alpha H x and y += alpha H^* x on gfx950.
//
// The reference evaluates the product by three recursions over its pointer
// trees (addeval_h2matrix_avector, Library/h2matrix.c:1268-1286):
// forward transform of x into the column basis coefficients
// (Library/clusterbasis.c:523-573), coupling and near-field blocks
// (fastaddeval_h2matrix_avector, Library/h2matrix.c:1210-1266), backward
// transform into y (Library/clusterbasis.c:748-797).  Here the trees are
// flattened into level lists and every small matrix-vector product is one
// wavefront, lane = output entry, looping over the input entries: each
// matrix is stored in the orientation that makes that loop read 64
// consecutive entries per step (leaf and transfer matrices in both
// orientations, couplings and near blocks column-major for H x).
//
//   1. forward, leaves:   xt_t = V_t^T x|_t               (one wave per leaf)
//   2. forward, upwards:  xt_t = sum_sons E_s^T xt_s       (one launch/level)
//   3. couplings:         yt_r += S_rc xt_c                (wave per block,
//                                                           FP64 atomics)
//   4. backward, down:    yt_s += E_s yt_father            (one launch/level)
//   5. backward, leaves:  y|_t += alpha V_t yt_t           (one wave per leaf)
//   6. near field:        y|_r += alpha N_rc x|_c          (wave per block,
//                                                           FP64 atomics)
// All phases are bound by HBM bandwidth: two flops per stored entry.
#include <algorithm>
#include <cstring>

#include "h2hip_internal.h"

namespace {

constexpr int NT = 256;
constexpr int WPB = NT / 64;

struct NodeD {
  uint32_t off, size, rank, father;
  uint64_t coef;  // offset of the node's coefficients
  uint64_t voff;  // offset of the leaf matrix
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
__device__ __forceinline__ void atomic_add(double *p, uint64_t i, double re,
                                           double im) {
  if constexpr (C) {
    atomicAdd(&p[2 * i], re);
    atomicAdd(&p[2 * i + 1], im);
  } else {
    atomicAdd(&p[i], re);
  }
}

__global__ void zero_kernel(double *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// xt_t = V_t^T x|_t; VT of leaf t is rank x size column-major
template <int C>
__global__ __launch_bounds__(NT) void fwd_leaf(
    const NodeD *__restrict__ nodes, const uint32_t *__restrict__ leaves,
    uint32_t nleaves, const uint32_t *__restrict__ idx,
    const double *__restrict__ VT, const double *__restrict__ x,
    double *__restrict__ xt) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nleaves) return;
  const NodeD t = nodes[leaves[w]];
  const double *Vt = VT + t.voff;
  for (uint32_t j = lane; j < t.rank; j += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t i = 0; i < t.size; i++) {
      const double v = Vt[j + (uint64_t)i * t.rank];
      double xr, xi;
      ld<C>(x, idx[t.off + i], xr, xi);
      ar = fma(v, xr, ar);
      ai = fma(v, xi, ai);
    }
    if constexpr (C) {
      xt[2 * (t.coef + j)] = ar;
      xt[2 * (t.coef + j) + 1] = ai;
    } else {
      xt[t.coef + j] = ar;
    }
  }
}

// xt_t = sum over sons s of E_s^T xt_s; ET of son s is rank_t x rank_s
// column-major
template <int C>
__global__ __launch_bounds__(NT) void up_level(
    const NodeD *__restrict__ nodes, const uint32_t *__restrict__ list,
    uint32_t n, const uint32_t *__restrict__ sonptr,
    const uint32_t *__restrict__ sons, const double *__restrict__ ET,
    uint64_t eld, double *__restrict__ xt) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= n) return;
  const uint32_t tn = list[w];
  const NodeD t = nodes[tn];
  for (uint32_t j = lane; j < t.rank; j += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t so = sonptr[tn]; so < sonptr[tn + 1]; so++) {
      const uint32_t sn = sons[so];
      const NodeD s = nodes[sn];
      const double *E = ET + (uint64_t)sn * eld;
      for (uint32_t i = 0; i < s.rank; i++) {
        const double e = E[j + (uint64_t)i * t.rank];
        double xr, xi;
        ld<C>(xt, s.coef + i, xr, xi);
        ar = fma(e, xr, ar);
        ai = fma(e, xi, ai);
      }
    }
    if constexpr (C) {
      xt[2 * (t.coef + j)] = ar;
      xt[2 * (t.coef + j) + 1] = ai;
    } else {
      xt[t.coef + j] = ar;
    }
  }
}

// yt_r += S_rc xt_c (trans: yt_c += S_rc^* xt_r)
template <int C, bool TRANS>
__global__ __launch_bounds__(NT) void coupling(
    const FarD *__restrict__ far, uint32_t nfar,
    const NodeD *__restrict__ rn, const NodeD *__restrict__ cn,
    const double *__restrict__ S, const double *__restrict__ xt,
    double *__restrict__ yt) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nfar) return;
  const FarD b = far[w];
  const NodeD r = rn[b.r], c = cn[b.c];
  const NodeD out = TRANS ? c : r, in = TRANS ? r : c;
  for (uint32_t i = lane; i < out.rank; i += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t j = 0; j < in.rank; j++) {
      // entry S[rr, cc], rr = row of S, cc = column of S
      const uint64_t e = TRANS ? (uint64_t)j + (uint64_t)i * r.rank
                               : (uint64_t)i + (uint64_t)j * r.rank;
      double sr, si, xr, xi;
      ld<C>(S, b.soff + e, sr, si);
      if (TRANS) si = -si;
      ld<C>(xt, in.coef + j, xr, xi);
      ar = fma(sr, xr, ar);
      if constexpr (C) {
        ar = fma(-si, xi, ar);
        ai = fma(sr, xi, fma(si, xr, ai));
      }
    }
    atomic_add<C>(yt, out.coef + i, ar, ai);
  }
}

// yt_s += E_s yt_father; E of node s is rank_s x rank_father column-major
template <int C>
__global__ __launch_bounds__(NT) void down_level(
    const NodeD *__restrict__ nodes, const uint32_t *__restrict__ list,
    uint32_t n, const double *__restrict__ E, uint64_t eld,
    double *__restrict__ yt) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= n) return;
  const uint32_t sn = list[w];
  const NodeD s = nodes[sn];
  const NodeD f = nodes[s.father];
  const double *Es = E + (uint64_t)sn * eld;
  for (uint32_t i = lane; i < s.rank; i += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t j = 0; j < f.rank; j++) {
      const double e = Es[i + (uint64_t)j * s.rank];
      double yr, yi;
      ld<C>(yt, f.coef + j, yr, yi);
      ar = fma(e, yr, ar);
      ai = fma(e, yi, ai);
    }
    if constexpr (C) {
      yt[2 * (s.coef + i)] += ar;
      yt[2 * (s.coef + i) + 1] += ai;
    } else {
      yt[s.coef + i] += ar;
    }
  }
}

// y|_t += alpha V_t yt_t; V of leaf t is size x rank column-major
template <int C>
__global__ __launch_bounds__(NT) void back_leaf(
    const NodeD *__restrict__ nodes, const uint32_t *__restrict__ leaves,
    uint32_t nleaves, const uint32_t *__restrict__ idx,
    const double *__restrict__ V, const double *__restrict__ yt, double are,
    double aim, double *__restrict__ y) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nleaves) return;
  const NodeD t = nodes[leaves[w]];
  const double *Vt = V + t.voff;
  for (uint32_t i = lane; i < t.size; i += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t j = 0; j < t.rank; j++) {
      const double v = Vt[i + (uint64_t)j * t.size];
      double yr, yi;
      ld<C>(yt, t.coef + j, yr, yi);
      ar = fma(v, yr, ar);
      ai = fma(v, yi, ai);
    }
    const uint32_t d = idx[t.off + i];
    if constexpr (C) {
      y[2 * d] += are * ar - aim * ai;
      y[2 * d + 1] += are * ai + aim * ar;
    } else {
      y[d] += are * ar;
    }
  }
}

// y|_r += alpha N x|_c (trans: y|_c += alpha N^* x|_r)
template <int C, bool TRANS>
__global__ __launch_bounds__(NT) void near_blocks(
    const NearD *__restrict__ nb, uint32_t nnear,
    const uint32_t *__restrict__ ridx, const uint32_t *__restrict__ cidx,
    const double *__restrict__ N, const double *__restrict__ x, double are,
    double aim, double *__restrict__ y) {
  const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nnear) return;
  const NearD b = nb[w];
  const uint32_t nout = TRANS ? b.cols : b.rows, nin = TRANS ? b.rows : b.cols;
  const uint32_t *oidx = TRANS ? cidx + b.coff : ridx + b.roff;
  const uint32_t *iidx = TRANS ? ridx + b.roff : cidx + b.coff;
  for (uint32_t i = lane; i < nout; i += 64) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t j = 0; j < nin; j++) {
      const uint64_t e = TRANS ? (uint64_t)j + (uint64_t)i * b.rows
                               : (uint64_t)i + (uint64_t)j * b.rows;
      double nr, ni, xr, xi;
      ld<C>(N, b.noff + e, nr, ni);
      if (TRANS) ni = -ni;
      ld<C>(x, iidx[j], xr, xi);
      ar = fma(nr, xr, ar);
      if constexpr (C) {
        ar = fma(-ni, xi, ar);
        ai = fma(nr, xi, fma(ni, xr, ai));
      }
    }
    atomic_add<C>(y, oidx[i], are * ar - aim * ai, are * ai + aim * ar);
  }
}

}  // namespace

struct h2hip_h2matrix {
  int cplx = 0;
  struct Side {
    uint32_t nnodes = 0, ndof = 0, depth = 0;
    std::vector<NodeD> nodes;
    std::vector<uint32_t> level;
    std::vector<uint32_t> leaves;
    std::vector<std::vector<uint32_t>> inner_at, all_at;  // per level
    std::vector<uint32_t> sonptr, sons;
    uint64_t ncoef = 0, vlen = 0, eld = 0;
    h2hip::DevBuf<NodeD> d_nodes;
    h2hip::DevBuf<uint32_t> d_idx, d_leaves, d_sonptr, d_sons, d_lists;
    std::vector<uint64_t> inner_off, all_off;  // into d_lists
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
  S.level.assign(nnodes, 0);
  std::vector<uint32_t> nsons(nnodes, 0);
  uint32_t root = 0xffffffffu;
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
    if (d.father == 0xffffffffu) {
      if (root != 0xffffffffu) {
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
  if (root == 0xffffffffu) {
    set_error("h2matrix: no root");
    return H2HIP_EINVAL;
  }
  // levels (father chains; nodes may come in any order)
  for (uint32_t i = 0; i < nnodes; i++) {
    uint32_t l = 0, f = S.nodes[i].father;
    while (f != 0xffffffffu) {
      l++;
      f = S.nodes[f].father;
      if (l > nnodes) {
        set_error("h2matrix: cycle in the tree");
        return H2HIP_EINVAL;
      }
    }
    S.level[i] = l;
    S.depth = std::max(S.depth, l);
  }
  S.sonptr.assign(nnodes + 1, 0);
  for (uint32_t i = 0; i < nnodes; i++) S.sonptr[i + 1] = S.sonptr[i] + nsons[i];
  S.sons.assign(S.sonptr[nnodes], 0);
  std::vector<uint32_t> fill(nnodes, 0);
  for (uint32_t i = 0; i < nnodes; i++) {
    const uint32_t f = S.nodes[i].father;
    if (f != 0xffffffffu) S.sons[S.sonptr[f] + fill[f]++] = i;
  }
  S.inner_at.assign(S.depth + 1, {});
  S.all_at.assign(S.depth + 1, {});
  S.leaves.clear();
  uint64_t coef = 0, v = 0, maxr = 0;
  for (uint32_t i = 0; i < nnodes; i++) {
    NodeD &d = S.nodes[i];
    d.coef = coef;
    coef += d.rank;
    maxr = std::max<uint64_t>(maxr, d.rank);
    if (nsons[i] == 0) {
      d.voff = v;
      v += (uint64_t)d.size * d.rank;
      S.leaves.push_back(i);
    } else {
      d.voff = 0;
      S.inner_at[S.level[i]].push_back(i);
    }
    S.all_at[S.level[i]].push_back(i);
  }
  S.ncoef = coef;
  S.vlen = v;
  S.eld = maxr * maxr;
  std::vector<uint32_t> lists;
  S.inner_off.assign(S.depth + 2, 0);
  S.all_off.assign(S.depth + 2, 0);
  for (uint32_t l = 0; l <= S.depth; l++) {
    S.inner_off[l] = lists.size();
    lists.insert(lists.end(), S.inner_at[l].begin(), S.inner_at[l].end());
  }
  S.inner_off[S.depth + 1] = lists.size();
  for (uint32_t l = 0; l <= S.depth; l++) {
    S.all_off[l] = lists.size();
    lists.insert(lists.end(), S.all_at[l].begin(), S.all_at[l].end());
  }
  S.all_off[S.depth + 1] = lists.size();
  if (S.d_nodes.upload(S.nodes.data(), nnodes) != hipSuccess ||
      S.d_idx.upload(idx, ndof) != hipSuccess ||
      S.d_leaves.upload(S.leaves.data(), S.leaves.size()) != hipSuccess ||
      S.d_sonptr.upload(S.sonptr.data(), S.sonptr.size()) != hipSuccess ||
      S.d_sons.upload(S.sons.data(), S.sons.size()) != hipSuccess ||
      S.d_lists.upload(lists.data(), lists.size()) != hipSuccess) {
    set_error("h2matrix: device allocation failed");
    return H2HIP_ENOMEM;
  }
  return H2HIP_OK;
}

inline dim3 waves(uint32_t n) { return dim3((n + WPB - 1) / WPB); }

template <int C>
int apply(const double *alpha, const h2hip_h2matrix *H, const double *x,
          double *y, hipStream_t st, bool trans) {
  // forward on the side x lives in, backward on the other one
  const auto &in = H->side[trans ? 0 : 1];
  const auto &out = H->side[trans ? 1 : 0];
  double *xt = H->d_xt.p, *yt = H->d_yt.p;
  const uint64_t nyt = (C ? 2 : 1) * out.ncoef;
  if (nyt)
    hipLaunchKernelGGL(zero_kernel, dim3((nyt + 255) / 256), dim3(256), 0, st,
                       yt, nyt);
  // 1. forward, leaves
  if (!in.leaves.empty())
    hipLaunchKernelGGL(fwd_leaf<C>, waves(in.leaves.size()), dim3(NT), 0, st,
                       in.d_nodes.p, in.d_leaves.p,
                       (uint32_t)in.leaves.size(), in.d_idx.p, in.d_VT.p, x,
                       xt);
  // 2. forward, upwards
  for (int l = (int)in.depth; l >= 0; l--) {
    const uint32_t n = (uint32_t)in.inner_at[l].size();
    if (!n) continue;
    hipLaunchKernelGGL(up_level<C>, waves(n), dim3(NT), 0, st, in.d_nodes.p,
                       in.d_lists.p + in.inner_off[l], n, in.d_sonptr.p,
                       in.d_sons.p, in.d_ET.p, in.eld, xt);
  }
  // 3. couplings
  if (H->nfar) {
    if (trans)
      hipLaunchKernelGGL((coupling<C, true>), waves(H->nfar), dim3(NT), 0, st,
                         H->d_far.p, H->nfar, H->side[0].d_nodes.p,
                         H->side[1].d_nodes.p, H->d_S.p, xt, yt);
    else
      hipLaunchKernelGGL((coupling<C, false>), waves(H->nfar), dim3(NT), 0,
                         st, H->d_far.p, H->nfar, H->side[0].d_nodes.p,
                         H->side[1].d_nodes.p, H->d_S.p, xt, yt);
  }
  // 4. backward, downwards
  for (uint32_t l = 1; l <= out.depth; l++) {
    const uint32_t n = (uint32_t)out.all_at[l].size();
    if (!n) continue;
    hipLaunchKernelGGL(down_level<C>, waves(n), dim3(NT), 0, st,
                       out.d_nodes.p, out.d_lists.p + out.all_off[l], n,
                       out.d_E.p, out.eld, yt);
  }
  // 5. backward, leaves
  if (!out.leaves.empty())
    hipLaunchKernelGGL(back_leaf<C>, waves(out.leaves.size()), dim3(NT), 0,
                       st, out.d_nodes.p, out.d_leaves.p,
                       (uint32_t)out.leaves.size(), out.d_idx.p, out.d_V.p,
                       yt, alpha[0], alpha[1], y);
  // 6. near field
  if (H->nnear) {
    if (trans)
      hipLaunchKernelGGL((near_blocks<C, true>), waves(H->nnear), dim3(NT), 0,
                         st, H->d_near.p, H->nnear, H->side[0].d_idx.p,
                         H->side[1].d_idx.p, H->d_N.p, x, alpha[0], alpha[1],
                         y);
    else
      hipLaunchKernelGGL((near_blocks<C, false>), waves(H->nnear), dim3(NT),
                         0, st, H->d_near.p, H->nnear, H->side[0].d_idx.p,
                         H->side[1].d_idx.p, H->d_N.p, x, alpha[0], alpha[1],
                         y);
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
    if (S.sonptr[i + 1] == S.sonptr[i]) {
      const double *src = V + voff[i];
      for (uint32_t c = 0; c < d.rank; c++)
        for (uint32_t r = 0; r < d.size; r++) {
          v[d.voff + r + (uint64_t)c * d.size] = src[r + (uint64_t)c * d.size];
          vt[d.voff + c + (uint64_t)r * d.rank] =
              src[r + (uint64_t)c * d.size];
        }
    }
    if (d.father != 0xffffffffu) {
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
      if (d.father != 0xffffffffu)
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
