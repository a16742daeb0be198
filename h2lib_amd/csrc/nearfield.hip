// Near-field Galerkin quadrature on gfx950: planning, execution and the
// C ABI.  The kernels live in nf_device.h (instantiated per kernel family
// in nf_laplace.hip and nf_helmholtz.hip).
//
// Planning cuts every requested block into tiles of at most 64 x 64 DOFs.
// A tile lists the support triangles of its rows and of its columns (for
// the piecewise constant basis the triangles themselves, for the piecewise
// linear basis all triangles touching one of the tile's vertices, each with
// the tile-local numbers of its three vertices); its work is the set of all
// pairs (row support triangle, column support triangle), exactly the pairs
// the reference's loops visit for the same block (Library/bem3d.c:1150-
// 1247, 1798-1971, 2394-2571, 3038-3250).  Pairs with common vertices are
// found from the vertex-to-triangle map (no all-pairs scan) and collected
// into the singular list with their rule class and vertex permutation.
#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "h2hip_internal.h"

namespace h2hip {

thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

static std::atomic<int> g_reproducible{0};
bool reproducible() {
  return g_reproducible.load(std::memory_order_relaxed) != 0;
}
void set_reproducible(int on) {
  g_reproducible.store(on ? 1 : 0, std::memory_order_relaxed);
}

namespace {

// Identity term of the constant-basis double layer for pairs with equal
// triangles at different block positions: the reference adds it at (t, t)
// (Library/bem3d.c:1234-1241).
__global__ void nf_fixup_kernel(const Fixup *fix, uint32_t nfix,
                                const double *g, double are, double aim,
                                int cplx, int ntrans, double *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfix) return;
  const Fixup f = fix[i];
  const double v = 0.5 * g[f.tri];
  if (cplx) {
    atomicAdd(&out[2 * f.pos], v * are);
    atomicAdd(&out[2 * f.pos + 1], v * ((ntrans | f.pad) ? -aim : aim));
  } else {
    atomicAdd(&out[f.pos], v * are);
  }
}

// Quadrature points of the regular rule on every triangle.
// FP32 offsets of the regular points from point 0 of their triangle (mixed
// precision); point 0 itself stays in FP64 in the regular point array.
__global__ void regoff_kernel(const double *pts, uint32_t nt, int qq,
                              float *off) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt * (uint32_t)qq) return;
  const uint32_t tri = i / qq;
  const double *p0 = pts + 3 * (size_t)tri * qq;
#pragma unroll
  for (int d = 0; d < 3; d++)
    off[3 * (size_t)i + d] = (float)(pts[3 * (size_t)i + d] - p0[d]);
}

__global__ void regpts_kernel(const double *x, const uint32_t *t, uint32_t nt,
                              const double *x0, const double *x1, int qq,
                              double *pts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt * (uint32_t)qq) return;
  const uint32_t tri = i / qq, a = i % qq;
  const double *A = x + 3 * t[3 * tri];
  const double *B = x + 3 * t[3 * tri + 1];
  const double *C = x + 3 * t[3 * tri + 2];
  const double u = x0[a], v = x1[a];
  const double ca = 1.0 - u, cb = u - v, cc = v;
#pragma unroll
  for (int d = 0; d < 3; d++)
    pts[3 * (size_t)i + d] = A[d] * ca + B[d] * cb + C[d] * cc;
}

// 1/(4 pi) (Library/laplacebem3d.c:33)
constexpr double KERNEL_CONST = 0.0795774715459476679;

// ---- planning ----------------------------------------------------------------

// Support list of one tile side.  Constant basis: the DOFs are triangles.
// Linear basis: all triangles touching a DOF vertex, each with the
// tile-local numbers of its vertices.  Uses and restores pos[] (vertex ->
// tile-local number, or -1) and tpos[] (triangle -> support position).
void tile_side(const h2hip_surface *s, int basis, const uint32_t *idx,
               uint32_t cnt, std::vector<int32_t> &pos,
               std::vector<int32_t> &tpos, std::vector<Supp> &out) {
  auto make = [&](uint32_t tri, uint32_t loc) {
    Supp e;
    e.tri = tri;
    e.loc = loc;
    e.v[0] = s->t[3 * tri];
    e.v[1] = s->t[3 * tri + 1];
    e.v[2] = s->t[3 * tri + 2];
    e.pad = 0;
    return e;
  };
  if (!basis) {
    for (uint32_t i = 0; i < cnt; i++)
      out.push_back(make(idx[i], i | 0xffff00u));
    return;
  }
  for (uint32_t i = 0; i < cnt; i++) pos[idx[i]] = (int32_t)i;
  const size_t first = out.size();
  for (uint32_t i = 0; i < cnt; i++) {
    const uint32_t v = idx[i];
    for (uint32_t p = s->v2t_ptr[v]; p < s->v2t_ptr[v + 1]; p++) {
      const uint32_t tri = s->v2t[p];
      if (tpos[tri] >= 0) continue;
      tpos[tri] = (int32_t)(out.size() - first);
      uint32_t packed = 0;
      for (int l = 0; l < 3; l++) {
        const int32_t q = pos[s->t[3 * tri + l]];
        packed |= (uint32_t)(q >= 0 ? q : 0xff) << (8 * l);
      }
      out.push_back(make(tri, packed));
    }
  }
  for (size_t k = first; k < out.size(); k++) tpos[out[k].tri] = -1;
  for (uint32_t i = 0; i < cnt; i++) pos[idx[i]] = -1;
}

// Linear-basis tile side of a block whose indices repeat a vertex.  The
// reference records for each support triangle the first three block
// positions whose vertex it touches (Library/bem3d.c:3071-3116), so a
// position past the third gets nothing from that triangle.  For the tile's
// positions t0 .. t0 + cnt - 1 of the block's index array blk: each
// triangle once with the first allowed position of each of its vertices,
// then once more for every further allowed position of a repeated vertex,
// with only that slot set.  Uses and restores tpos[] (-1) and used[] (0).
void tile_side_repeated(const h2hip_surface *s, const uint32_t *blk,
                        uint32_t t0, uint32_t cnt, std::vector<int32_t> &tpos,
                        std::vector<int32_t> &used, std::vector<Supp> &out) {
  std::vector<uint32_t> tris;
  for (uint32_t i = t0; i < t0 + cnt; i++)
    for (uint32_t p = s->v2t_ptr[blk[i]]; p < s->v2t_ptr[blk[i] + 1]; p++) {
      const uint32_t tri = s->v2t[p];
      if (tpos[tri] >= 0) continue;
      tpos[tri] = (int32_t)tris.size();
      tris.push_back(tri);
    }
  // the block positions before the tile that these triangles record
  for (uint32_t i = 0; i < t0; i++)
    for (uint32_t p = s->v2t_ptr[blk[i]]; p < s->v2t_ptr[blk[i] + 1]; p++)
      if (tpos[s->v2t[p]] >= 0) used[s->v2t[p]]++;
  std::vector<uint32_t> slots(tris.size(), 0xffffffu);
  std::vector<std::pair<uint32_t, uint32_t>> extra;  // (triangle, slots)
  for (uint32_t i = t0; i < t0 + cnt; i++) {
    const uint32_t v = blk[i];
    for (uint32_t p = s->v2t_ptr[v]; p < s->v2t_ptr[v + 1]; p++) {
      const uint32_t tri = s->v2t[p];
      if (used[tri] >= 3) continue;
      used[tri]++;
      int l = 0;
      while (s->t[3 * tri + l] != v) l++;
      const uint32_t sh = 8 * l, loc = i - t0;
      uint32_t &m = slots[tpos[tri]];
      if (((m >> sh) & 0xffu) == 0xffu)
        m = (m & ~(0xffu << sh)) | (loc << sh);
      else
        extra.emplace_back(tri, (0xffffffu & ~(0xffu << sh)) | (loc << sh));
    }
  }
  auto push = [&](uint32_t tri, uint32_t loc) {
    Supp e;
    e.tri = tri;
    e.loc = loc;
    e.v[0] = s->t[3 * tri];
    e.v[1] = s->t[3 * tri + 1];
    e.v[2] = s->t[3 * tri + 2];
    e.pad = 0;
    out.push_back(e);
  };
  for (size_t k = 0; k < tris.size(); k++)
    if (slots[k] != 0xffffffu) push(tris[k], slots[k]);
  for (auto &e : extra) push(e.first, e.second);
  for (uint32_t tri : tris) {
    tpos[tri] = -1;
    used[tri] = 0;
  }
}

// Shared vertices of two triangles and the reference's vertex permutation:
// shared vertices first in the order of the first triangle, then the
// remaining ones in their original order (Library/singquad2d.c:549-633).
int classify(const uint32_t *tv, const uint32_t *sv, uint16_t &perm) {
  int tp[3], sp[3], p = 0;
  bool tu[3] = {false, false, false}, su[3] = {false, false, false};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (tv[i] == sv[j]) {
        tp[p] = i;
        sp[p] = j;
        tu[i] = su[j] = true;
        p++;
        break;
      }
  if (p == 3) {  // identical triangles keep their order
    for (int i = 0; i < 3; i++) tp[i] = sp[i] = i;
  } else {
    int ct = p, cs = p;
    for (int i = 0; i < 3; i++) {
      if (!tu[i]) tp[ct++] = i;
      if (!su[i]) sp[cs++] = i;
    }
  }
  perm = 0;
  for (int i = 0; i < 3; i++)
    perm |= (uint16_t)((tp[i] << (2 * i)) | (sp[i] << (6 + 2 * i)));
  return p;
}

// Runs f(t, lo, hi) on up to `threads` host threads over [0, n) cut into
// contiguous ranges in order (thread t gets the t-th range).
template <class F>
void parallel_ranges(uint32_t n, unsigned threads, F &&f) {
  threads = std::max(1u, std::min<unsigned>(threads, n ? n : 1));
  if (threads == 1) {
    f(0u, 0u, n);
    return;
  }
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; t++) {
    const uint32_t lo = (uint32_t)((uint64_t)n * t / threads);
    const uint32_t hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
    pool.emplace_back([&f, t, lo, hi]() { f(t, lo, hi); });
  }
  for (auto &th : pool) th.join();
}

// host threads of the planner: up to 16, one for small plans
unsigned plan_threads(uint64_t work) {
  if (work < 4096) return 1;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return std::min(16u, hw);
}

// ntr: per-block ntrans flags (nullptr: ntrans for every block); st: the
// uploads keep large enough device buffers of P and are ordered on st
// (otherwise new buffers, complete on return).
//
// The blocks and then the tiles are split between host threads in
// contiguous ranges, each with its own vertex / triangle scratch maps; the
// pieces are joined in range order, so the plan is the one a single thread
// builds.
int build_plan(h2hip_nfplan *P, const h2hip_surface *s, int rb, int cb,
               const uint32_t *ridx_all, uint32_t nridx,
               const uint32_t *cidx_all, uint32_t ncidx, uint32_t nblocks,
               const uint32_t *blocks, int ntrans,
               const uint8_t *ntr = nullptr, bool grow = false,
               hipStream_t st = nullptr) {
  P->off.clear();
  P->tiles.clear();
  P->supp.clear();
  P->fix.clear();
  P->spairs.clear();
  P->tpairs.clear();
  // a fused schedule is set again for the new plan
  P->fused = 0;
  P->d_tiles_f.release();
  P->d_tpairs.release();
  for (auto &c : P->pairs) c = 0;
  P->s = s;
  P->rb = rb;
  P->cb = cb;
  P->nblocks = nblocks;
  P->off.resize(nblocks);
  const uint32_t nrdof = rb ? s->nv : s->nt;
  const uint32_t ncdof = cb ? s->nv : s->nt;
  // tile edge (at most TILE, the size of the kernels' LDS tile)
  const uint32_t tile = std::min<uint32_t>(TILE, std::max<uint32_t>(8, P->tile));

  // block offsets in the packed output, and the checks of the blocks
  uint64_t total = 0, dofs = 0;
  for (uint32_t b = 0; b < nblocks; b++) {
    const uint32_t roff = blocks[4 * b], rows = blocks[4 * b + 1];
    const uint32_t coff = blocks[4 * b + 2], cols = blocks[4 * b + 3];
    if ((uint64_t)roff + rows > nridx || (uint64_t)coff + cols > ncidx) {
      set_error("block exceeds the index arrays");
      return H2HIP_EINVAL;
    }
    P->off[b] = total;
    total += (uint64_t)rows * cols;
    dofs += (uint64_t)rows + cols;
  }
  P->entries = total;

  // ---- tiles and support lists, per range of blocks ----
  struct Part {
    std::vector<Tile> tiles;
    std::vector<Supp> supp;
    std::vector<Fixup> fix;
    std::string err;
  };
  const unsigned nthr = plan_threads(dofs);
  std::vector<Part> parts(std::max(1u, std::min<unsigned>(nthr, nblocks)));
  parallel_ranges(nblocks, (unsigned)parts.size(),
                  [&](unsigned t, uint32_t lo, uint32_t hi) {
    Part &W = parts[t];
    std::vector<int32_t> pos(s->nv, -1), tpos(s->nt, -1);
    // A tile side's support list depends only on its index range, and a
    // cluster takes part in many near blocks: each range's list is built
    // once per thread and shared by the tiles that use it (key: offset in
    // the index array and count; rows and columns apart).
    using Memo = std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>>;
    Memo rmemo, cmemo;
    std::vector<int32_t> used;  // per triangle, for sides with repeats
    auto side = [&](Memo &memo, int basis, const uint32_t *all, uint32_t off,
                    uint32_t cnt, uint32_t &n) {
      const uint64_t key = (uint64_t)off << 32 | cnt;
      auto it = memo.find(key);
      if (it != memo.end()) {
        n = it->second.second;
        return it->second.first;
      }
      const uint32_t s0 = (uint32_t)W.supp.size();
      tile_side(s, basis, all + off, cnt, pos, tpos, W.supp);
      n = (uint32_t)W.supp.size() - s0;
      memo.emplace(key, std::make_pair(s0, n));
      return s0;
    };
    for (uint32_t b = lo; b < hi; b++) {
      const uint32_t roff = blocks[4 * b], rows = blocks[4 * b + 1];
      const uint32_t coff = blocks[4 * b + 2], cols = blocks[4 * b + 3];
      const uint32_t *ri = ridx_all + roff, *ci = cidx_all + coff;
      for (uint32_t i = 0; i < rows; i++)
        if (ri[i] >= nrdof) {
          W.err = "row index out of range";
          return;
        }
      for (uint32_t j = 0; j < cols; j++)
        if (ci[j] >= ncdof) {
          W.err = "column index out of range";
          return;
        }
      const uint64_t boff = P->off[b];
      const int bnt = ntr ? (ntr[b] ? 1 : 0) : ntrans;
      const uint32_t ld = bnt ? cols : rows;
      // a linear-basis side that repeats a vertex follows the reference's
      // rule of three positions per triangle, tile by tile
      auto repeats = [&](const uint32_t *ix, uint32_t n) {
        uint32_t k = 0;
        for (; k < n && pos[ix[k]] < 0; k++) pos[ix[k]] = 0;
        for (uint32_t j = 0; j < k; j++) pos[ix[j]] = -1;
        return k < n;
      };
      const bool rrep = rb && repeats(ri, rows);
      const bool crep = cb && repeats(ci, cols);
      if ((rrep || crep) && used.empty()) used.assign(s->nt, 0);
      auto side_rep = [&](const uint32_t *blk, uint32_t t0, uint32_t cnt,
                          uint32_t &n) {
        const uint32_t s0 = (uint32_t)W.supp.size();
        tile_side_repeated(s, blk, t0, cnt, tpos, used, W.supp);
        n = (uint32_t)W.supp.size() - s0;
        return s0;
      };
      for (uint32_t r0 = 0; r0 < rows; r0 += tile) {
        const uint32_t nr = std::min<uint32_t>(tile, rows - r0);
        uint32_t nrs;
        const uint32_t rs0 =
            rrep ? side_rep(ri, r0, nr, nrs)
                 : side(rmemo, rb, ridx_all, roff + r0, nr, nrs);
        for (uint32_t c0 = 0; c0 < cols; c0 += tile) {
          const uint32_t nc = std::min<uint32_t>(tile, cols - c0);
          uint32_t ncs;
          const uint32_t cs0 =
              crep ? side_rep(ci, c0, nc, ncs)
                   : side(cmemo, cb, cidx_all, coff + c0, nc, ncs);
          if (nrs > 0xffff || ncs > 0xffff ||
              (uint64_t)nrs * ncs >= (1ull << 31)) {
            W.err = "tile support too large";
            return;
          }
          Tile T;
          T.sbeg = 0;
          T.ns[0] = T.ns[1] = T.ns[2] = 0;
          T.out_off = boff + (bnt ? (uint64_t)c0 + (uint64_t)r0 * ld
                                  : (uint64_t)r0 + (uint64_t)c0 * ld);
          T.ld = ld;
          T.rsup = rs0;  // relative to the part's support list
          T.csup = cs0;
          T.nrsup = (uint16_t)nrs;
          T.ncsup = (uint16_t)ncs;
          T.row0 = r0;
          T.col0 = c0;
          T.nr = (uint8_t)nr;
          T.nc = (uint8_t)nc;
          T.ntrans = (uint8_t)bnt;
          // The reference's lc identity term is skipped for blocks with at
          // most three columns (Library/bem3d.c:2533 tests a loop variable
          // that equals 3 there).
          T.flags = (cols > 3) ? 1 : 0;
          W.tiles.push_back(T);
        }
      }
      // cc identity terms the reference leaves at (t, t) with t != s
      if (!rb && !cb) {
        // i == j: added with the entry itself.  The reference visits the
        // columns in order and overwrites (t, t) when it computes column
        // t, so a term added from column s < t does not survive; only
        // s > t leaves it in place (and keeps (t, t) inside the window).
        auto fixup = [&](uint32_t i, uint32_t j) {
          if (j <= i) return;
          Fixup f;
          f.pos = boff + (uint64_t)i + (uint64_t)i * ld;
          f.tri = ri[i];
          f.pad = (uint32_t)bnt;  // the block's ntrans
          W.fix.push_back(f);
        };
        // the column of each triangle through the thread's triangle map;
        // a block that repeats a column index takes a map of lists
        bool unique = true;
        uint32_t set = 0;
        for (; set < cols; set++) {
          if (tpos[ci[set]] >= 0) {
            unique = false;
            break;
          }
          tpos[ci[set]] = (int32_t)set;
        }
        if (unique) {
          for (uint32_t i = 0; i < rows; i++) {
            const int32_t j = tpos[ri[i]];
            if (j >= 0) fixup(i, (uint32_t)j);
          }
        }
        for (uint32_t j = 0; j < set; j++) tpos[ci[j]] = -1;
        if (!unique) {
          std::unordered_map<uint32_t, std::vector<uint32_t>> where;
          for (uint32_t j = 0; j < cols; j++) where[ci[j]].push_back(j);
          for (uint32_t i = 0; i < rows; i++) {
            auto it = where.find(ri[i]);
            if (it == where.end()) continue;
            for (uint32_t j : it->second) fixup(i, j);
          }
        }
      }
    }
  });
  std::vector<Tile> tiles;
  for (auto &W : parts) {
    if (!W.err.empty()) {
      set_error(W.err);
      return H2HIP_EINVAL;
    }
    const uint32_t base = (uint32_t)P->supp.size();
    for (auto &T : W.tiles) {
      T.rsup += base;
      T.csup += base;
    }
    tiles.insert(tiles.end(), W.tiles.begin(), W.tiles.end());
    P->supp.insert(P->supp.end(), W.supp.begin(), W.supp.end());
    P->fix.insert(P->fix.end(), W.fix.begin(), W.fix.end());
  }
  parts.clear();

  // largest tiles first: the tail of the launch runs the small ones
  std::vector<uint32_t> order(tiles.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return (uint64_t)tiles[x].nrsup * tiles[x].ncsup >
           (uint64_t)tiles[y].nrsup * tiles[y].ncsup;
  });
  P->tiles.resize(tiles.size());
  for (size_t i = 0; i < order.size(); i++) P->tiles[i] = tiles[order[i]];
  tiles.clear();

  // ---- pairs with common vertices, found through the vertex-to-triangle
  // map, per range of tiles ----
  struct SPart {
    std::vector<SPair> by_cls[4];
    std::vector<SPair> tp;  // grouped by tile
    uint64_t pairs0 = 0;
    std::string err;
  };
  const uint32_t ntiles = (uint32_t)P->tiles.size();
  std::vector<SPart> sparts(
      std::max(1u, std::min<unsigned>(plan_threads(P->supp.size()), ntiles)));
  parallel_ranges(ntiles, (unsigned)sparts.size(),
                  [&](unsigned t, uint32_t lo, uint32_t hi) {
    SPart &W = sparts[t];
    std::vector<int32_t> cpos(s->nt, -1), cnext;
    for (uint32_t ti = lo; ti < hi; ti++) {
      Tile &T = P->tiles[ti];
      W.pairs0 += (uint64_t)T.nrsup * T.ncsup;
      size_t b0[4];
      for (int c = 1; c < 4; c++) b0[c] = W.by_cls[c].size();
      // the first column position of each triangle; the further positions
      // of a triangle that the column indices repeat (constant basis)
      // chained behind it in increasing order
      cnext.assign(T.ncsup, -1);
      for (uint32_t c = T.ncsup; c-- > 0;) {
        const uint32_t tri = P->supp[T.csup + c].tri;
        cnext[c] = cpos[tri];
        cpos[tri] = (int32_t)c;
      }
      for (uint32_t r = 0; r < T.nrsup; r++) {
        const Supp &sr = P->supp[T.rsup + r];
        uint32_t seen[64];
        int nseen = 0;
        for (int l = 0; l < 3; l++) {
          const uint32_t v = sr.v[l];
          for (uint32_t p = s->v2t_ptr[v]; p < s->v2t_ptr[v + 1]; p++) {
            const uint32_t tri = s->v2t[p];
            if (cpos[tri] < 0) continue;
            bool dup = false;
            for (int k = 0; k < nseen; k++) dup |= (seen[k] == tri);
            if (dup) continue;
            if (nseen < 64) seen[nseen++] = tri;
            for (int32_t c = cpos[tri]; c >= 0; c = cnext[c]) {
              SPair q;
              q.tile = ti;
              q.r = (uint16_t)r;
              q.c = (uint16_t)c;
              q.cls =
                  (uint16_t)classify(sr.v, P->supp[T.csup + c].v, q.perm);
              W.by_cls[q.cls].push_back(q);
            }
          }
        }
      }
      for (uint32_t c = 0; c < T.ncsup; c++)
        cpos[P->supp[T.csup + c].tri] = -1;
      // the tile's pairs again, grouped by tile: identical, edge, vertex
      T.sbeg = (uint32_t)W.tp.size();  // relative to the part's list
      for (int c = 3; c >= 1; c--) {
        const size_t cnt = W.by_cls[c].size() - b0[c];
        if (cnt > 0xffff) {
          W.err = "too many pairs with common vertices in one tile";
          return;
        }
        T.ns[3 - c] = (uint16_t)cnt;
        W.tp.insert(W.tp.end(), W.by_cls[c].begin() + b0[c],
                    W.by_cls[c].end());
      }
    }
  });
  for (auto &W : sparts) {
    if (!W.err.empty()) {
      set_error(W.err);
      return H2HIP_EINVAL;
    }
  }
  // where each part's pairs go: the tile-grouped list in part order; the
  // class lists (identical, edge, vertex) one after the other, each in part
  // order; then the parts copy their pieces at the same time
  const unsigned np = (unsigned)sparts.size();
  std::vector<size_t> tdst(np), cdst(4 * (size_t)np);
  size_t ntp = 0, nsp = 0;
  for (unsigned part = 0; part < np; part++) {
    tdst[part] = ntp;
    ntp += sparts[part].tp.size();
    P->pairs[0] += sparts[part].pairs0;
  }
  for (int c = 3; c >= 1; c--) {
    for (unsigned part = 0; part < np; part++) {
      cdst[4 * part + c] = nsp;
      nsp += sparts[part].by_cls[c].size();
      P->pairs[c] += sparts[part].by_cls[c].size();
    }
    P->pairs[0] -= P->pairs[c];
  }
  if (ntp > 0xffffffffull) {
    set_error("too many pairs with common vertices");
    return H2HIP_EINVAL;
  }
  P->tpairs.resize(ntp);
  P->spairs.resize(nsp);
  parallel_ranges(np, np, [&](unsigned, uint32_t lo, uint32_t hi) {
    for (uint32_t part = lo; part < hi; part++) {
      SPart &W = sparts[part];
      const uint32_t t0 = (uint32_t)((uint64_t)ntiles * part / np);
      const uint32_t t1 = (uint32_t)((uint64_t)ntiles * (part + 1) / np);
      for (uint32_t ti = t0; ti < t1; ti++)
        P->tiles[ti].sbeg += (uint32_t)tdst[part];
      std::copy(W.tp.begin(), W.tp.end(), P->tpairs.begin() + tdst[part]);
      for (int c = 1; c < 4; c++)
        std::copy(W.by_cls[c].begin(), W.by_cls[c].end(),
                  P->spairs.begin() + cdst[4 * part + c]);
      W = SPart();
    }
  });
  sparts.clear();
  if (grow) {
    H2HIP_CHECK(P->d_tiles.upload_grow(P->tiles.data(), P->tiles.size(), st));
    H2HIP_CHECK(P->d_supp.upload_grow(P->supp.data(), P->supp.size(), st));
    H2HIP_CHECK(P->d_fix.upload_grow(P->fix.data(), P->fix.size(), st));
    H2HIP_CHECK(
        P->d_spairs.upload_grow(P->spairs.data(), P->spairs.size(), st));
    return H2HIP_OK;
  }
  H2HIP_CHECK(P->d_tiles.upload(P->tiles.data(), P->tiles.size()));
  H2HIP_CHECK(P->d_supp.upload(P->supp.data(), P->supp.size()));
  H2HIP_CHECK(P->d_fix.upload(P->fix.data(), P->fix.size()));
  H2HIP_CHECK(P->d_spairs.upload(P->spairs.data(), P->spairs.size()));
  return H2HIP_OK;
}

int execute(const h2hip_op *op, const h2hip_nfplan *P, int far, void *d_out,
            hipStream_t stream) {
  if (op->s != P->s || op->rb != P->rb || op->cb != P->cb) {
    set_error("plan and operator do not match (surface or bases)");
    return H2HIP_EINVAL;
  }
  if (P->tiles.empty()) return H2HIP_OK;
  const h2hip_surface *s = op->s;
  nf::KArgs a;
  a.tiles = P->d_tiles.p;
  a.supp = P->d_supp.p;
  a.x = s->d_x.p;
  a.n = s->d_n.p;
  a.g = s->d_g.p;
  a.regpts = op->d_regpts.p;
  a.regw = op->d_regw.p;
  a.gauss = op->d_gauss.p;
  a.regoff = op->d_regoff.p;
  a.regwf = op->d_regwf.p;
  a.q2 = op->q2;
  a.mp = op->precision;
  a.kc = KERNEL_CONST;
  a.kr = op->k[0];
  a.ki = -op->k[1];
  a.krf = (float)a.kr;
  a.krev = (float)(a.kr / (2.0 * M_PI));
  a.kil2 = (float)(a.ki * M_LOG2E);
  a.are = op->layer != H2HIP_SLP ? op->alpha[0] : 0.0;
  a.aim = op->layer != H2HIP_SLP ? op->alpha[1] : 0.0;
  a.adj = op->layer == H2HIP_ADLP;
  for (int i = 0; i < 9; i++) a.mass[i] = op->mass[i];
  a.far = far;
  a.out = (double *)d_out;
  a.counter = P->persistent ? P->d_counter.p : nullptr;
  a.tlist = nullptr;
  if (P->fused && !far && op->q == 2 && op->q2 == 4) {
    a.tiles = P->d_tiles_f.p;
    a.tlist = P->d_tpairs.p;
  }
  const uint32_t ntiles = (uint32_t)P->tiles.size();
  uint32_t ncls[4];
  for (int c = 0; c < 4; c++) ncls[c] = (uint32_t)P->pairs[c];
  const int rc =
      op->kernel == H2HIP_HELMHOLTZ
          ? nf::launch_helmholtz(op->layer, op->rb, op->cb, op->q, a, ntiles,
                                 P->d_spairs.p, ncls, stream)
          : nf::launch_laplace(op->layer, op->rb, op->cb, op->q, a, ntiles,
                               P->d_spairs.p, ncls, stream);
  if (rc != H2HIP_OK) {
    set_error("unsupported regular quadrature order (1..4 supported)");
    return rc;
  }
  H2HIP_CHECK(hipGetLastError());
  if (!far && !P->fix.empty() && (a.are != 0.0 || a.aim != 0.0)) {
    const uint32_t nfix = (uint32_t)P->fix.size();
    const int ntr = 0;  // each fix carries its block's ntrans
    hipLaunchKernelGGL(nf_fixup_kernel, dim3((nfix + 255) / 256), dim3(256),
                       0, stream, P->d_fix.p, nfix, s->d_g.p, a.are, a.aim,
                       op->kernel == H2HIP_HELMHOLTZ ? 1 : 0, ntr,
                       (double *)d_out);
    H2HIP_CHECK(hipGetLastError());
  }
  return H2HIP_OK;
}

}  // namespace

// One context of the blocking callback: its own stream, plan buffers,
// device output and pinned staging memory, all kept between calls.
struct CbCtx {
  hipStream_t st = nullptr;
  h2hip_nfplan plan;
  DevBuf<unsigned char> out;
  void *stage = nullptr;
  size_t stage_cap = 0;
  bool busy = false;
  ~CbCtx() {
    if (stage) (void)hipHostFree(stage);
    if (st) (void)hipStreamDestroy(st);
  }
  int stage_size(size_t bytes) {
    if (stage_cap >= bytes) return H2HIP_OK;
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_cap = 0;
    if (hipHostMalloc(&stage, bytes, hipHostMallocDefault) != hipSuccess) {
      stage = nullptr;
      set_error("h2hip_nearfield: pinned host allocation failed");
      return H2HIP_ENOMEM;
    }
    stage_cap = bytes;
    return H2HIP_OK;
  }
};

// Blocks recorded by h2hip_nearfield_defer: their index sets (appended to
// ridx / cidx), (roff, rows, coff, cols), ntrans and the host windows.
struct Deferred {
  std::vector<uint32_t> ridx, cidx, blocks;
  std::vector<uint8_t> ntr;
  std::vector<void *> dst;
  std::vector<size_t> ld;
  void clear() {
    ridx.clear();
    cidx.clear();
    blocks.clear();
    ntr.clear();
    dst.clear();
    ld.clear();
  }
};

namespace {

constexpr int CB_CONTEXTS = 4;

// Copies the packed column-major blocks of `stage` into their host
// windows (block b: ld[b], written in the block's storage orientation),
// split over a few threads for large copies.
void scatter_blocks(const unsigned char *stage, const uint64_t *off,
                    const uint32_t *blocks, const uint8_t *ntr,
                    void *const *dst, const size_t *ld, size_t nblocks,
                    size_t es) {
  auto one = [&](size_t b) {
    const uint32_t rows = blocks[4 * b + 1], cols = blocks[4 * b + 3];
    const size_t wr = ntr[b] ? cols : rows, wc = ntr[b] ? rows : cols;
    const unsigned char *src = stage + es * off[b];
    unsigned char *d = (unsigned char *)dst[b];
    if (ld[b] == wr) {
      std::memcpy(d, src, es * wr * wc);
    } else {
      for (size_t j = 0; j < wc; j++)
        std::memcpy(d + es * j * ld[b], src + es * j * wr, es * wr);
    }
  };
  uint64_t bytes = 0;
  for (size_t b = 0; b < nblocks; b++)
    bytes += (uint64_t)blocks[4 * b + 1] * blocks[4 * b + 3];
  bytes *= es;
  const unsigned hw = std::max(1u, std::min(16u,
                                            std::thread::hardware_concurrency()));
  const unsigned nt = bytes < (16u << 20) ? 1u : hw;
  if (nt <= 1 || nblocks < 2) {
    for (size_t b = 0; b < nblocks; b++) one(b);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++)
    pool.emplace_back([&]() {
      for (size_t b; (b = next.fetch_add(1)) < nblocks;) one(b);
    });
  for (auto &th : pool) th.join();
}

// Plans, computes and writes one group of blocks (one rule) through a
// callback context.
int run_blocks(const h2hip_op *op, CbCtx &C, int far,
               const uint32_t *ridx_all, uint32_t nridx,
               const uint32_t *cidx_all, uint32_t ncidx, uint32_t nblocks,
               const uint32_t *blocks, const uint8_t *ntr, void *const *dst,
               const size_t *ld) {
  if (!C.st && hipStreamCreateWithFlags(&C.st, hipStreamNonBlocking) !=
                   hipSuccess) {
    C.st = nullptr;
    set_error("h2hip_nearfield: stream creation failed");
    return H2HIP_EDEVICE;
  }
  int rc = build_plan(&C.plan, op->s, op->rb, op->cb, ridx_all, nridx,
                      cidx_all, ncidx, nblocks, blocks, 0, ntr, true, C.st);
  if (rc != H2HIP_OK) return rc;
  const size_t es = op->kernel == H2HIP_HELMHOLTZ ? 16 : 8;
  const size_t bytes = es * C.plan.entries;
  if (bytes == 0) return H2HIP_OK;
  if (C.out.n < bytes && C.out.alloc(bytes) != hipSuccess) {
    set_error("h2hip_nearfield: device allocation failed");
    return H2HIP_ENOMEM;
  }
  rc = C.stage_size(bytes);
  if (rc != H2HIP_OK) return rc;
  {
    std::lock_guard<std::mutex> guard(op->lock);  // precision stays fixed
    rc = execute(op, &C.plan, far, C.out.p, C.st);
  }
  if (rc != H2HIP_OK) return rc;
  H2HIP_CHECK(hipMemcpyAsync(C.stage, C.out.p, bytes, hipMemcpyDeviceToHost,
                             C.st));
  H2HIP_CHECK(hipStreamSynchronize(C.st));
  scatter_blocks((const unsigned char *)C.stage, C.plan.off.data(), blocks,
                 ntr, dst, ld, nblocks, es);
  return H2HIP_OK;
}

// A free callback context of op (waits while all are busy).
CbCtx *take_context(const h2hip_op *op) {
  std::unique_lock<std::mutex> lk(op->cb_lock);
  if (op->cb_ctx.empty())
    for (int i = 0; i < CB_CONTEXTS; i++) op->cb_ctx.push_back(new CbCtx);
  for (;;) {
    for (CbCtx *c : op->cb_ctx)
      if (!c->busy) {
        c->busy = true;
        return c;
      }
    op->cb_cv.wait(lk);
  }
}

void give_context(const h2hip_op *op, CbCtx *c) {
  {
    std::lock_guard<std::mutex> lk(op->cb_lock);
    c->busy = false;
  }
  op->cb_cv.notify_one();
}

}  // namespace
}  // namespace h2hip

h2hip_op::~h2hip_op() {
  for (h2hip::CbCtx *c : cb_ctx) delete c;
  delete[] defer;
}

using namespace h2hip;

// ---- C ABI ---------------------------------------------------------------------

extern "C" {

const char *h2hip_last_error(void) { return g_last_error.c_str(); }

int h2hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int h2hip_set_reproducible(int on) {
  set_reproducible(on);
  return H2HIP_OK;
}

int h2hip_get_reproducible(void) { return reproducible() ? 1 : 0; }

int h2hip_surface_create(h2hip_surface **out, uint32_t vertices,
                         const double *x, uint32_t triangles,
                         const uint32_t *t, const double *n,
                         const double *g) {
  if (!out || !x || !t || vertices == 0 || triangles == 0) {
    set_error("h2hip_surface_create: invalid arguments");
    return H2HIP_EINVAL;
  }
  for (size_t i = 0; i < 3 * (size_t)triangles; i++)
    if (t[i] >= vertices) {
      set_error("h2hip_surface_create: vertex index out of range");
      return H2HIP_EINVAL;
    }
  h2hip_surface *s = new h2hip_surface;
  H2HIP_CHECK(hipGetDevice(&s->device));
  s->nv = vertices;
  s->nt = triangles;
  s->x.assign(x, x + 3 * (size_t)vertices);
  s->t.assign(t, t + 3 * (size_t)triangles);
  s->n.resize(3 * (size_t)triangles);
  s->g.resize(triangles);
  if (n && g) {
    std::memcpy(s->n.data(), n, sizeof(double) * 3 * triangles);
    std::memcpy(s->g.data(), g, sizeof(double) * triangles);
  } else {
    // unit normals and twice the areas (Library/surface3d.c:51-108)
    for (uint32_t i = 0; i < triangles; i++) {
      const double *p0 = &s->x[3 * s->t[3 * i]];
      const double *p1 = &s->x[3 * s->t[3 * i + 1]];
      const double *p2 = &s->x[3 * s->t[3 * i + 2]];
      double e1[3], e2[3];
      for (int d = 0; d < 3; d++) {
        e1[d] = p1[d] - p0[d];
        e2[d] = p2[d] - p0[d];
      }
      double c[3] = {e1[1] * e2[2] - e1[2] * e2[1],
                     e1[2] * e2[0] - e1[0] * e2[2],
                     e1[0] * e2[1] - e1[1] * e2[0]};
      const double len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      s->g[i] = len;
      for (int d = 0; d < 3; d++) s->n[3 * i + d] = c[d] / len;
    }
  }
  // triangles of each vertex, in decreasing triangle order as the
  // reference's prepend-built lists
  s->v2t_ptr.assign(vertices + 1, 0);
  for (size_t i = 0; i < 3 * (size_t)triangles; i++) s->v2t_ptr[t[i] + 1]++;
  for (uint32_t v = 0; v < vertices; v++) s->v2t_ptr[v + 1] += s->v2t_ptr[v];
  s->v2t.resize(3 * (size_t)triangles);
  std::vector<uint32_t> fill(vertices, 0);
  for (uint32_t i = triangles; i-- > 0;)
    for (int l = 0; l < 3; l++) {
      const uint32_t v = t[3 * i + l];
      s->v2t[s->v2t_ptr[v] + fill[v]++] = i;
    }
  int rc = H2HIP_OK;
  if (s->d_x.upload(s->x.data(), s->x.size()) != hipSuccess ||
      s->d_t.upload(s->t.data(), s->t.size()) != hipSuccess ||
      s->d_n.upload(s->n.data(), s->n.size()) != hipSuccess ||
      s->d_g.upload(s->g.data(), s->g.size()) != hipSuccess) {
    set_error("h2hip_surface_create: device allocation failed");
    rc = H2HIP_ENOMEM;
  }
  if (rc != H2HIP_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return H2HIP_OK;
}

void h2hip_surface_destroy(h2hip_surface *s) { delete s; }

int h2hip_surface_geometry(const h2hip_surface *s, double *n, double *g) {
  if (!s) return H2HIP_EINVAL;
  if (n) std::memcpy(n, s->n.data(), sizeof(double) * s->n.size());
  if (g) std::memcpy(g, s->g.data(), sizeof(double) * s->g.size());
  return H2HIP_OK;
}

int h2hip_op_create(h2hip_op **out, const h2hip_surface *s, int kernel,
                    int layer, int row_basis, int col_basis,
                    uint32_t q_regular, uint32_t q_singular,
                    const double *alpha, const double *k) {
  if (!out || !s || (kernel != H2HIP_LAPLACE && kernel != H2HIP_HELMHOLTZ) ||
      (layer != H2HIP_SLP && layer != H2HIP_DLP && layer != H2HIP_ADLP) ||
      row_basis < 0 ||
      row_basis > 1 || col_basis < 0 || col_basis > 1 || q_regular < 1 ||
      q_regular > 4 || q_singular < 1 || q_singular > 16) {
    set_error("h2hip_op_create: invalid arguments");
    return H2HIP_EINVAL;
  }
  h2hip_op *op = new h2hip_op;
  op->s = s;
  op->kernel = kernel;
  op->layer = layer;
  op->rb = row_basis;
  op->cb = col_basis;
  op->q = q_regular;
  op->q2 = q_singular;
  if (alpha) {
    op->alpha[0] = alpha[0];
    op->alpha[1] = kernel == H2HIP_HELMHOLTZ ? alpha[1] : 0.0;
  }
  if (k && kernel == H2HIP_HELMHOLTZ) {
    op->k[0] = k[0];
    op->k[1] = k[1];
  }
  // mass matrices of the basis pairs (Library/laplacebem3d.c:928-994)
  if (row_basis && col_basis) {
    for (int i = 0; i < 9; i++)
      op->mass[i] = (i == 0 || i == 4 || i == 8) ? 1.0 / 12.0 : 1.0 / 24.0;
  } else if (row_basis || col_basis) {
    op->mass[0] = op->mass[1] = op->mass[2] = 1.0 / 6.0;
  }

  // Gauss nodes of the singular rules (the kernels form the Sauter-Schwab
  // points from them)
  std::vector<double> gq, gqw;
  gauss_legendre01(q_singular, gq, gqw);
  gq.insert(gq.end(), gqw.begin(), gqw.end());
  // regular rule: single-triangle points and (basis-weighted) weights
  TriRule tr = make_triangle_rule(q_regular);
  const int qq = (int)tr.w.size();
  op->regw.assign(4 * qq, 0.0);
  for (int i = 0; i < qq; i++) {
    const double u = tr.x0[i], v = tr.x1[i], w = tr.w[i];
    op->regw[i] = w;
    op->regw[qq + i] = w * (1.0 - u);
    op->regw[2 * qq + i] = w * (u - v);
    op->regw[3 * qq + i] = w * v;
  }
  DevBuf<double> d_x0, d_x1;
  int rc = H2HIP_OK;
  if (op->d_gauss.upload(gq.data(), gq.size()) != hipSuccess ||
      op->d_regw.upload(op->regw.data(), op->regw.size()) != hipSuccess ||
      d_x0.upload(tr.x0.data(), qq) != hipSuccess ||
      d_x1.upload(tr.x1.data(), qq) != hipSuccess ||
      op->d_regpts.alloc(3 * (size_t)qq * s->nt) != hipSuccess) {
    set_error("h2hip_op_create: device allocation failed");
    rc = H2HIP_ENOMEM;
  }
  if (rc == H2HIP_OK) {
    const uint32_t n = s->nt * (uint32_t)qq;
    hipLaunchKernelGGL(regpts_kernel, dim3((n + 255) / 256), dim3(256), 0, 0,
                       s->d_x.p, s->d_t.p, s->nt, d_x0.p, d_x1.p, qq,
                       op->d_regpts.p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      set_error("h2hip_op_create: point kernel failed");
      rc = H2HIP_EDEVICE;
    }
  }
  if (rc != H2HIP_OK) {
    delete op;
    return rc;
  }
  *out = op;
  return H2HIP_OK;
}

void h2hip_op_destroy(h2hip_op *op) { delete op; }

int h2hip_op_set_precision(h2hip_op *op, int mode) {
  if (!op || mode < 0 || mode > 2) {
    set_error("h2hip_op_set_precision: invalid arguments");
    return H2HIP_EINVAL;
  }
  std::lock_guard<std::mutex> guard(op->lock);
  if (mode >= 1 && !op->d_regoff.p) {
    const uint32_t qq = op->q * op->q;
    std::vector<float> wf(op->regw.begin(), op->regw.end());
    if (op->d_regwf.upload(wf.data(), wf.size()) != hipSuccess ||
        op->d_regoff.alloc(3 * (size_t)qq * op->s->nt) != hipSuccess) {
      op->d_regwf.release();
      op->d_regoff.release();
      set_error("h2hip_op_set_precision: device allocation failed");
      return H2HIP_ENOMEM;
    }
    const uint32_t n = op->s->nt * qq;
    hipLaunchKernelGGL(regoff_kernel, dim3((n + 255) / 256), dim3(256), 0, 0,
                       op->d_regpts.p, op->s->nt, qq, op->d_regoff.p);
    if (hipGetLastError() != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
      op->d_regoff.release();
      set_error("h2hip_op_set_precision: offset kernel failed");
      return H2HIP_EDEVICE;
    }
  }
  op->precision = mode;
  return H2HIP_OK;
}

int h2hip_nearfield(const h2hip_op *op, int far, const uint32_t *ridx,
                    uint32_t rows, const uint32_t *cidx, uint32_t cols,
                    int ntrans, void *N, size_t ld) {
  if (!op || (!N && rows && cols)) {
    set_error("h2hip_nearfield: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (rows == 0 || cols == 0) return H2HIP_OK;
  const size_t wr = ntrans ? cols : rows;
  if (ld < wr) {
    set_error("h2hip_nearfield: leading dimension too small");
    return H2HIP_EINVAL;
  }
  std::vector<uint32_t> ri, ci;
  if (!ridx) {
    ri.resize(rows);
    for (uint32_t i = 0; i < rows; i++) ri[i] = i;
    ridx = ri.data();
  }
  if (!cidx) {
    ci.resize(cols);
    for (uint32_t j = 0; j < cols; j++) ci[j] = j;
    cidx = ci.data();
  }
  const uint32_t blk[4] = {0, rows, 0, cols};
  const uint8_t ntr = ntrans ? 1 : 0;
  void *dst = N;
  CbCtx *C = take_context(op);
  const int rc = run_blocks(op, *C, far ? 1 : 0, ridx, rows, cidx, cols, 1,
                            blk, &ntr, &dst, &ld);
  give_context(op, C);
  return rc;
}

int h2hip_nearfield_defer(const h2hip_op *op, int far, const uint32_t *ridx,
                          uint32_t rows, const uint32_t *cidx, uint32_t cols,
                          int ntrans, void *N, size_t ld) {
  if (!op || (!N && rows && cols)) {
    set_error("h2hip_nearfield_defer: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (rows == 0 || cols == 0) return H2HIP_OK;
  const size_t wr = ntrans ? cols : rows;
  if (ld < wr) {
    set_error("h2hip_nearfield_defer: leading dimension too small");
    return H2HIP_EINVAL;
  }
  const uint32_t nr = op->rb ? op->s->nv : op->s->nt;
  const uint32_t nc = op->cb ? op->s->nv : op->s->nt;
  for (uint32_t i = 0; ridx && i < rows; i++)
    if (ridx[i] >= nr) {
      set_error("h2hip_nearfield_defer: row index out of range");
      return H2HIP_EINVAL;
    }
  for (uint32_t j = 0; cidx && j < cols; j++)
    if (cidx[j] >= nc) {
      set_error("h2hip_nearfield_defer: column index out of range");
      return H2HIP_EINVAL;
    }
  std::lock_guard<std::mutex> guard(op->defer_lock);
  if (!op->defer) op->defer = new Deferred[2];
  Deferred &D = op->defer[far ? 1 : 0];
  if (D.ridx.size() + rows > 0xffffffffull ||
      D.cidx.size() + cols > 0xffffffffull) {
    set_error("h2hip_nearfield_defer: too many deferred indices; flush "
              "first");
    return H2HIP_EINVAL;
  }
  D.blocks.push_back((uint32_t)D.ridx.size());
  D.blocks.push_back(rows);
  D.blocks.push_back((uint32_t)D.cidx.size());
  D.blocks.push_back(cols);
  for (uint32_t i = 0; i < rows; i++) D.ridx.push_back(ridx ? ridx[i] : i);
  for (uint32_t j = 0; j < cols; j++) D.cidx.push_back(cidx ? cidx[j] : j);
  D.ntr.push_back(ntrans ? 1 : 0);
  D.dst.push_back(N);
  D.ld.push_back(ld);
  return H2HIP_OK;
}

int h2hip_nearfield_flush(const h2hip_op *op) {
  if (!op) {
    set_error("h2hip_nearfield_flush: invalid arguments");
    return H2HIP_EINVAL;
  }
  Deferred D[2];
  {
    std::lock_guard<std::mutex> guard(op->defer_lock);
    if (!op->defer) return H2HIP_OK;
    std::swap(D[0], op->defer[0]);
    std::swap(D[1], op->defer[1]);
  }
  for (int far = 0; far < 2; far++) {
    const size_t nb = D[far].ntr.size();
    if (nb == 0) continue;
    CbCtx *C = take_context(op);
    const int rc = run_blocks(
        op, *C, far, D[far].ridx.data(), (uint32_t)D[far].ridx.size(),
        D[far].cidx.data(), (uint32_t)D[far].cidx.size(), (uint32_t)nb,
        D[far].blocks.data(), D[far].ntr.data(), D[far].dst.data(),
        D[far].ld.data());
    give_context(op, C);
    if (rc != H2HIP_OK) return rc;
  }
  return H2HIP_OK;
}

int h2hip_nfplan_create(h2hip_nfplan **out, const h2hip_surface *s,
                        int row_basis, int col_basis,
                        const uint32_t *ridx_all, uint32_t nridx,
                        const uint32_t *cidx_all, uint32_t ncidx,
                        uint32_t nblocks, const uint32_t *blocks) {
  if (!out || !s || !ridx_all || !cidx_all || (!blocks && nblocks)) {
    set_error("h2hip_nfplan_create: invalid arguments");
    return H2HIP_EINVAL;
  }
  h2hip_nfplan *P = new h2hip_nfplan;
  int rc = build_plan(P, s, row_basis ? 1 : 0, col_basis ? 1 : 0, ridx_all,
                      nridx, cidx_all, ncidx, nblocks, blocks, 0);
  if (rc != H2HIP_OK) {
    delete P;
    return rc;
  }
  // the inputs, for a new tiling (h2hip_nfplan_set_tile)
  P->in_ridx.assign(ridx_all, ridx_all + nridx);
  P->in_cidx.assign(cidx_all, cidx_all + ncidx);
  P->in_blocks.assign(blocks, blocks + 4 * (size_t)nblocks);
  *out = P;
  return H2HIP_OK;
}

int h2hip_nfplan_set_tile(h2hip_nfplan *p, uint32_t tile) {
  if (!p || tile < 8 || tile > (uint32_t)TILE) {
    set_error("h2hip_nfplan_set_tile: invalid arguments (tile 8..64)");
    return H2HIP_EINVAL;
  }
  if (tile == p->tile) return H2HIP_OK;
  const int sorted = p->sorted, persistent = p->persistent,
            fused = p->fused;
  const uint32_t old = p->tile;
  p->tile = tile;
  int rc = build_plan(p, p->s, p->rb, p->cb, p->in_ridx.data(),
                      (uint32_t)p->in_ridx.size(), p->in_cidx.data(),
                      (uint32_t)p->in_cidx.size(), p->nblocks,
                      p->in_blocks.data(), 0);
  if (rc != H2HIP_OK) {
    // the previous tiling again (it was valid)
    p->tile = old;
    (void)build_plan(p, p->s, p->rb, p->cb, p->in_ridx.data(),
                     (uint32_t)p->in_ridx.size(), p->in_cidx.data(),
                     (uint32_t)p->in_cidx.size(), p->nblocks,
                     p->in_blocks.data(), 0);
  }
  // the new tiles are in plan order: the schedule is applied again
  p->sorted = 0;
  p->tiles0.clear();
  p->spairs0.clear();
  int rc2 = h2hip_nfplan_set_schedule(p, sorted, persistent);
  if (rc2 == H2HIP_OK) rc2 = h2hip_nfplan_set_fused(p, fused);
  return rc != H2HIP_OK ? rc : rc2;
}

void h2hip_nfplan_destroy(h2hip_nfplan *p) { delete p; }

uint64_t h2hip_nfplan_entries(const h2hip_nfplan *p) {
  return p ? p->entries : 0;
}

int h2hip_nfplan_offsets(const h2hip_nfplan *p, uint64_t *off) {
  if (!p || !off) return H2HIP_EINVAL;
  std::memcpy(off, p->off.data(), sizeof(uint64_t) * p->off.size());
  return H2HIP_OK;
}

int h2hip_nfplan_set_schedule(h2hip_nfplan *p, int sorted, int persistent) {
  if (!p || sorted < 0 || sorted > 1 || persistent < 0 || persistent > 1) {
    set_error("h2hip_nfplan_set_schedule: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (persistent && !p->d_counter.p)
    H2HIP_CHECK(p->d_counter.alloc(1));
  p->persistent = persistent;
  if (sorted == p->sorted) return H2HIP_OK;
  if (sorted) {
    // tiles by decreasing number of support pairs (the regular work), the
    // singular pairs renumbered to follow their tiles
    p->tiles0 = p->tiles;
    p->spairs0 = p->spairs;
    const size_t nt = p->tiles.size();
    std::vector<uint32_t> perm(nt), pos(nt);
    for (size_t i = 0; i < nt; i++) perm[i] = (uint32_t)i;
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) {
      const uint64_t cx = (uint64_t)p->tiles0[x].nrsup * p->tiles0[x].ncsup;
      const uint64_t cy = (uint64_t)p->tiles0[y].nrsup * p->tiles0[y].ncsup;
      return cx > cy;
    });
    for (size_t i = 0; i < nt; i++) {
      p->tiles[i] = p->tiles0[perm[i]];
      pos[perm[i]] = (uint32_t)i;
    }
    for (auto &sp : p->spairs) sp.tile = pos[sp.tile];
  } else {
    p->tiles = p->tiles0;
    p->spairs = p->spairs0;
  }
  H2HIP_CHECK(p->d_tiles.upload(p->tiles.data(), p->tiles.size()));
  H2HIP_CHECK(p->d_spairs.upload(p->spairs.data(), p->spairs.size()));
  p->sorted = sorted;
  return H2HIP_OK;
}

int h2hip_nfplan_set_fused(h2hip_nfplan *p, int fused) {
  if (!p || fused < 0 || fused > 1) {
    set_error("h2hip_nfplan_set_fused: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (fused && !p->d_tiles_f.p && !p->tiles.empty()) {
    // tiles by decreasing work at regular order 2 and singular order 4
    // (16 evaluations per regular pair, 1536 per identical or edge pair,
    // 512 per vertex pair)
    std::vector<Tile> tf = p->tiles;
    auto work = [](const Tile &T) {
      return (uint64_t)T.nrsup * T.ncsup * 16 +
             ((uint64_t)T.ns[0] + T.ns[1]) * 1536 + (uint64_t)T.ns[2] * 512;
    };
    std::stable_sort(tf.begin(), tf.end(), [&](const Tile &x, const Tile &y) {
      return work(x) > work(y);
    });
    H2HIP_CHECK(p->d_tiles_f.upload(tf.data(), tf.size()));
    H2HIP_CHECK(p->d_tpairs.upload(p->tpairs.data(), p->tpairs.size()));
  }
  p->fused = fused;
  return H2HIP_OK;
}

int h2hip_nfplan_work(const h2hip_nfplan *p, const h2hip_op *op, int far,
                      double *out) {
  if (!p || !op || !out) return H2HIP_EINVAL;
  double pairs[4];
  for (int c = 0; c < 4; c++) pairs[c] = (double)p->pairs[c];
  if (far) {
    pairs[0] += pairs[1] + pairs[2] + pairs[3];
    pairs[1] = pairs[2] = pairs[3] = 0.0;
  }
  const double q4 = (double)op->q * op->q * op->q * op->q;
  const double s4 = (double)op->q2 * op->q2 * op->q2 * op->q2;
  const double nq[4] = {q4, 2.0 * s4, 6.0 * s4, 6.0 * s4};
  for (int c = 0; c < 4; c++) {
    out[c] = pairs[c] * nq[c];
    out[4 + c] = pairs[c];
  }
  return H2HIP_OK;
}

int h2hip_nearfield_execute(const h2hip_op *op, const h2hip_nfplan *p,
                            int far, void *d_out, void *stream) {
  if (!op || !p || (!d_out && p->entries)) {
    set_error("h2hip_nearfield_execute: invalid arguments");
    return H2HIP_EINVAL;
  }
  return execute(op, p, far, d_out, (hipStream_t)stream);
}

int h2hip_stream_sync(void *stream) {
  H2HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return H2HIP_OK;
}

int h2hip_malloc(void **ptr, size_t bytes) {
  if (!ptr) return H2HIP_EINVAL;
  *ptr = nullptr;
  if (bytes == 0) return H2HIP_OK;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    set_error("h2hip_malloc: out of device memory");
    return H2HIP_ENOMEM;
  }
  return H2HIP_OK;
}

int h2hip_free(void *ptr) {
  if (ptr) H2HIP_CHECK(hipFree(ptr));
  return H2HIP_OK;
}

// The blocking copies and fills complete before they return: work that the
// caller launches afterwards on a non-blocking stream (which is not ordered
// after the default stream) sees their result.
int h2hip_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  if (bytes == 0) return H2HIP_OK;
  const hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  H2HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, nullptr));
  H2HIP_CHECK(hipStreamSynchronize(nullptr));
  return H2HIP_OK;
}

int h2hip_memset(void *ptr, int value, size_t bytes) {
  if (bytes == 0) return H2HIP_OK;
  H2HIP_CHECK(hipMemsetAsync(ptr, value, bytes, nullptr));
  H2HIP_CHECK(hipStreamSynchronize(nullptr));
  return H2HIP_OK;
}

int h2hip_memset_async(void *ptr, int value, size_t bytes, void *stream) {
  if (bytes == 0) return H2HIP_OK;
  H2HIP_CHECK(hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream));
  return H2HIP_OK;
}

int h2hip_stream_create(void **stream) {
  if (!stream) return H2HIP_EINVAL;
  hipStream_t s;
  H2HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return H2HIP_OK;
}

int h2hip_stream_destroy(void *stream) {
  if (stream) H2HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
  return H2HIP_OK;
}

int h2hip_event_create(void **event) {
  if (!event) return H2HIP_EINVAL;
  hipEvent_t e;
  H2HIP_CHECK(hipEventCreate(&e));
  *event = (void *)e;
  return H2HIP_OK;
}

int h2hip_event_destroy(void *event) {
  if (event) H2HIP_CHECK(hipEventDestroy((hipEvent_t)event));
  return H2HIP_OK;
}

int h2hip_event_record(void *event, void *stream) {
  H2HIP_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return H2HIP_OK;
}

int h2hip_event_elapsed(void *start, void *stop, float *ms) {
  if (!ms) return H2HIP_EINVAL;
  H2HIP_CHECK(hipEventSynchronize((hipEvent_t)stop));
  H2HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return H2HIP_OK;
}

}  // extern "C"
