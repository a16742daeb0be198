// Internal declarations of the h2hip library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "h2hip.h"

namespace h2hip {

// ---- errors -----------------------------------------------------------------

void set_error(const std::string &msg);

// Reproducible mode (h2hip_set_reproducible): sums in a fixed order instead
// of atomics in the H2 products and the Krylov reductions.
bool reproducible();
void set_reproducible(int on);

#define H2HIP_CHECK(call)                                                    \
  do {                                                                       \
    hipError_t err__ = (call);                                               \
    if (err__ != hipSuccess) {                                               \
      ::h2hip::set_error(std::string(#call) + ": " +                        \
                         hipGetErrorString(err__));                          \
      return H2HIP_EDEVICE;                                                  \
    }                                                                        \
  } while (0)

// Device buffer owned by the library.
template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)&p, sizeof(T) * count);
  }
  hipError_t upload(const T *src, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice);
  }
  // as upload, but keeps an allocation that is large enough (n is then the
  // capacity), ordered on a stream
  hipError_t upload_grow(const T *src, size_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    if (!p || n < count) {
      hipError_t e = alloc(count);
      if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice,
                          st);
  }
};

// ---- quadrature rules (host) -------------------------------------------------

// Single-triangle rule of order q: points (x0, x1), weights w.
struct TriRule {
  std::vector<double> x0, x1, w;
};

void gauss_legendre01(unsigned m, std::vector<double> &x,
                      std::vector<double> &w);
TriRule make_triangle_rule(unsigned q);

// ---- device-side records -------------------------------------------------------

// Rule classes by number of shared vertices.
enum { CLS_REGULAR = 0, CLS_VERTEX = 1, CLS_EDGE = 2, CLS_IDENTICAL = 3 };

// Support entry of a tile side: triangle, its vertices, and the tile-local
// DOF numbers of its three local basis functions (0xff: not in the tile).
// For the constant basis loc0 is the tile-local triangle number.
struct Supp {
  uint32_t tri;
  uint32_t loc;  // loc0 | loc1 << 8 | loc2 << 16
  uint32_t v[3];
  uint32_t pad;
};

// A pair of support triangles with common vertices, evaluated by the
// singular kernel: tile, row / column support positions, rule class
// (1 vertex, 2 edge, 3 identical) and the vertex permutations that put the
// shared vertices first (tp | sp << 2 | ... two bits per entry).
struct SPair {
  uint32_t tile;
  uint16_t r, c;
  uint16_t cls;
  uint16_t perm;  // tp0 | tp1 << 2 | tp2 << 4 | sp0 << 6 | sp1 << 8 | sp2 << 10
};

constexpr int TILE = 64;

struct Tile {
  uint64_t out_off;  // element offset of tile entry (0, 0)
  uint32_t ld;       // leading dimension of the output block
  uint32_t rsup;     // first row support entry
  uint32_t csup;     // first column support entry
  uint32_t row0, col0;  // tile position inside the block (multiples of 64)
  uint16_t nrsup, ncsup;
  uint8_t nr, nc;       // tile size
  uint8_t ntrans;
  uint8_t flags;        // bit 0: identity term allowed (lc quirk)
  // the tile's pairs with common vertices in the list grouped by tile:
  // from sbeg, ns[0] identical, then ns[1] edge, then ns[2] vertex pairs
  uint32_t sbeg;
  uint16_t ns[3];
};

struct Fixup {  // identity term of the cc double layer at (t, t), t != s
  uint64_t pos;
  uint32_t tri;
  uint32_t pad;
};

namespace nf {

// Arguments of the near-field kernels.
struct KArgs {
  const Tile *tiles;
  const Supp *supp;
  const double *x;       // vertices
  const double *n;       // unit normals
  const double *g;       // Gram determinants
  const double *regpts;  // regular-rule points, 3 q^2 per triangle
  const double *regw;    // regular weights: w | phi0 w | phi1 w | phi2 w
  const double *gauss;   // Gauss-Legendre nodes and weights on [0, 1], q2 each
  // mixed precision only: offsets of the regular points from point 0 of
  // their triangle and the regular weights, in FP32
  const float *regoff;
  const float *regwf;
  uint32_t q2;           // order of the singular rules
  int mp;                // 1: kernel values in FP32 (mixed precision)
  double kc;             // 1 / (4 pi)
  double kr, ki;         // e^{i k r} with k = kr - i ki
  float krf, krev, kil2;  // FP32: kr, kr / (2 pi), ki log2(e)
  double are, aim;       // identity coefficient alpha
  double mass[9];
  int far;
  // adjoint double layer: the kernels of the double layer take -n_x (the
  // row triangle's normal) in place of n_y, <y - x, n_x> = <x - y, -n_x>
  int adj;
  double *out;
  // persistent schedule of the regular kernel: tile counter (nullptr: one
  // workgroup per tile)
  uint32_t *counter;
  // fused tiles (regular order 2, singular order 4): the singular pairs
  // grouped by tile, evaluated by the tile's workgroup (nullptr: separate
  // launches per rule class over the class-sorted list)
  const SPair *tlist;
};

// Launches the regular tile kernel over `ntiles` tiles and then (unless
// far) the singular kernel once per rule class over the pairs sp (sorted
// identical, edge, vertex; ncls[c] pairs of class c), on `stream`.
// Returns H2HIP_EINVAL if the order q has no kernel.
int launch_laplace(int layer, int rb, int cb, uint32_t q, const KArgs &a,
                   uint32_t ntiles, const SPair *sp, const uint32_t *ncls,
                   hipStream_t stream);
int launch_helmholtz(int layer, int rb, int cb, uint32_t q, const KArgs &a,
                     uint32_t ntiles, const SPair *sp, const uint32_t *ncls,
                     hipStream_t stream);

}  // namespace nf
}  // namespace h2hip

struct h2hip_surface {
  int device = 0;
  uint32_t nv = 0, nt = 0;
  std::vector<double> x, n, g;
  std::vector<uint32_t> t;
  std::vector<uint32_t> v2t_ptr, v2t;  // triangles of each vertex
  h2hip::DevBuf<double> d_x, d_n, d_g;
  h2hip::DevBuf<uint32_t> d_t;
  // data of the L2 projection onto the linear basis, built on first use:
  // vertex -> (triangle, local vertex) lists, the mass matrix in CSR form
  // and the workspace of the projection kernel
  mutable std::mutex fl_lock;
  mutable bool fl_ready = false;
  mutable h2hip::DevBuf<uint32_t> d_vt_ptr, d_vt, d_mrow, d_mcol, d_pflag;
  mutable h2hip::DevBuf<double> d_mval, d_fwork;
  // the projection's looks at its stopping flag: two pinned slots, events
  mutable std::mutex proj_lock;
  mutable uint32_t *h_plook = nullptr;
  mutable hipEvent_t pev[2] = {nullptr, nullptr};
  ~h2hip_surface() {
    if (h_plook) (void)hipHostFree(h_plook);
    for (auto e : pev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace h2hip {
// Builds (once) the surface's vertex -> (4 triangle + local vertex) lists
// and the linear-basis mass matrix on the device (d_vt_ptr, d_vt, ...).
int surface_vertex_lists(const h2hip_surface *s);
}  // namespace h2hip

namespace h2hip {
struct CbCtx;
struct Deferred;
}  // namespace h2hip

struct h2hip_op {
  const h2hip_surface *s = nullptr;
  int kernel = 0, layer = 0, rb = 0, cb = 0;
  uint32_t q = 2, q2 = 4;
  double alpha[2] = {0, 0}, k[2] = {0, 0};
  double mass[9] = {0};
  int precision = 0;
  // Gauss-Legendre nodes and weights of order q2 on [0, 1]: [g | w]; the
  // singular kernels form the Sauter-Schwab points from them
  h2hip::DevBuf<double> d_gauss;
  // regular rule: per-triangle points (3 * q * q doubles per triangle)
  h2hip::DevBuf<double> d_regpts;
  // mixed precision: regular points relative to point 0 of their triangle
  // and the regular weights, in FP32
  h2hip::DevBuf<float> d_regoff, d_regwf;
  // regular weights: [w (q*q) | phi_0 (q*q) | phi_1 | phi_2], basis factors
  // premultiplied by the weights
  h2hip::DevBuf<double> d_regw;
  std::vector<double> regw;
  mutable std::mutex lock;  // guards set_precision against the callbacks
  // contexts of the blocking callback h2hip_nearfield (created on first
  // use): calls from several host threads each take a free one and run
  // concurrently on its stream
  mutable std::mutex cb_lock;
  mutable std::condition_variable cb_cv;
  mutable std::vector<h2hip::CbCtx *> cb_ctx;
  // blocks recorded by h2hip_nearfield_defer (0 near, 1 far rule) until
  // h2hip_nearfield_flush
  mutable std::mutex defer_lock;
  mutable h2hip::Deferred *defer = nullptr;
  ~h2hip_op();
};

struct h2hip_nfplan {
  const h2hip_surface *s = nullptr;
  int rb = 0, cb = 0;
  uint32_t nblocks = 0;
  uint64_t entries = 0;
  std::vector<uint64_t> off;
  std::vector<h2hip::Tile> tiles;
  std::vector<h2hip::Supp> supp;
  std::vector<h2hip::Fixup> fix;
  std::vector<h2hip::SPair> spairs;  // sorted by class, identical first
  std::vector<h2hip::SPair> tpairs;  // grouped by tile (Tile::sbeg, ns)
  uint64_t pairs[4] = {0, 0, 0, 0};  // support pairs per class
  h2hip::DevBuf<h2hip::Tile> d_tiles;
  h2hip::DevBuf<h2hip::Supp> d_supp;
  h2hip::DevBuf<h2hip::Fixup> d_fix;
  h2hip::DevBuf<h2hip::SPair> d_spairs;
  // schedule: tiles in decreasing order of their pair count, and the
  // persistent regular kernel with its tile counter
  int sorted = 0, persistent = 0;
  std::vector<h2hip::Tile> tiles0;    // plan order, kept once sorted
  std::vector<h2hip::SPair> spairs0;
  h2hip::DevBuf<uint32_t> d_counter;
  // fused schedule: tiles by decreasing total work (regular and singular
  // evaluations) and the singular pairs grouped by tile
  int fused = 0;
  // tile edge and the inputs of h2hip_nfplan_create (for a new tiling)
  uint32_t tile = h2hip::TILE;
  std::vector<uint32_t> in_ridx, in_cidx, in_blocks;
  h2hip::DevBuf<h2hip::Tile> d_tiles_f;
  h2hip::DevBuf<h2hip::SPair> d_tpairs;
};
