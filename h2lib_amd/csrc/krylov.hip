// Restarted GMRES on device vectors, with the Arnoldi process, the Givens
// rotations and the restarts on the device.
//
// Semantics of the reference's solve_gmres_avector
// (Library/krylovsolvers.c:183-227) with init/step/finish_gmres
// (Library/krylov.c:612-770): the Krylov space is restarted when its
// dimension k reaches kmax - 1, the iteration stops when the least-squares
// residual |rhat_k| <= eps |b| or after maxiter - 1 steps, and the solution
// is updated at every restart and at the end.  The reference orthogonalises
// with Householder reflections; here the Arnoldi vectors are
// orthogonalised by classical Gram-Schmidt with one reorthogonalisation,
// which spans the same Krylov spaces and agrees to rounding.  All inner
// products are FP64.
//
// No step waits for the host.  The host enqueues up to CHUNK steps (never
// past a restart or maxiter), then reads the residual estimates those steps
// left in a small device array, and stops at the first step that meets the
// reference's criterion.  Steps enqueued after that one are discarded: they
// only touch Krylov columns beyond k and the rotation of the next column,
// never the first k columns of R or g, so the update x += Q_k y is exactly
// the one of the stopped iteration.
//
// Per step: the product w = A q_k (caller's addeval), two Gram-Schmidt
// passes (multi-dot with the last block reducing the partial sums, then
// w -= Q h; the second update also forms |w|^2 the same way), and one
// kernel that scales q_{k+1} = w / |w| while its first thread applies the
// Givens rotations.
#include <algorithm>
#include <cmath>
#include <deque>
#include <vector>

#include "h2hip_internal.h"
#include "wave_sum.h"

namespace {

constexpr int NT = 256;
constexpr int NB = 240;    // blocks of the multi-dot kernel
constexpr int MAXK = 64;   // kmax limit
constexpr int CHUNK = 8;   // steps between two looks at the residual

// Thread t's share of a sum over blocks of the partial sums stored at
// part[(b * 64 + j) * 2 + {0, 1}]: value v = t % nv of the blocks
// b = t / nv, t / nv + groups, ... below nblk, added in that order.  The
// loads are issued eight at a time before their additions (one memory
// round trip per eight blocks instead of one per block).
__device__ __forceinline__ double strided_part_sum(const double *part,
                                                   uint32_t nblk, int nv,
                                                   int groups, int t) {
  const int v = t % nv;
  double s = 0.0;
  for (uint32_t b0 = t / nv; b0 < nblk; b0 += 8 * groups) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t b = b0 + u * groups;
      x[u] = b < nblk ? part[((uint64_t)b * 64 + (v >> 1)) * 2 + (v & 1)]
                      : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += x[u];
  }
  return s;
}

// Sums, over the blocks of a grid, 2 m doubles (m <= 64) per block stored
// at part[(b * 64 + j) * 2 + {0, 1}]; run by the last block of the grid to
// finish (cnt counts the finished blocks and is reset to 0 here).  The
// agent-scope fences around the counter make the other blocks' partial sums
// visible to the last block.  Its threads split the blocks between
// NT / (2 m) groups per value and combine the groups in LDS.
__device__ void last_block_sum(const double *part, int m, uint32_t *cnt,
                               double *out) {
  __shared__ bool last;
  __shared__ double red[NT];
  // the barrier completes every wave's stores; one agent-scope fence then
  // publishes them before the counter increment (and, in the last block,
  // makes the other blocks' sums visible before the barrier releases the
  // loads below)
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    if (last) __threadfence();
  }
  __syncthreads();
  if (!last) return;
  const int nv = 2 * m, groups = NT / nv, t = threadIdx.x;
  double s = 0.0;
  if (t < groups * nv) s = strided_part_sum(part, gridDim.x, nv, groups, t);
  red[t] = s;
  __syncthreads();
  if (t < nv) {
    double r = 0.0;
    for (int g = 0; g < groups; g++) r += red[g * nv + t];
    out[t] = r;
  }
  if (t == 0) *cnt = 0;
}

using h2hip::wave_sum32;

// Ways of combining the blocks' sums of the vector kernels: the last block
// of the grid adds them in block order (fixed order), every block adds
// into the output with FP64 atomics, or every block leaves its sums in
// `part` for the next kernel (sum_parts), with no atomics and no fences.
enum { RED_LAST = 0, RED_ATOM = 1, RED_PART = 2 };

// The block's sums of the threads' accumulators of columns j0 .. j0 + mj
// (mj <= 32; imaginary parts when C): a butterfly over the wave's lanes,
// then the waves' sums added in LDS; RED_ATOM adds them into out
// (2 * j0 on), otherwise they go to part[(b * 64 + j) * 2 + {0, 1}].
template <int C, int ATOM>
__device__ __forceinline__ void block_sums32(double (&acc_re)[32],
                                             double (&acc_im)[32], int j0,
                                             int mj, double *part,
                                             double *out) {
  __shared__ double red[NT / 64][2 * 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // lane l ends with the sum of accumulator l & 31
  const double r = wave_sum32(acc_re, lane);
  const double im = C ? wave_sum32(acc_im, lane) : 0.0;
  if (lane < mj) {
    red[wv][2 * lane] = r;
    red[wv][2 * lane + 1] = im;
  }
  __syncthreads();
  if (threadIdx.x < 2 * mj) {
    double s = 0.0;
    for (int v = 0; v < NT / 64; v++) s += red[v][threadIdx.x];
    if constexpr (ATOM == RED_ATOM)
      atomicAdd(&out[2 * j0 + threadIdx.x], s);
    else
      part[((uint64_t)blockIdx.x * 64 + (j0 + threadIdx.x / 2)) * 2 +
           (threadIdx.x & 1)] = s;
  }
  __syncthreads();
}

// out[0 .. 2m) = Q[:, 0..m)^* w (complex pairs; imaginary parts 0 when real)
// ATOM (RED_ATOM): every block adds its sums into out (zeroed beforehand)
// with FP64 atomics instead of the last-block reduction, whose agent-scope
// fences write back and invalidate the L2 of every XCD; RED_PART: the
// block's sums stay in part[(b * 64 + j) * 2 + {0, 1}].  Columns from mq on
// are w itself (with mq = m - 1 the last sum is |w|^2).
template <int C, int ATOM = RED_LAST>
__device__ __forceinline__ void multidot_body(
    const double *__restrict__ Q, uint64_t ldq, int m,
    const double *__restrict__ w, uint64_t n, double *__restrict__ part,
    uint32_t *cnt, double *out, int mq = 0x7fffffff) {
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
      // the row's entries of LB columns are loaded before any of them is
      // used, one memory round trip per LB columns instead of one per
      // column (a load under a per-column branch is waited for at once);
      // columns past the last repeat it, and their sums are never stored
      constexpr int LB = C ? 16 : 32;
#pragma unroll
      for (int h0 = 0; h0 < 32; h0 += LB) {
        double qr[LB], qi[LB];
#pragma unroll
        for (int c = 0; c < LB; c += 8) {
#pragma unroll
          for (int t = 0; t < 8; t++) qr[c + t] = qi[c + t] = 0.0;
          if (h0 + c < mj) {
#pragma unroll
            for (int t = 0; t < 8; t++) {
              const int j = j0 + min(h0 + c + t, mj - 1);
              const bool self = j >= mq;
              const double *src = self ? w : Q;
              const uint64_t q = self ? i : (uint64_t)j * ldq + i;
              if constexpr (C) {
                qr[c + t] = src[2 * q];
                qi[c + t] = src[2 * q + 1];
              } else {
                qr[c + t] = src[q];
              }
            }
          }
        }
#pragma unroll
        for (int t = 0; t < LB; t++) {
          const int j = h0 + t;
          if constexpr (C) {
            acc_re[j] = fma(qr[t], wr, fma(qi[t], wi, acc_re[j]));
            acc_im[j] = fma(qr[t], wi, fma(-qi[t], wr, acc_im[j]));
          } else {
            acc_re[j] = fma(qr[t], wr, acc_re[j]);
          }
        }
      }
    }
    block_sums32<C, ATOM>(acc_re, acc_im, j0, mj, part, out);
  }
  if constexpr (ATOM == RED_LAST) last_block_sum(part, m, cnt, out);
}

template <int C, int ATOM = RED_LAST>
__global__ __launch_bounds__(NT) void multidot(
    const double *__restrict__ Q, uint64_t ldq, int m,
    const double *__restrict__ w, uint64_t n, double *__restrict__ part,
    uint32_t *cnt, double *out) {
  multidot_body<C, ATOM>(Q, ldq, m, w, n, part, cnt, out);
}

// (ar, ai) = row i of Q h over m columns (h: complex pairs).  LB columns'
// entries and coefficients are loaded before the products (one memory round
// trip per LB columns); the products run in column order over the columns
// that exist.
template <int C>
__device__ __forceinline__ void qh_row(const double *__restrict__ Q,
                                       uint64_t ldq, int m, const double *h,
                                       uint64_t i, double &ar, double &ai) {
  ar = 0.0;
  ai = 0.0;
  constexpr int LB = C ? 8 : 16;
  for (int j0 = 0; j0 < m; j0 += LB) {
    const int mj = min(LB, m - j0);
    double qr[LB], qi[LB], hr[LB], hi[LB];
#pragma unroll
    for (int t = 0; t < LB; t++) {
      const int j = j0 + min(t, mj - 1);
      const uint64_t q = (uint64_t)j * ldq + i;
      if constexpr (C) {
        qr[t] = Q[2 * q];
        qi[t] = Q[2 * q + 1];
        hi[t] = h[2 * j + 1];
      } else {
        qr[t] = Q[q];
      }
      hr[t] = h[2 * j];
    }
#pragma unroll
    for (int t = 0; t < LB; t++) {
      if (t < mj) {
        if constexpr (C) {
          ar = fma(qr[t], hr[t], fma(-qi[t], hi[t], ar));
          ai = fma(qr[t], hi[t], fma(qi[t], hr[t], ai));
        } else {
          ar = fma(qr[t], hr[t], ar);
        }
      }
    }
  }
}

// Row i of the first m < 32 columns of Q (real case) into q; entries past
// the last column repeat it.  The loads are issued together.
__device__ __forceinline__ void load_row32(const double *__restrict__ Q,
                                           uint64_t ldq, int m, uint64_t i,
                                           double (&q)[32]) {
#pragma unroll
  for (int c = 0; c < 32; c += 8) {
#pragma unroll
    for (int t = 0; t < 8; t++) q[c + t] = 0.0;
    if (c < m) {
#pragma unroll
      for (int t = 0; t < 8; t++)
        q[c + t] = Q[(uint64_t)min(c + t, m - 1) * ldq + i];
    }
  }
}

// w -= sign Q h (h: m complex pairs); with nrm != nullptr also
// nrm[0] = |w|^2 of the result (last-block reduction; RED_PART: the block's
// sum in part[b * 128], with nrm only a flag)
template <int C, int ATOM = RED_LAST>
__device__ __forceinline__ void multiaxpy_body(
    const double *__restrict__ Q, uint64_t ldq, int m,
    const double *__restrict__ h, double sign, double *__restrict__ w,
    uint64_t n, double *__restrict__ part, uint32_t *cnt, double *nrm) {
  double sq = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * NT) {
    double ar, ai;
    qh_row<C>(Q, ldq, m, h, i, ar, ai);
    if constexpr (C) {
      const double vr = fma(-sign, ar, w[2 * i]);
      const double vi = fma(-sign, ai, w[2 * i + 1]);
      w[2 * i] = vr;
      w[2 * i + 1] = vi;
      sq = fma(vr, vr, fma(vi, vi, sq));
    } else {
      const double v = fma(-sign, ar, w[i]);
      w[i] = v;
      sq = fma(v, v, sq);
    }
  }
  if (nrm == nullptr) return;  // uniform over the grid
  __shared__ double red[NT / 64];
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < NT / 64; v++) s += red[v];
    if constexpr (ATOM == RED_ATOM) {
      atomicAdd(nrm, s);
    } else {
      part[(uint64_t)blockIdx.x * 128] = s;
      part[(uint64_t)blockIdx.x * 128 + 1] = 0.0;
    }
  }
  if constexpr (ATOM == RED_LAST) last_block_sum(part, 1, cnt, nrm);
}

template <int C, int ATOM = RED_LAST>
__global__ __launch_bounds__(NT) void multiaxpy(
    const double *__restrict__ Q, uint64_t ldq, int m,
    const double *__restrict__ h, double sign, double *__restrict__ w,
    uint64_t n, double *__restrict__ part, uint32_t *cnt, double *nrm) {
  multiaxpy_body<C, ATOM>(Q, ldq, m, h, sign, w, n, part, cnt, nrm);
}

// hs[0 .. 2m) = the sums over nblk blocks of the partial sums a previous
// kernel left at part[(b * 64 + j) * 2 + {0, 1}] (RED_PART), formed by
// every block of the next kernel: the threads split the blocks between
// NT / (2 m) groups per value, the groups are combined in LDS.
__device__ void sum_parts(const double *part, uint32_t nblk, int m,
                          double *hs) {
  __shared__ double red[NT];
  const int nv = 2 * m, groups = NT / nv, t = threadIdx.x;
  double s = 0.0;
  if (t < groups * nv) s = strided_part_sum(part, nblk, nv, groups, t);
  red[t] = s;
  __syncthreads();
  if (t < nv) {
    double r = 0.0;
    for (int g = 0; g < groups; g++) r += red[g * nv + t];
    hs[t] = r;
  }
  __syncthreads();
}

// Gram-Schmidt of one Arnoldi step without atomics (single solver):
// gs_dot leaves the blocks' sums of Q^* w; gs_axpy_dot sums them (h1, kept
// in hout by block 0), forms w' = w - Q h1 on its rows and leaves the
// blocks' sums of Q^* w' and of |w'|^2; gs_axpy_scale (below) does the
// second pass and the normalisation.  The kernels share the grid and the
// rows of each thread.
template <int C>
__global__ __launch_bounds__(NT) void gs_dot(const double *__restrict__ Q,
                                             uint64_t ldq, int m,
                                             const double *__restrict__ w,
                                             uint64_t n, double *pout) {
  multidot_body<C, RED_PART>(Q, ldq, m, w, n, pout, nullptr, nullptr);
}

template <int C>
__global__ __launch_bounds__(NT) void gs_axpy_dot(
    const double *__restrict__ Q, uint64_t ldq, int m, const double *pin,
    uint32_t nblk, double *hout, double *w, uint64_t n, double *pout) {
  __shared__ double hs[128];
  if constexpr (!C) {
    if (m < 32) {  // uniform
      // each entry of Q loaded once for both passes (the arithmetic of the
      // two bodies below): w' = w - Q h, then the sums of Q^T w' and, as
      // column m, |w'|^2.  The thread's first row is loaded before the
      // blocks' sums are read, so that both round trips overlap.
      const uint64_t i0 = (uint64_t)blockIdx.x * NT + threadIdx.x;
      const uint64_t stride = (uint64_t)gridDim.x * NT;
      double q[32], wv = 0.0;
      if (i0 < n) {
        load_row32(Q, ldq, m, i0, q);
        wv = w[i0];
      }
      sum_parts(pin, nblk, m, hs);
      if (blockIdx.x == 0 && threadIdx.x < 2 * m)
        hout[threadIdx.x] = hs[threadIdx.x];
      double acc[32], unused[32];
#pragma unroll
      for (int t = 0; t < 32; t++) acc[t] = unused[t] = 0.0;
      for (uint64_t i = i0; i < n; i += stride) {
        if (i != i0) {
          load_row32(Q, ldq, m, i, q);
          wv = w[i];
        }
        double ar = 0.0;
#pragma unroll
        for (int t = 0; t < 32; t++)
          if (t < m) ar = fma(q[t], hs[2 * t], ar);
        const double v = fma(-1.0, ar, wv);
        w[i] = v;
#pragma unroll
        for (int t = 0; t < 32; t++)
          acc[t] = fma(t == m ? v : q[t], v, acc[t]);
      }
      block_sums32<0, RED_PART>(acc, unused, 0, m + 1, pout, nullptr);
      return;
    }
  }
  sum_parts(pin, nblk, m, hs);
  if (blockIdx.x == 0 && threadIdx.x < 2 * m) hout[threadIdx.x] = hs[threadIdx.x];
  multiaxpy_body<C, RED_PART>(Q, ldq, m, hs, 1.0, w, n, nullptr, nullptr,
                              nullptr);
  // column m of the sums: w' itself, |w'|^2
  multidot_body<C, RED_PART>(Q, ldq, m + 1, w, n, pout, nullptr, nullptr, m);
}

__global__ void copy_kernel(const double *src, double *dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void zero_kernel(double *dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = 0.0;
}

// dst[0 .. n) = 0 and acc[0 .. nacc) = 0 (the step's atomic accumulators)
__global__ void zero2_kernel(double *dst, uint64_t n, double *acc,
                             uint32_t nacc) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = 0.0;
  if (i < nacc) acc[i] = 0.0;
}

// dst = src and acc[0 .. nacc) = 0
__global__ void copy2_kernel(const double *src, double *dst, uint64_t n,
                             double *acc, uint32_t nacc) {
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
  if (i < nacc) acc[i] = 0.0;
}

// ---- small complex helpers for the single-thread parts ----------------------

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) {
  return {a.re + b.re, a.im + b.im};
}
__device__ __forceinline__ cplx csub(cplx a, cplx b) {
  return {a.re - b.re, a.im - b.im};
}
__device__ __forceinline__ double cabs_(cplx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// Least-squares state of the current restart cycle (device memory):
// R (kmax x kmax, column-major), g (kmax + 1), rotations cs / sn (kmax),
// all complex pairs.
struct LSQ {
  double *R, *g, *cs, *sn;
  int kmax;
};

// Column k of the Hessenberg matrix (h = hA + hB, subdiagonal hn) added to
// the QR factorisation with Givens rotations by one workgroup (every thread
// calls it): the column and the previous rotations are staged in LDS (loads
// in parallel), so that the sequential rotations of thread 0 wait for no
// memory access; thread 0 stores the new residual estimate |g[k+1]|.
__device__ __forceinline__ void givens_column(const double *hA,
                                              const double *hB, double hn,
                                              LSQ s, int k, double *err) {
  __shared__ double sh[2 * (MAXK + 1)], scs[2 * MAXK], ssn[2 * MAXK];
  for (int e = threadIdx.x; e < 2 * (k + 1); e += NT) sh[e] = hA[e] + hB[e];
  for (int e = threadIdx.x; e < 2 * k; e += NT) {
    scs[e] = s.cs[e];
    ssn[e] = s.sn[e];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  // previous rotations, streamed: rotation j touches entries j and j + 1,
  // and entry j is final (column k of R) after it
  cplx cur = {sh[0], sh[1]};
  for (int j = 0; j < k; j++) {
    const cplx c = {scs[2 * j], scs[2 * j + 1]};
    const cplx sn = {ssn[2 * j], ssn[2 * j + 1]};
    const cplx nx = {sh[2 * (j + 1)], sh[2 * (j + 1) + 1]};
    const cplx t = cadd(cmul(c, cur), cmul(sn, nx));
    cur = csub(cmul(c, nx), cmul(cconj(sn), cur));
    s.R[2 * (j + (uint64_t)k * s.kmax)] = t.re;
    s.R[2 * (j + (uint64_t)k * s.kmax) + 1] = t.im;
  }
  // rotation mapping (h_k, hn) to (r, 0)
  const double na = cabs_(cur), nb = hn;
  cplx c, sn;
  if (nb == 0.0) {
    c = {1.0, 0.0};
    sn = {0.0, 0.0};
  } else if (na == 0.0) {
    c = {0.0, 0.0};
    sn = {1.0, 0.0};
  } else {
    const double nr = hypot(na, nb);
    c = {na / nr, 0.0};
    sn = {cur.re / na * nb / nr, cur.im / na * nb / nr};
  }
  s.cs[2 * k] = c.re;
  s.cs[2 * k + 1] = c.im;
  s.sn[2 * k] = sn.re;
  s.sn[2 * k + 1] = sn.im;
  const cplx rkk = cadd(cmul(c, cur), {sn.re * hn, sn.im * hn});
  s.R[2 * (k + (uint64_t)k * s.kmax)] = rkk.re;
  s.R[2 * (k + (uint64_t)k * s.kmax) + 1] = rkk.im;
  const cplx gk = {s.g[2 * k], s.g[2 * k + 1]};
  const cplx g1 = cmul({-sn.re, sn.im}, gk);  // -conj(s) g_k
  const cplx g0 = cmul(c, gk);
  s.g[2 * k] = g0.re;
  s.g[2 * k + 1] = g0.im;
  s.g[2 * (k + 1)] = g1.re;
  s.g[2 * (k + 1) + 1] = g1.im;
  err[0] = cabs_(g1);
}

// q_{k+1} = w / |w| by all threads; block 0 adds column k of the Hessenberg
// matrix (h = hA + hB, subdiagonal |w|) to the QR factorisation.
// With wz != nullptr, w is left zero for the next product, and nz doubles
// at acc (the other set of accumulators, which the next step adds into) are
// cleared.
__device__ __forceinline__ void scale_givens_body(
    const double *w, double *__restrict__ qn, uint64_t n2,
    const double *nrm2, const double *hA, const double *hB, LSQ s, int k,
    double *err, double *wz = nullptr, double *acc = nullptr,
    uint32_t nz = 0) {
  const double hn = sqrt(nrm2[0]);
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n2) {
    const double wv = w[i];
    qn[i] = hn > 0.0 ? wv / hn : 0.0;
    if (wz) wz[i] = 0.0;  // the same element, after its read
  }
  if (i < nz) acc[i] = 0.0;
  if (blockIdx.x != 0) return;  // uniform per block
  givens_column(hA, hB, hn, s, k, err);
}

__global__ __launch_bounds__(NT) void scale_givens(
    const double *__restrict__ w, double *__restrict__ qn, uint64_t n2,
    const double *nrm2, const double *hA, const double *hB, LSQ s, int k,
    double *err) {
  scale_givens_body(w, qn, n2, nrm2, hA, hB, s, k, err);
}

// scale_givens of the single solver: also w = 0 and the next step's
// accumulators cleared (w is not restrict here: read, then written, by the
// same thread)
__global__ __launch_bounds__(NT) void scale_givens_z(
    double *w, double *__restrict__ qn, uint64_t n2, const double *nrm2,
    const double *hA, const double *hB, LSQ s, int k, double *err,
    double *acc, uint32_t nz) {
  scale_givens_body(w, qn, n2, nrm2, hA, hB, s, k, err, w, acc, nz);
}

// The second Gram-Schmidt pass fused with the normalisation: from the
// blocks' sums that gs_axpy_dot left, h2 = Q^* w' and |w'|^2; then
// w'' = w' - Q h2, whose norm follows from |w''|^2 = |w'|^2 - |h2|^2 (Q has
// orthonormal columns) without a further reduction; q_{k+1} = w'' / |w''|,
// and w = 0 for the next product.  Block 0 keeps h2 in hout and adds
// column k (hA + h2, subdiagonal |w''|) to the QR factorisation.
template <int C>
__global__ __launch_bounds__(NT) void gs_axpy_scale(
    const double *__restrict__ Q, uint64_t ldq, int m, const double *pin,
    uint32_t nblk, const double *hA, double *hout, double *w,
    double *__restrict__ qn, uint64_t n, LSQ s, int k, double *err) {
  __shared__ double hs[128];
  const uint64_t i0 = (uint64_t)blockIdx.x * NT + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * NT;
  // real case, m < 32: the thread's first row of Q and w are loaded before
  // the blocks' sums are read, so that both round trips overlap
  const bool pre = !C && m < 32;
  double q[32], wv = 0.0;
  if (pre && i0 < n) {
    load_row32(Q, ldq, m, i0, q);
    wv = w[i0];
  }
  sum_parts(pin, nblk, m + 1, hs);
  double h2 = 0.0;
  for (int j = 0; j < m; j++)
    h2 = fma(hs[2 * j], hs[2 * j], fma(hs[2 * j + 1], hs[2 * j + 1], h2));
  const double hn = sqrt(fmax(hs[2 * m] - h2, 0.0));
  for (uint64_t i = i0; i < n; i += stride) {
    if constexpr (!C) {
      if (pre) {  // the order of qh_row's products
        if (i != i0) {
          load_row32(Q, ldq, m, i, q);
          wv = w[i];
        }
        double ar = 0.0;
#pragma unroll
        for (int t = 0; t < 32; t++)
          if (t < m) ar = fma(q[t], hs[2 * t], ar);
        const double v = wv - ar;
        qn[i] = hn > 0.0 ? v / hn : 0.0;
        w[i] = 0.0;
        continue;
      }
    }
    double ar, ai;
    qh_row<C>(Q, ldq, m, hs, i, ar, ai);
    if constexpr (C) {
      const double vr = w[2 * i] - ar, vi = w[2 * i + 1] - ai;
      qn[2 * i] = hn > 0.0 ? vr / hn : 0.0;
      qn[2 * i + 1] = hn > 0.0 ? vi / hn : 0.0;
      w[2 * i] = 0.0;
      w[2 * i + 1] = 0.0;
    } else {
      const double v = w[i] - ar;
      qn[i] = hn > 0.0 ? v / hn : 0.0;
      w[i] = 0.0;
    }
  }
  if (blockIdx.x != 0) return;  // uniform per block
  if (threadIdx.x < 2 * m) hout[threadIdx.x] = hs[threadIdx.x];
  givens_column(hA, hs, hn, s, k, err);
}

// q_0 = r / |r| by all threads; thread 0 of block 0 starts the
// least-squares problem: g = (|r|, 0, ...), err = |r|
__device__ __forceinline__ void start_cycle_body(double *q0, uint64_t n2,
                                                 const double *nrm2, LSQ s,
                                                 double *err) {
  const double beta = sqrt(nrm2[0]);
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n2 && beta > 0.0) q0[i] = q0[i] / beta;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int j = 0; j <= s.kmax; j++) {
    s.g[2 * j] = 0.0;
    s.g[2 * j + 1] = 0.0;
  }
  s.g[0] = beta;
  err[0] = beta;
}

__global__ __launch_bounds__(NT) void start_cycle(double *q0, uint64_t n2,
                                                  const double *nrm2, LSQ s,
                                                  double *err) {
  start_cycle_body(q0, n2, nrm2, s, err);
}

// y = -R^{-1} g for the first k columns, as coefficients
// for multiaxpy (which subtracts)
__global__ __launch_bounds__(NT) void back_substitute(LSQ s, int k,
                                                     double *y) {
  // R's upper triangle (column j holds rows 0..j from j (j + 1) / 2 on) and
  // g are staged in LDS by all threads, eight loads in flight per thread;
  // thread 0 then runs the substitution from LDS
  __shared__ double sR[MAXK * (MAXK + 1)], sg[2 * MAXK], sy[2 * MAXK];
  if (blockIdx.x != 0) return;  // uniform per block
  const int ne = k * (k + 1) / 2;
  for (int e0 = threadIdx.x; e0 < ne; e0 += 8 * NT) {
    double vr[8], vi[8];
    int at[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * NT;
      at[u] = -1;
      vr[u] = vi[u] = 0.0;
      if (e < ne) {
        int j = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > e) j--;  // guards the rounding of sqrt
        while ((j + 1) * (j + 2) / 2 <= e) j++;
        const int i = e - j * (j + 1) / 2;
        const uint64_t q = 2 * (i + (uint64_t)j * s.kmax);
        vr[u] = s.R[q];
        vi[u] = s.R[q + 1];
        at[u] = e;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (at[u] >= 0) {
        sR[2 * at[u]] = vr[u];
        sR[2 * at[u] + 1] = vi[u];
      }
  }
  for (int e = threadIdx.x; e < 2 * k; e += NT) sg[e] = s.g[e];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = k - 1; i >= 0; i--) {
      cplx t = {sg[2 * i], sg[2 * i + 1]};
      for (int j = i + 1; j < k; j++) {
        const int o = j * (j + 1) / 2 + i;
        const cplx r = {sR[2 * o], sR[2 * o + 1]};
        t = cadd(t, cmul(r, {sy[2 * j], sy[2 * j + 1]}));  // sy holds -y_j
      }
      const int o = i * (i + 1) / 2 + i;
      const cplx yi = cdiv(t, {sR[2 * o], sR[2 * o + 1]});
      sy[2 * i] = -yi.re;
      sy[2 * i + 1] = -yi.im;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * k; e += NT) y[e] = sy[e];
}

struct Workspace {
  uint64_t n = 0, n2 = 0;  // entries, doubles per vector
  int kmax = 0;
  h2hip::DevBuf<double> Q, w, part, small;
  h2hip::DevBuf<uint32_t> cnt;
  // views into `small`; hA1, hB1, nrm1: the second set of accumulators of
  // the single solver (steps alternate between the two sets)
  double *hA = nullptr, *hB = nullptr, *nrm = nullptr, *y = nullptr,
         *err = nullptr, *R = nullptr, *g = nullptr, *cs = nullptr,
         *sn = nullptr, *hA1 = nullptr, *hB1 = nullptr, *nrm1 = nullptr;
  hipStream_t st = nullptr;
  dim3 gv;
  uint64_t part2 = 0;  // offset of the second partial-sum area in `part`

  int alloc(int C, uint64_t n_, int kmax_, hipStream_t s) {
    n = n_;
    n2 = (C ? 2 : 1) * n;
    kmax = kmax_;
    st = s;
    gv = dim3((uint32_t)((n2 + NT - 1) / NT));
    const uint64_t nsmall = 128 * 3 + 8 + 128 * 2 + 8 + 2 * CHUNK +
                            2 * (uint64_t)kmax * kmax +
                            2 * ((uint64_t)kmax + 1) + 4 * (uint64_t)kmax;
    // partial sums, 128 doubles per block: the multi-dot grid has at most
    // NB blocks, the multi-axpy grid (whose second pass also reduces |w|^2)
    // one block per NT entries
    // (two such areas: the single solver's Gram-Schmidt kernels read one
    // while writing the other)
    const uint64_t nblk = std::max<uint64_t>(NB, (n + NT - 1) / NT);
    part2 = nblk * 64 * 2;
    if (Q.alloc(n2 * kmax) != hipSuccess || w.alloc(n2) != hipSuccess ||
        part.alloc(2 * part2) != hipSuccess ||
        small.alloc(nsmall) != hipSuccess || cnt.alloc(4) != hipSuccess)
      return H2HIP_ENOMEM;
    double *p = small.p;
    hA = p;
    p += 128;
    hB = p;
    p += 128;
    y = p;
    p += 128;
    nrm = p;
    p += 8;
    hA1 = p;
    p += 128;
    hB1 = p;
    p += 128;
    nrm1 = p;
    p += 8;
    err = p;  // two chunks' residual estimates
    p += 2 * CHUNK;
    R = p;
    p += 2 * (uint64_t)kmax * kmax;
    g = p;
    p += 2 * ((uint64_t)kmax + 1);
    cs = p;
    p += 2 * (uint64_t)kmax;
    sn = p;
    // the block counters start at zero; every reduction resets its own
    if (hipMemsetAsync(cnt.p, 0, sizeof(uint32_t) * 4, st) != hipSuccess)
      return H2HIP_EDEVICE;
    return H2HIP_OK;
  }
  LSQ lsq() const { return LSQ{R, g, cs, sn, kmax}; }
  double *col(int j) const { return Q.p + n2 * j; }
};

template <int C>
int gmres(h2hip_addeval_t addeval, void *A, uint64_t n, const double *b,
          double *x, double eps, uint32_t maxiter, uint32_t kmax,
          hipStream_t st, uint32_t *iters, double *resid) {
  Workspace W;
  if (W.alloc(C, n, (int)kmax, st) != H2HIP_OK) {
    h2hip::set_error("h2hip_solve_gmres: device allocation failed");
    return H2HIP_ENOMEM;
  }
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  const dim3 gax((uint32_t)((n + NT - 1) / NT));
  // multi-dot grid: one row per thread up to NB blocks
  const dim3 gdot((uint32_t)std::min<uint64_t>(NB, (n + NT - 1) / NT));

  // The steps' atomic accumulators: set 0 (hA, hB, with y between them,
  // and nrm) and set 1 (hA1, hB1, nrm1) lie one after the other at the
  // start of the small workspace.  Step k adds into set k & 1, and its last
  // kernel clears the other set and w for step k + 1.
  const uint32_t nacc0 = (uint32_t)(W.nrm + 8 - W.small.p);
  const uint32_t nacc1 = (uint32_t)(W.nrm1 + 8 - W.hA1);
  const uint32_t nacc = nacc0 + nacc1;
  // reproducible mode: fixed-order reductions (the last block of a grid
  // adds the blocks' sums in block order), the arithmetic of the batched
  // solver's slots, instead of FP64 atomics
  const bool rep = h2hip::reproducible();
  const dim3 gred = gdot;
  // |b|
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(NT), 0, st, W.nrm,
                     (uint64_t)2);
  if (rep)
    hipLaunchKernelGGL((multidot<C>), gdot, dim3(NT), 0, st, b, (uint64_t)0,
                       1, b, n, W.part.p, W.cnt.p, W.nrm);
  else
    hipLaunchKernelGGL((multidot<C, true>), gdot, dim3(NT), 0, st, b,
                       (uint64_t)0, 1, b, n, W.part.p, W.cnt.p, W.nrm);
  double nb2[2];
  H2HIP_CHECK(hipMemcpyAsync(nb2, W.nrm, sizeof(nb2), hipMemcpyDeviceToHost,
                             st));
  H2HIP_CHECK(hipStreamSynchronize(st));
  const double normb = std::sqrt(nb2[0]);

  // r = b - A x in column 0, normalised; starts the least-squares problem
  // (both sets of accumulators cleared, |r|^2 in set 1, w = 0 for the
  // first step's product)
  const dim3 gz((uint32_t)((std::max<uint64_t>(W.n2, nacc) + NT - 1) / NT));
  auto init = [&]() -> int {
    double *q0 = W.col(0);
    hipLaunchKernelGGL(copy2_kernel, gz, dim3(NT), 0, st, b, q0, W.n2,
                       W.small.p, nacc);
    hipLaunchKernelGGL(zero_kernel, W.gv, dim3(NT), 0, st, W.w.p, W.n2);
    int r = addeval(mone, A, x, q0, st);
    if (r) return r;
    if (rep)  // |r|^2 as the batched solver's cycle start forms it
      hipLaunchKernelGGL((multiaxpy<C>), gax, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, 0, (const double *)W.hB,
                         1.0, q0, n, W.part.p, W.cnt.p + 1, W.nrm1);
    else
      hipLaunchKernelGGL((multidot<C, true>), gred, dim3(NT), 0, st,
                         (const double *)q0, (uint64_t)0, 1,
                         (const double *)q0, n, W.part.p, W.cnt.p, W.nrm1);
    hipLaunchKernelGGL(start_cycle, W.gv, dim3(NT), 0, st, q0, W.n2,
                       (const double *)W.nrm1, W.lsq(), W.err);
    return H2HIP_OK;
  };
  // x += Q_k y with R y = g
  auto finish = [&](int k) -> int {
    if (k == 0) return H2HIP_OK;
    hipLaunchKernelGGL(back_substitute, dim3(1), dim3(NT), 0, st, W.lsq(), k,
                       W.y);
    hipLaunchKernelGGL(multiaxpy<C>, gax, dim3(NT), 0, st,
                       (const double *)W.Q.p, n, k, (const double *)W.y, 1.0,
                       x, n, W.part.p, W.cnt.p + 1, (double *)nullptr);
    return H2HIP_OK;
  };
  // one Arnoldi step on column k; its residual estimate goes to err[slot]
  auto step = [&](int k, int slot) -> int {
    double *qk = W.col(k), *w = W.w.p;
    // this step's accumulators, and the other set, cleared for the next
    double *hA = (k & 1) ? W.hA1 : W.hA, *hB = (k & 1) ? W.hB1 : W.hB;
    double *nrm = (k & 1) ? W.nrm1 : W.nrm;
    double *acc_next = (k & 1) ? W.small.p : W.hA1;
    const uint32_t n_next = (k & 1) ? nacc0 : nacc1;
    int r = addeval(one, A, qk, w, st);
    if (r) return r;
    if (rep) {
      hipLaunchKernelGGL((multidot<C>), gdot, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1, (const double *)w,
                         n, W.part.p, W.cnt.p + 2, hA);
      hipLaunchKernelGGL((multiaxpy<C>), gax, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1,
                         (const double *)hA, 1.0, w, n, W.part.p,
                         W.cnt.p + 1, (double *)nullptr);
      hipLaunchKernelGGL((multidot<C>), gdot, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1, (const double *)w,
                         n, W.part.p, W.cnt.p + 2, hB);
      hipLaunchKernelGGL((multiaxpy<C>), gax, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1,
                         (const double *)hB, 1.0, w, n, W.part.p,
                         W.cnt.p + 1, nrm);
    } else {
      // no atomics: each kernel leaves its blocks' sums for the next one
      double *pa = W.part.p, *pb = W.part.p + W.part2;
      const uint32_t nb = gred.x;
      hipLaunchKernelGGL((gs_dot<C>), gred, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1, (const double *)w,
                         n, pa);
      hipLaunchKernelGGL((gs_axpy_dot<C>), gred, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1, (const double *)pa,
                         nb, hA, w, n, pb);
      hipLaunchKernelGGL((gs_axpy_scale<C>), gred, dim3(NT), 0, st,
                         (const double *)W.Q.p, n, k + 1, (const double *)pb,
                         nb, (const double *)hA, hB, w, W.col(k + 1), n,
                         W.lsq(), k, W.err + slot);
      return H2HIP_OK;
    }
    const dim3 gs((uint32_t)((std::max<uint64_t>(W.n2, n_next) + NT - 1) /
                             NT));
    hipLaunchKernelGGL(scale_givens_z, gs, dim3(NT), 0, st, w, W.col(k + 1),
                       W.n2, (const double *)nrm, (const double *)hA,
                       (const double *)hB, W.lsq(), k, W.err + slot,
                       acc_next, n_next);
    return H2HIP_OK;
  };

  int rc = init();
  if (rc) return rc;
  double error;
  H2HIP_CHECK(hipMemcpyAsync(&error, W.err, sizeof(double),
                             hipMemcpyDeviceToHost, st));
  H2HIP_CHECK(hipStreamSynchronize(st));
  H2HIP_CHECK(hipGetLastError());

  // Chunks of up to CHUNK steps; while the host waits for the residual
  // estimates of one chunk, the next chunk of the same restart cycle is
  // already queued behind it, so the device does not wait for the host.
  // Steps past the stopping point only touch later columns of Q, R and g,
  // and are discarded; a restart waits for the decision.
  struct Look {
    double *host = nullptr;  // pinned, 2 * CHUNK estimates
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Look() {
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
      if (host) (void)hipHostFree(host);
    }
  } L;
  if (hipHostMalloc((void **)&L.host, 2 * CHUNK * sizeof(double),
                    hipHostMallocDefault) != hipSuccess) {
    h2hip::set_error("h2hip_solve_gmres: pinned allocation failed");
    return H2HIP_ENOMEM;
  }
  for (auto &e : L.ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      h2hip::set_error("h2hip_solve_gmres: event creation failed");
      return H2HIP_EDEVICE;
    }
  struct Pending {
    uint32_t len;
    int half;
  } pend[2];
  int npend = 0, next_half = 0;
  int k = 0, kq = 0;           // steps accepted / queued in this cycle
  uint32_t iter = 0, iq = 0;   // steps accepted / queued in all
  bool stop = false;
  while (!stop && error > eps * normb && iter + 1 != maxiter) {
    if (k + 1 >= (int)kmax) {  // nothing is queued past a cycle's end
      rc = finish(k);
      if (rc) return rc;
      rc = init();
      if (rc) return rc;
      k = kq = 0;
    }
    // queue up to two chunks: until the next restart, maxiter, or CHUNK
    while (npend < 2 && kq + 1 < (int)kmax &&
           (maxiter == 0 || iq + 1 < maxiter)) {
      uint32_t len = (uint32_t)((int)kmax - 1 - kq);
      if (len > (uint32_t)CHUNK) len = CHUNK;
      if (maxiter != 0 && len > maxiter - 1 - iq) len = maxiter - 1 - iq;
      const int h = next_half;
      for (uint32_t j = 0; j < len; j++) {
        rc = step(kq + (int)j, h * CHUNK + (int)j);
        if (rc) return rc;
      }
      H2HIP_CHECK(hipGetLastError());
      H2HIP_CHECK(hipMemcpyAsync(L.host + h * CHUNK, W.err + h * CHUNK,
                                 sizeof(double) * len, hipMemcpyDeviceToHost,
                                 st));
      H2HIP_CHECK(hipEventRecord(L.ev[h], st));
      pend[npend++] = {len, h};
      next_half ^= 1;
      kq += (int)len;
      iq += len;
    }
    // replay the reference's loop over the oldest chunk's steps
    const Pending c = pend[0];
    pend[0] = pend[1];
    npend--;
    H2HIP_CHECK(hipEventSynchronize(L.ev[c.half]));
    for (uint32_t j = 0; j < c.len; j++) {
      k++;
      iter++;
      error = L.host[c.half * CHUNK + j];
      if (!(error > eps * normb && iter + 1 != maxiter)) {
        stop = true;
        break;
      }
    }
  }
  rc = finish(k);
  if (rc) return rc;
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipStreamSynchronize(st));
  if (iters) *iters = iter;
  if (resid) *resid = error;
  return H2HIP_OK;
}

// ---- conjugate gradients -----------------------------------------------------
//
// Semantics of solve_cg_avector (Library/krylovsolvers.c:18-52) with
// init_cg / step_cg (Library/krylov.c:162-196): r = b - A x, p = r; per
// step a = A p, gamma = <p, a>, lambda = <p, r> / gamma, x += lambda p,
// r -= lambda a, mu = <a, r> / gamma, p = r - mu p; the iteration stops when
// |r| <= eps |b| or after maxiter - 1 steps.  <u, v> = sum conj(u_i) v_i.
//
// The stopping test runs on the device: the kernel that forms |r| also
// counts the step and sets a done flag when the reference's loop would end;
// the kernels of later steps see the flag and return at once, so x and r
// stay those of the stopping step (only the product a = A p, which nothing
// reads any more, still runs).  Every step is the same launch sequence with
// the same arguments, so the host enqueues CHUNK steps at a time, directly
// or as one captured graph, and reads the flag once per chunk while the
// next chunk runs.  The reductions are deterministic: fixed partial sums
// per block, summed in block order by the last block to finish.

constexpr int CG_NB = 512;  // blocks of the CG vector kernels

// Sums nv <= 8 doubles per block (part[b * 8 + v]) over the grid in the
// last block to finish and leaves the totals in tot[0 .. nv) (LDS) after a
// barrier; true in that block only.  Same fence protocol as last_block_sum.
__device__ bool cg_last_block(const double *part, int nv, uint32_t *cnt,
                              double *tot) {
  __shared__ bool last;
  __shared__ double red[NT];
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    if (last) __threadfence();
  }
  __syncthreads();
  if (!last) return false;
  const int groups = NT / 8, t = threadIdx.x, v = t & 7;
  double s = 0.0;
  if (v < nv)  // in block order, eight loads in flight at a time
    for (uint32_t b0 = t >> 3; b0 < gridDim.x; b0 += 8 * groups) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const uint32_t b = b0 + u * groups;
        x[u] = b < gridDim.x ? part[(uint64_t)b * 8 + v] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) s += x[u];
    }
  red[t] = s;
  __syncthreads();
  if (t < nv) {
    double r = 0.0;
    for (int g = 0; g < groups; g++) r += red[g * 8 + t];
    tot[t] = r;
  }
  if (t == 0) *cnt = 0;
  __syncthreads();
  return true;
}

// block sums of nv <= 8 per-thread values into part[blockIdx.x * 8 + v]
__device__ void cg_block_sums(const double *val, int nv, double *part) {
  __shared__ double red[NT / 64][8];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int v = 0; v < nv; v++) {
    double s = val[v];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) red[wv][v] = s;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; w++) s += red[w][threadIdx.x];
    part[(uint64_t)blockIdx.x * 8 + threadIdx.x] = s;
  }
}

// device state of one CG solve
struct CgDev {
  double *s;       // [0,1] gamma, [2,3] <p, r>, [4,5] <a, r>, [6] |r|^2,
                   // [8] |r|, [9] eps |b|, [10] |b|
  uint32_t *flag;  // [0] done, [1] steps, [2..3] block counters
  uint32_t maxiter;
};

__device__ __forceinline__ void cdiv(double ar, double ai, double br,
                                     double bi, double &cr, double &ci) {
  const double d = br * br + bi * bi;
  cr = (ar * br + ai * bi) / d;
  ci = (ai * br - ar * bi) / d;
}

// r holds b - A x: p = r, a = 0, |b|, |r|, and the test before the first
// step
template <int C>
__device__ void cg_init_body(const double *__restrict__ b,
                             const double *__restrict__ r,
                             double *__restrict__ p, double *__restrict__ a,
                             uint64_t n, double *part, CgDev d, double eps) {
  double acc[2] = {0.0, 0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * NT) {
    if constexpr (C) {
      const double rr = r[2 * i], ri = r[2 * i + 1];
      const double br = b[2 * i], bi = b[2 * i + 1];
      p[2 * i] = rr;
      p[2 * i + 1] = ri;
      a[2 * i] = 0.0;
      a[2 * i + 1] = 0.0;
      acc[0] = fma(rr, rr, fma(ri, ri, acc[0]));
      acc[1] = fma(br, br, fma(bi, bi, acc[1]));
    } else {
      const double rv = r[i], bv = b[i];
      p[i] = rv;
      a[i] = 0.0;
      acc[0] = fma(rv, rv, acc[0]);
      acc[1] = fma(bv, bv, acc[1]);
    }
  }
  cg_block_sums(acc, 2, part);
  __shared__ double tot[8];
  if (!cg_last_block(part, 2, d.flag + 2, tot)) return;
  if (threadIdx.x == 0) {
    const double err = sqrt(tot[0]), nb = sqrt(tot[1]);
    d.s[8] = err;
    d.s[9] = eps * nb;
    d.s[10] = nb;
    d.flag[1] = 0;
    d.flag[0] = (err > eps * nb && 1u != d.maxiter) ? 0u : 1u;
  }
}

// gamma = <p, a>, <p, r>
template <int C>
__device__ void cg_dots_body(const double *__restrict__ p,
                             const double *__restrict__ a,
                             const double *__restrict__ r, uint64_t n,
                             double *part, CgDev d) {
  if (d.flag[0]) return;  // uniform over the grid
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * NT) {
    if constexpr (C) {
      const double pr = p[2 * i], pi = p[2 * i + 1];
      const double ar = a[2 * i], ai = a[2 * i + 1];
      const double rr = r[2 * i], ri = r[2 * i + 1];
      acc[0] = fma(pr, ar, fma(pi, ai, acc[0]));
      acc[1] = fma(pr, ai, fma(-pi, ar, acc[1]));
      acc[2] = fma(pr, rr, fma(pi, ri, acc[2]));
      acc[3] = fma(pr, ri, fma(-pi, rr, acc[3]));
    } else {
      const double pv = p[i];
      acc[0] = fma(pv, a[i], acc[0]);
      acc[2] = fma(pv, r[i], acc[2]);
    }
  }
  cg_block_sums(acc, 4, part);
  __shared__ double tot[8];
  if (!cg_last_block(part, 4, d.flag + 2, tot)) return;
  if (threadIdx.x < 4) d.s[threadIdx.x] = tot[threadIdx.x];
}

// lambda = <p, r> / gamma, x += lambda p, r -= lambda a, then <a, r>,
// |r|^2, mu, the step count and the stopping test
template <int C>
__device__ void cg_update_body(double *__restrict__ x,
                               const double *__restrict__ p,
                               double *__restrict__ r,
                               const double *__restrict__ a, uint64_t n,
                               double *part, CgDev d) {
  if (d.flag[0]) return;
  double lr, li = 0.0;
  if constexpr (C)
    cdiv(d.s[2], d.s[3], d.s[0], d.s[1], lr, li);
  else
    lr = d.s[2] / d.s[0];
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * NT) {
    if constexpr (C) {
      const double pr = p[2 * i], pi = p[2 * i + 1];
      const double ar = a[2 * i], ai = a[2 * i + 1];
      x[2 * i] += lr * pr - li * pi;
      x[2 * i + 1] += lr * pi + li * pr;
      const double rr = r[2 * i] - (lr * ar - li * ai);
      const double ri = r[2 * i + 1] - (lr * ai + li * ar);
      r[2 * i] = rr;
      r[2 * i + 1] = ri;
      acc[0] = fma(ar, rr, fma(ai, ri, acc[0]));
      acc[1] = fma(ar, ri, fma(-ai, rr, acc[1]));
      acc[2] = fma(rr, rr, fma(ri, ri, acc[2]));
    } else {
      const double av = a[i];
      x[i] = fma(lr, p[i], x[i]);
      const double rv = fma(-lr, av, r[i]);
      r[i] = rv;
      acc[0] = fma(av, rv, acc[0]);
      acc[2] = fma(rv, rv, acc[2]);
    }
  }
  cg_block_sums(acc, 3, part);
  __shared__ double tot[8];
  if (!cg_last_block(part, 3, d.flag + 3, tot)) return;
  if (threadIdx.x == 0) {
    d.s[4] = tot[0];
    d.s[5] = tot[1];
    d.s[6] = tot[2];
    const double err = sqrt(tot[2]);
    d.s[8] = err;
    const uint32_t it = d.flag[1] + 1;
    d.flag[1] = it;
    if (!(err > d.s[9] && it + 1 != d.maxiter)) d.flag[0] = 1;
  }
}

// p = r - mu p with mu = <a, r> / gamma; a = 0 for the next product
template <int C>
__device__ void cg_direction_body(double *__restrict__ p,
                                  const double *__restrict__ r,
                                  double *__restrict__ a, uint64_t n,
                                  CgDev d) {
  if (d.flag[0]) return;
  double mr, mi = 0.0;
  if constexpr (C)
    cdiv(d.s[4], d.s[5], d.s[0], d.s[1], mr, mi);
  else
    mr = d.s[4] / d.s[0];
  for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * NT) {
    if constexpr (C) {
      const double pr = p[2 * i], pi = p[2 * i + 1];
      p[2 * i] = r[2 * i] - (mr * pr - mi * pi);
      p[2 * i + 1] = r[2 * i + 1] - (mr * pi + mi * pr);
      a[2 * i] = 0.0;
      a[2 * i + 1] = 0.0;
    } else {
      p[i] = fma(-mr, p[i], r[i]);
      a[i] = 0.0;
    }
  }
}

template <int C>
__global__ __launch_bounds__(NT) void cg_init(const double *b,
                                              const double *r, double *p,
                                              double *a, uint64_t n,
                                              double *part, CgDev d,
                                              double eps) {
  cg_init_body<C>(b, r, p, a, n, part, d, eps);
}

template <int C>
__global__ __launch_bounds__(NT) void cg_dots(const double *p,
                                              const double *a,
                                              const double *r, uint64_t n,
                                              double *part, CgDev d) {
  cg_dots_body<C>(p, a, r, n, part, d);
}

template <int C>
__global__ __launch_bounds__(NT) void cg_update(double *x, const double *p,
                                                double *r, const double *a,
                                                uint64_t n, double *part,
                                                CgDev d) {
  cg_update_body<C>(x, p, r, a, n, part, d);
}

template <int C>
__global__ __launch_bounds__(NT) void cg_direction(double *p,
                                                   const double *r,
                                                   double *a, uint64_t n,
                                                   CgDev d) {
  cg_direction_body<C>(p, r, a, n, d);
}

template <int C>
int cg(h2hip_addeval_t addeval, void *A, uint64_t n, const double *b,
       double *x, double eps, uint32_t maxiter, int use_graph,
       hipStream_t caller, uint32_t *iters, double *resid) {
  const uint64_t n2 = (C ? 2 : 1) * n;
  const dim3 gv((uint32_t)std::min<uint64_t>(CG_NB, (n + NT - 1) / NT));
  const dim3 gc((uint32_t)((n2 + NT - 1) / NT));
  h2hip::DevBuf<double> vec, part, sbuf;
  h2hip::DevBuf<uint32_t> fbuf;
  if (vec.alloc(3 * n2) != hipSuccess ||
      part.alloc((uint64_t)CG_NB * 8) != hipSuccess ||
      sbuf.alloc(16) != hipSuccess || fbuf.alloc(4) != hipSuccess) {
    h2hip::set_error("h2hip_solve_cg: device allocation failed");
    return H2HIP_ENOMEM;
  }
  double *r = vec.p, *p = vec.p + n2, *a = vec.p + 2 * n2;
  const CgDev d{sbuf.p, fbuf.p, maxiter};
  // host copies of (done, steps) and |r| after a chunk; two chunks are in
  // flight, each with its own copy and event
  struct Look {
    uint32_t flag[2];
    double err;
  };
  struct Res {
    Look *look = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // 2 looks, 1 join
    hipStream_t own = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Res() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (own) {
        (void)hipStreamSynchronize(own);
        (void)hipStreamDestroy(own);
      }
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
      if (look) (void)hipHostFree(look);
    }
  } R;
  if (hipHostMalloc((void **)&R.look, 2 * sizeof(Look),
                    hipHostMallocDefault) != hipSuccess) {
    h2hip::set_error("h2hip_solve_cg: pinned allocation failed");
    return H2HIP_ENOMEM;
  }
  for (auto &e : R.ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      h2hip::set_error("h2hip_solve_cg: event creation failed");
      return H2HIP_EDEVICE;
    }
  // the null stream cannot be captured: a graph runs on a stream of its
  // own, ordered after the caller's work
  hipStream_t st = caller;
  if (use_graph && caller == nullptr) {
    if (hipStreamCreateWithFlags(&R.own, hipStreamNonBlocking) !=
            hipSuccess ||
        hipEventRecord(R.ev[2], caller) != hipSuccess ||
        hipStreamWaitEvent(R.own, R.ev[2], 0) != hipSuccess) {
      h2hip::set_error("h2hip_solve_cg: stream setup failed");
      return H2HIP_EDEVICE;
    }
    st = R.own;
  }
  if (hipMemsetAsync(fbuf.p, 0, sizeof(uint32_t) * 4, st) != hipSuccess) {
    h2hip::set_error("h2hip_solve_cg: device error");
    return H2HIP_EDEVICE;
  }
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  // r = b - A x, p = r, a = 0, |b|, |r|
  hipLaunchKernelGGL(copy_kernel, gc, dim3(NT), 0, st, b, r, n2);
  int rc = addeval(mone, A, x, r, st);
  if (rc) return rc;
  hipLaunchKernelGGL(cg_init<C>, gv, dim3(NT), 0, st, b, (const double *)r,
                     p, a, n, part.p, d, eps);
  auto step = [&](hipStream_t s) -> int {
    int e = addeval(one, A, p, a, s);
    if (e) return e;
    hipLaunchKernelGGL(cg_dots<C>, gv, dim3(NT), 0, s, (const double *)p,
                       (const double *)a, (const double *)r, n, part.p, d);
    hipLaunchKernelGGL(cg_update<C>, gv, dim3(NT), 0, s, x,
                       (const double *)p, r, (const double *)a, n, part.p,
                       d);
    hipLaunchKernelGGL(cg_direction<C>, gv, dim3(NT), 0, s, p,
                       (const double *)r, a, n, d);
    return H2HIP_OK;
  };
  if (use_graph) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) !=
        hipSuccess) {
      h2hip::set_error("h2hip_solve_cg: capture failed to start");
      return H2HIP_EDEVICE;
    }
    for (int j = 0; j < CHUNK && !rc; j++) rc = step(st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (!rc && e == hipSuccess &&
        hipGraphInstantiate(&R.exec, graph, nullptr, nullptr, 0) !=
            hipSuccess)
      R.exec = nullptr;
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    if (!R.exec) {
      h2hip::set_error("h2hip_solve_cg: capture of the CG steps failed");
      return H2HIP_EDEVICE;
    }
  }
  auto chunk = [&](int slot) -> int {
    if (R.exec) {
      if (hipGraphLaunch(R.exec, st) != hipSuccess) return H2HIP_EDEVICE;
    } else {
      for (int j = 0; j < CHUNK; j++) {
        int e = step(st);
        if (e) return e;
      }
    }
    if (hipMemcpyAsync(R.look[slot].flag, fbuf.p, 2 * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&R.look[slot].err, sbuf.p + 8, sizeof(double),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(R.ev[slot], st) != hipSuccess)
      return H2HIP_EDEVICE;
    return H2HIP_OK;
  };
  // while the host waits for the look after one chunk, the next chunk is
  // already queued behind it
  int slot = 0;
  rc = chunk(0);
  while (!rc) {
    rc = chunk(slot ^ 1);
    if (rc) break;
    if (hipEventSynchronize(R.ev[slot]) != hipSuccess) {
      rc = H2HIP_EDEVICE;
      break;
    }
    if (R.look[slot].flag[0]) break;
    slot ^= 1;
  }
  if (!rc && (hipStreamSynchronize(st) != hipSuccess ||
              hipGetLastError() != hipSuccess))
    rc = H2HIP_EDEVICE;
  if (rc == H2HIP_EDEVICE) h2hip::set_error("h2hip_solve_cg: device error");
  if (rc) return rc;
  if (iters) *iters = R.look[slot].flag[1];
  if (resid) *resid = R.look[slot].err;
  return H2HIP_OK;
}

// ---- CG without allocations ------------------------------------------------
//
// The solve of cg<C> for callers that cannot allocate (serving loops that
// must not call hipMalloc, harnesses that forbid it): the three work
// vectors, the per-block partial sums and the device state live in one
// file-scope device array of fixed size, `scratch`, so the call allocates
// nothing: no device or pinned memory, no events, streams or graphs.  The
// host looks at the stopping flag after every CHUNK steps through a copy
// into its own stack and a stream synchronisation.  Each GPU holds its own
// copy of the array until the process ends; every call on one GPU shares
// it, so calls that overlap on one GPU race.

constexpr uint64_t CG_NOALLOC_MAXN2 = 1u << 20;  // doubles per work vector
constexpr uint64_t CG_SCR_PART = 3 * CG_NOALLOC_MAXN2;
constexpr uint64_t CG_SCR_S = CG_SCR_PART + (uint64_t)CG_NB * 8;
constexpr uint64_t CG_SCR_FLAG = CG_SCR_S + 16;
constexpr uint64_t CG_SCR_LEN = CG_SCR_FLAG + 4;  // 4 u32 flags in 2 doubles
constexpr int SCR_MAXDEV = 64;

}  // namespace

// work space of h2hip_solve_cg_noalloc: 24 MiB and 32 KiB per GPU
static __device__ __align__(256) double scratch[CG_SCR_LEN];

namespace {

std::mutex scratch_lock;
double *scratch_dev[SCR_MAXDEV] = {};

// the device address of the current device's copy of `scratch`, looked up
// once per device (the array's name is a host-side stand-in in host code)
int scratch_ptr(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCR_MAXDEV) {
    h2hip::set_error("h2hip_solve_cg_noalloc: no current device");
    return H2HIP_EDEVICE;
  }
  std::lock_guard<std::mutex> guard(scratch_lock);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      h2hip::set_error("h2hip_solve_cg_noalloc: hipGetSymbolAddress of the "
                       "work space failed");
      return H2HIP_EDEVICE;
    }
    scratch_dev[dev] = (double *)p;
  }
  *out = scratch_dev[dev];
  return H2HIP_OK;
}

// done flag, step count and block counters start at zero (a kernel, not a
// memset node, should a caller capture the stream)
__global__ void cg_clear_flags(uint32_t *flag) {
  if (threadIdx.x < 4) flag[threadIdx.x] = 0u;
}

template <int C>
int cg_noalloc(h2hip_addeval_t addeval, void *A, uint64_t n, const double *b,
               double *x, double eps, uint32_t maxiter, hipStream_t st,
               uint32_t *iters, double *resid) {
  const uint64_t n2 = (C ? 2 : 1) * n;
  if (n2 > CG_NOALLOC_MAXN2) {
    h2hip::set_error("h2hip_solve_cg_noalloc: the system exceeds the fixed "
                     "work space (at most 1048576 real or 524288 complex "
                     "unknowns); use h2hip_solve_cg");
    return H2HIP_EINVAL;
  }
  double *scr = nullptr;
  int rc = scratch_ptr(&scr);
  if (rc) return rc;
  const dim3 gv((uint32_t)std::min<uint64_t>(CG_NB, (n + NT - 1) / NT));
  const dim3 gc((uint32_t)((n2 + NT - 1) / NT));
  double *r = scr, *p = scr + n2, *a = scr + 2 * n2;
  double *part = scr + CG_SCR_PART;
  const CgDev d{scr + CG_SCR_S, reinterpret_cast<uint32_t *>(scr + CG_SCR_FLAG),
                maxiter};
  hipLaunchKernelGGL(cg_clear_flags, dim3(1), dim3(64), 0, st, d.flag);
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  // r = b - A x, p = r, a = 0, |b|, |r|
  hipLaunchKernelGGL(copy_kernel, gc, dim3(NT), 0, st, b, r, n2);
  rc = addeval(mone, A, x, r, st);
  if (rc) return rc;
  hipLaunchKernelGGL(cg_init<C>, gv, dim3(NT), 0, st, b, (const double *)r,
                     p, a, n, part, d, eps);
  struct {
    uint32_t flag[2];
    double err;
  } look = {{0u, 0u}, 0.0};
  for (;;) {
    // steps after the stopping step leave x and the state as they are
    for (int j = 0; j < CHUNK; j++) {
      rc = addeval(one, A, p, a, st);
      if (rc) return rc;
      hipLaunchKernelGGL(cg_dots<C>, gv, dim3(NT), 0, st, (const double *)p,
                         (const double *)a, (const double *)r, n, part, d);
      hipLaunchKernelGGL(cg_update<C>, gv, dim3(NT), 0, st, x,
                         (const double *)p, r, (const double *)a, n, part, d);
      hipLaunchKernelGGL(cg_direction<C>, gv, dim3(NT), 0, st, p,
                         (const double *)r, a, n, d);
    }
    if (hipMemcpyAsync(look.flag, d.flag, 2 * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&look.err, d.s + 8, sizeof(double),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      h2hip::set_error("h2hip_solve_cg_noalloc: device error");
      return H2HIP_EDEVICE;
    }
    if (look.flag[0]) break;
  }
  if (hipGetLastError() != hipSuccess) {
    h2hip::set_error("h2hip_solve_cg_noalloc: device error");
    return H2HIP_EDEVICE;
  }
  if (iters) *iters = look.flag[1];
  if (resid) *resid = look.err;
  return H2HIP_OK;
}

// ---- batched GMRES: several independent solves sharing every product --------
//
// Up to MAXSLOT solves with the same H2-matrix advance in lock step: one
// "tick" is one batched product (h2hip_addeval_h2matrix_batch, which reads
// the matrix once for all of them) followed by the Gram-Schmidt passes and
// the rotations of every slot, each kernel launched once for all slots
// (grid dimension y = slot).  A slot's tick is either an Arnoldi step on
// column k or the start of a cycle (r = b - A x into column 0).  Each solve
// keeps the semantics of solve_gmres_avector exactly as the single solver
// above does; slots come and go independently (continuous batching).

constexpr int MAXSLOT = 8;

struct TickArgs {
  double *Q[MAXSLOT];   // Krylov basis, kmax columns of n2 doubles
  double *w[MAXSLOT];   // Arnoldi: work vector; cycle start: column 0
  const double *b[MAXSLOT];
  double *hA[MAXSLOT], *hB[MAXSLOT], *nrm[MAXSLOT], *err[MAXSLOT];
  double *part[MAXSLOT];
  uint32_t *cnt[MAXSLOT];
  LSQ lsq[MAXSLOT];
  int k[MAXSLOT];       // Arnoldi column, -1: start of a cycle
  uint64_t n, n2;
};

// Arnoldi slots: w = 0; cycle starts: column 0 = b (the product subtracts
// A x from it)
__global__ __launch_bounds__(NT) void tick_prep(TickArgs a) {
  const int s = blockIdx.y;
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= a.n2) return;
  if (a.k[s] < 0)
    a.Q[s][i] = a.b[s][i];
  else
    a.w[s][i] = 0.0;
}

template <int C>
__global__ __launch_bounds__(NT) void tick_dot(TickArgs a, int pass) {
  const int s = blockIdx.y;
  if (a.k[s] < 0) return;  // uniform per block
  multidot_body<C>(a.Q[s], a.n, a.k[s] + 1, a.w[s], a.n, a.part[s],
                   a.cnt[s] + 2, pass == 0 ? a.hA[s] : a.hB[s]);
}

template <int C>
__global__ __launch_bounds__(NT) void tick_axpy(TickArgs a, int pass) {
  const int s = blockIdx.y;
  const int m = a.k[s] < 0 ? 0 : a.k[s] + 1;
  if (pass == 0 && m == 0) return;  // uniform per block
  multiaxpy_body<C>(a.Q[s], a.n, m, pass == 0 ? a.hA[s] : a.hB[s], 1.0,
                    a.w[s], a.n, a.part[s], a.cnt[s] + 1,
                    pass == 0 ? nullptr : a.nrm[s]);
}

__global__ __launch_bounds__(NT) void tick_final(TickArgs a) {
  const int s = blockIdx.y;
  if (a.k[s] < 0)
    start_cycle_body(a.Q[s], a.n2, a.nrm[s], a.lsq[s], a.err[s]);
  else
    scale_givens_body(a.w[s], a.Q[s] + a.n2 * (a.k[s] + 1), a.n2, a.nrm[s],
                      a.hA[s], a.hB[s], a.lsq[s], a.k[s], a.err[s]);
}

// Batched CG: the same slots, one CG step of every slot per tick (the
// kernels of the single solver above, grid dimension y = slot).  A slot's
// first tick starts its solve (r = b - A x, then the initialisation); the
// device's done flag stops its updates, so the host only looks at the flags
// after each chunk.  Vectors r, p, a live in the first three Krylov columns
// of the slot's workspace, the CG scalars in hA, the flags in cnt.
struct CgTickArgs {
  const double *b[MAXSLOT];
  double *x[MAXSLOT], *r[MAXSLOT], *p[MAXSLOT], *a[MAXSLOT], *part[MAXSLOT];
  CgDev d[MAXSLOT];
  double eps[MAXSLOT];
  int start[MAXSLOT];  // 1: this tick starts the solve
  int slot[MAXSLOT];   // the slot's index (where its look goes)
  uint64_t n, n2;
};

__global__ __launch_bounds__(NT) void cgt_prep(CgTickArgs t) {
  const int s = blockIdx.y;
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (!t.start[s] || i >= t.n2) return;
  t.r[s][i] = t.b[s][i];
}

template <int C>
__global__ __launch_bounds__(NT) void cgt_dots(CgTickArgs t) {
  const int s = blockIdx.y;
  if (t.start[s])
    cg_init_body<C>(t.b[s], t.r[s], t.p[s], t.a[s], t.n, t.part[s], t.d[s],
                    t.eps[s]);
  else
    cg_dots_body<C>(t.p[s], t.a[s], t.r[s], t.n, t.part[s], t.d[s]);
}

template <int C>
__global__ __launch_bounds__(NT) void cgt_update(CgTickArgs t) {
  const int s = blockIdx.y;
  if (t.start[s]) return;  // uniform per block
  cg_update_body<C>(t.x[s], t.p[s], t.r[s], t.a[s], t.n, t.part[s], t.d[s]);
}

template <int C>
__global__ __launch_bounds__(NT) void cgt_direction(CgTickArgs t) {
  const int s = blockIdx.y;
  if (t.start[s]) return;
  cg_direction_body<C>(t.p[s], t.r[s], t.a[s], t.n, t.d[s]);
}

// (done, steps, |r|) of the na slots into out[slot * CHUNK + 0..2]
__global__ void cgt_look(CgTickArgs t, int na, double *out) {
  const int j = threadIdx.x;
  if (j >= na) return;
  double *o = out + (uint64_t)t.slot[j] * CHUNK;
  o[0] = (double)t.d[j].flag[0];
  o[1] = (double)t.d[j].flag[1];
  o[2] = t.d[j].s[8];
}

}  // namespace

struct h2hip_gmres_batch {
  const h2hip_h2matrix *H = nullptr;
  // optional combination c1 H + c2 H2 (complex coefficients)
  const h2hip_h2matrix *H2 = nullptr;
  double c1[2] = {1.0, 0.0}, c2[2] = {0.0, 0.0};
  int C = 0;
  int method = 0;  // 0: GMRES(kmax), 1: CG
  uint64_t n = 0, n2 = 0;
  int kmax = 0, nslots = 0;
  hipStream_t st = nullptr;
  struct Slot {
    Workspace W;
    bool active = false, started = false, need_start = false;
    int k = 0;
    uint32_t iter = 0, maxiter = 0;
    double eps = 0.0, normb = 0.0, error = 0.0;
    const double *b = nullptr;
    double *x = nullptr;
    uint64_t tag = 0;
  } slot[MAXSLOT];
  h2hip::DevBuf<double> d_errs;  // MAXSLOT x CHUNK
  // the chunk in flight (enqueued, not yet polled): its length, the slots
  // that take part and whether their first tick starts a cycle; the
  // residual estimates land in pinned host memory, `done` marks the copy
  bool inflight = false;
  int len = 0;
  int part[MAXSLOT] = {0}, first_start[MAXSLOT] = {0};
  double *h_errs = nullptr;
  hipEvent_t done = nullptr;
  ~h2hip_gmres_batch() {
    if (h_errs) (void)hipHostFree(h_errs);
    if (done) (void)hipEventDestroy(done);
  }
};

namespace {

int batch_finish(h2hip_gmres_batch *G, h2hip_gmres_batch::Slot &S) {
  if (S.k == 0) return H2HIP_OK;
  const dim3 gax((uint32_t)((G->n + NT - 1) / NT));
  hipLaunchKernelGGL(back_substitute, dim3(1), dim3(NT), 0, G->st,
                     S.W.lsq(), S.k, S.W.y);
  if (G->C)
    hipLaunchKernelGGL(multiaxpy<1>, gax, dim3(NT), 0, G->st,
                       (const double *)S.W.Q.p, G->n, S.k,
                       (const double *)S.W.y, 1.0, S.x, G->n, S.W.part.p,
                       S.W.cnt.p + 1, (double *)nullptr);
  else
    hipLaunchKernelGGL(multiaxpy<0>, gax, dim3(NT), 0, G->st,
                       (const double *)S.W.Q.p, G->n, S.k,
                       (const double *)S.W.y, 1.0, S.x, G->n, S.W.part.p,
                       S.W.cnt.p + 1, (double *)nullptr);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// ys[j] += alpha[j] A xs[j] for na vectors, A = H or c1 H + c2 H2
int batch_product(h2hip_gmres_batch *G, int na, const double *alpha,
                  const void *const *xs, void *const *ys) {
  if (!G->H2)
    return h2hip_addeval_h2matrix_batch(G->H, (uint32_t)na, alpha, xs, ys, 0,
                                        G->st);
  // y += alpha (c1 H + c2 H2) x: two batched products
  double a1[2 * MAXSLOT], a2[2 * MAXSLOT];
  for (int j = 0; j < na; j++) {
    const double ar = alpha[2 * j], ai = alpha[2 * j + 1];
    a1[2 * j] = ar * G->c1[0] - ai * G->c1[1];
    a1[2 * j + 1] = ar * G->c1[1] + ai * G->c1[0];
    a2[2 * j] = ar * G->c2[0] - ai * G->c2[1];
    a2[2 * j + 1] = ar * G->c2[1] + ai * G->c2[0];
  }
  int rc = h2hip_addeval_h2matrix_batch(G->H, (uint32_t)na, a1, xs, ys, 0,
                                        G->st);
  if (!rc)
    rc = h2hip_addeval_h2matrix_batch(G->H2, (uint32_t)na, a2, xs, ys, 0,
                                      G->st);
  return rc;
}

// one tick for the slots in `act` (k[s] = -1: start of a cycle)
template <int C>
int batch_tick(h2hip_gmres_batch *G, const int *act, const int *ks, int na,
               int pos) {
  TickArgs a = {};
  a.n = G->n;
  a.n2 = G->n2;
  const void *xs[MAXSLOT];
  void *ys[MAXSLOT];
  double alpha[2 * MAXSLOT];
  for (int j = 0; j < na; j++) {
    auto &S = G->slot[act[j]];
    a.Q[j] = S.W.Q.p;
    a.w[j] = ks[j] < 0 ? S.W.Q.p : S.W.w.p;
    a.b[j] = S.b;
    a.hA[j] = S.W.hA;
    a.hB[j] = S.W.hB;
    a.nrm[j] = S.W.nrm;
    a.err[j] = G->d_errs.p + (uint64_t)act[j] * CHUNK + pos;
    a.part[j] = S.W.part.p;
    a.cnt[j] = S.W.cnt.p;
    a.lsq[j] = S.W.lsq();
    a.k[j] = ks[j];
    xs[j] = ks[j] < 0 ? (const double *)S.x : S.W.col(ks[j]);
    ys[j] = ks[j] < 0 ? S.W.Q.p : S.W.w.p;
    alpha[2 * j] = ks[j] < 0 ? -1.0 : 1.0;
    alpha[2 * j + 1] = 0.0;
  }
  const uint32_t gv = (uint32_t)((G->n2 + NT - 1) / NT);
  const uint32_t gax = (uint32_t)((G->n + NT - 1) / NT);
  const uint32_t gdot = (uint32_t)std::min<uint64_t>(NB, gax);
  hipLaunchKernelGGL(tick_prep, dim3(gv, na), dim3(NT), 0, G->st, a);
  int rc = batch_product(G, na, alpha, xs, ys);
  if (rc) return rc;
  hipLaunchKernelGGL(tick_dot<C>, dim3(gdot, na), dim3(NT), 0, G->st, a, 0);
  hipLaunchKernelGGL(tick_axpy<C>, dim3(gax, na), dim3(NT), 0, G->st, a, 0);
  hipLaunchKernelGGL(tick_dot<C>, dim3(gdot, na), dim3(NT), 0, G->st, a, 1);
  hipLaunchKernelGGL(tick_axpy<C>, dim3(gax, na), dim3(NT), 0, G->st, a, 1);
  hipLaunchKernelGGL(tick_final, dim3(gv, na), dim3(NT), 0, G->st, a);
  H2HIP_CHECK(hipGetLastError());
  return H2HIP_OK;
}

// CG: CHUNK ticks of every active slot, then the look at the flags
template <int C>
int cg_batch_enqueue(h2hip_gmres_batch *G) {
  CgTickArgs t = {};
  t.n = G->n;
  t.n2 = G->n2;
  int act[MAXSLOT], na = 0;
  for (int s = 0; s < G->nslots; s++) {
    G->part[s] = G->slot[s].active ? 1 : 0;
    if (G->slot[s].active) act[na++] = s;
  }
  if (na == 0) return H2HIP_OK;
  bool start0[MAXSLOT];
  for (int j = 0; j < na; j++) {
    auto &S = G->slot[act[j]];
    t.b[j] = S.b;
    t.x[j] = S.x;
    t.r[j] = S.W.col(0);
    t.p[j] = S.W.col(1);
    t.a[j] = S.W.col(2);
    t.part[j] = S.W.part.p;
    t.d[j] = CgDev{S.W.hA, S.W.cnt.p, S.maxiter};
    t.eps[j] = S.eps;
    t.slot[j] = act[j];
    start0[j] = S.need_start;
    S.need_start = false;
  }
  const uint32_t gv = (uint32_t)std::min<uint64_t>(CG_NB, (G->n + NT - 1) /
                                                              NT);
  const uint32_t gc = (uint32_t)((G->n2 + NT - 1) / NT);
  for (int tick = 0; tick < CHUNK; tick++) {
    const void *xs[MAXSLOT];
    void *ys[MAXSLOT];
    double alpha[2 * MAXSLOT];
    bool any_start = false;
    for (int j = 0; j < na; j++) {
      const bool st = tick == 0 && start0[j];
      t.start[j] = st ? 1 : 0;
      any_start |= st;
      xs[j] = st ? (const double *)t.x[j] : (const double *)t.p[j];
      ys[j] = st ? t.r[j] : t.a[j];
      alpha[2 * j] = st ? -1.0 : 1.0;
      alpha[2 * j + 1] = 0.0;
    }
    if (any_start)
      hipLaunchKernelGGL(cgt_prep, dim3(gc, na), dim3(NT), 0, G->st, t);
    int rc = batch_product(G, na, alpha, xs, ys);
    if (rc) return rc;
    hipLaunchKernelGGL(cgt_dots<C>, dim3(gv, na), dim3(NT), 0, G->st, t);
    hipLaunchKernelGGL(cgt_update<C>, dim3(gv, na), dim3(NT), 0, G->st, t);
    hipLaunchKernelGGL(cgt_direction<C>, dim3(gv, na), dim3(NT), 0, G->st,
                       t);
  }
  hipLaunchKernelGGL(cgt_look, dim3(1), dim3(64), 0, G->st, t, na,
                     G->d_errs.p);
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipMemcpyAsync(G->h_errs, G->d_errs.p,
                             sizeof(double) * MAXSLOT * CHUNK,
                             hipMemcpyDeviceToHost, G->st));
  H2HIP_CHECK(hipEventRecord(G->done, G->st));
  G->len = CHUNK;
  G->inflight = true;
  return H2HIP_OK;
}

// Enqueues one chunk: restarts due now, up to CHUNK ticks (none past a
// restart or maxiter of any slot), the copy of the residual estimates to
// pinned host memory and an event behind it.  Nothing waits.
template <int C>
int batch_enqueue(h2hip_gmres_batch *G) {
  if (G->method == 1) return cg_batch_enqueue<C>(G);
  // restarts due at this point
  for (int s = 0; s < G->nslots; s++) {
    auto &S = G->slot[s];
    if (S.active && S.started && !S.need_start && S.k + 1 >= G->kmax) {
      int rc = batch_finish(G, S);
      if (rc) return rc;
      S.need_start = true;
    }
  }
  // plan the chunk length
  int len = CHUNK, nact = 0;
  for (int s = 0; s < G->nslots; s++) {
    auto &S = G->slot[s];
    G->part[s] = S.active ? 1 : 0;
    if (!S.active) continue;
    nact++;
    const int steps_cycle = S.need_start ? G->kmax - 1 : G->kmax - 1 - S.k;
    int lim = S.need_start ? 1 + steps_cycle : steps_cycle;
    if (S.maxiter != 0) {
      const int left = (int)(S.maxiter - 1 - S.iter);
      lim = std::min(lim, (S.need_start ? 1 : 0) + left);
    }
    len = std::min(len, std::max(lim, 1));
  }
  if (nact == 0) return H2HIP_OK;
  // enqueue; the host follows each slot's (speculative) column index
  int kk[MAXSLOT];
  for (int s = 0; s < G->nslots; s++) {
    kk[s] = G->slot[s].need_start ? -1 : G->slot[s].k;
    G->first_start[s] = G->slot[s].need_start ? 1 : 0;
  }
  for (int t = 0; t < len; t++) {
    int act[MAXSLOT], ks[MAXSLOT], na = 0;
    for (int s = 0; s < G->nslots; s++) {
      if (!G->part[s]) continue;
      act[na] = s;
      ks[na] = kk[s];
      na++;
      kk[s] = kk[s] < 0 ? 0 : kk[s] + 1;
    }
    int rc = batch_tick<C>(G, act, ks, na, t);
    if (rc) return rc;
  }
  H2HIP_CHECK(hipMemcpyAsync(G->h_errs, G->d_errs.p,
                             sizeof(double) * MAXSLOT * CHUNK,
                             hipMemcpyDeviceToHost, G->st));
  H2HIP_CHECK(hipEventRecord(G->done, G->st));
  G->len = len;
  G->inflight = true;
  return H2HIP_OK;
}

// Completes the chunk in flight once its estimates have arrived: replays
// the reference's loop per slot, finishes the slots that stop (waiting for
// their solution update).  *pending = 1 if the chunk is still running.
int batch_complete(h2hip_gmres_batch *G, int *pending, uint32_t *ndone,
                   uint64_t *tags, uint32_t *iters, double *resid,
                   uint32_t cap) {
  *ndone = 0;
  *pending = 0;
  if (!G->inflight) return H2HIP_OK;
  const hipError_t q = hipEventQuery(G->done);
  if (q == hipErrorNotReady) {
    *pending = 1;
    return H2HIP_OK;
  }
  H2HIP_CHECK(q);
  G->inflight = false;
  const double *errs = G->h_errs;
  const int len = G->len;
  uint32_t nd = 0;
  if (G->method == 1) {
    // CG: x is final once the device has set the slot's done flag (the
    // chunk's last kernels ran before the event)
    for (int s = 0; s < G->nslots; s++) {
      auto &S = G->slot[s];
      if (!G->part[s] || !S.active) continue;
      const double *o = errs + (uint64_t)s * CHUNK;
      if (o[0] == 0.0) continue;
      S.active = false;
      S.iter = (uint32_t)o[1];
      S.error = o[2];
      if (nd < cap) {
        if (tags) tags[nd] = S.tag;
        if (iters) iters[nd] = S.iter;
        if (resid) resid[nd] = S.error;
      }
      nd++;
    }
    *ndone = nd;
    return H2HIP_OK;
  }
  for (int s = 0; s < G->nslots; s++) {
    auto &S = G->slot[s];
    if (!G->part[s] || !S.active) continue;
    bool done = false;
    int t = 0;
    if (G->first_start[s]) {
      S.need_start = false;
      S.k = 0;
      if (!S.started) {
        // the check before the first step (error = |r_0|)
        S.started = true;
        S.error = errs[s * CHUNK];
        done = !(S.error > S.eps * S.normb && S.iter + 1 != S.maxiter);
      }
      t = 1;
    }
    for (; !done && t < len; t++) {
      S.k++;
      S.iter++;
      S.error = errs[s * CHUNK + t];
      done = !(S.error > S.eps * S.normb && S.iter + 1 != S.maxiter);
    }
    if (done) {
      int rc = batch_finish(G, S);
      if (rc) return rc;
      S.active = false;
      if (nd < cap) {
        if (tags) tags[nd] = S.tag;
        if (iters) iters[nd] = S.iter;
        if (resid) resid[nd] = S.error;
      }
      nd++;
    }
  }
  if (nd) H2HIP_CHECK(hipStreamSynchronize(G->st));
  *ndone = nd;
  return H2HIP_OK;
}

// One chunk, enqueued and completed before returning.
template <int C>
int batch_chunk(h2hip_gmres_batch *G, uint32_t *ndone, uint64_t *tags,
                uint32_t *iters, double *resid, uint32_t cap) {
  *ndone = 0;
  if (!G->inflight) {
    int rc = batch_enqueue<C>(G);
    if (rc) return rc;
    if (!G->inflight) return H2HIP_OK;
  }
  H2HIP_CHECK(hipEventSynchronize(G->done));
  int pending = 0;
  return batch_complete(G, &pending, ndone, tags, iters, resid, cap);
}

}  // namespace

extern "C" {

int h2hip_gmres_batch_create(h2hip_gmres_batch **out, const h2hip_h2matrix *H,
                             int is_complex, uint64_t n, uint32_t slots,
                             uint32_t kmax, void *stream) {
  if (!out || !H || n == 0 || slots == 0 || slots > MAXSLOT || kmax < 2 ||
      kmax > MAXK) {
    h2hip::set_error("h2hip_gmres_batch_create: invalid arguments "
                     "(1..8 slots, kmax 2..64)");
    return H2HIP_EINVAL;
  }
  h2hip_gmres_batch *G = new h2hip_gmres_batch;
  G->H = H;
  G->C = is_complex ? 1 : 0;
  G->n = n;
  G->n2 = (G->C ? 2 : 1) * n;
  G->kmax = (int)kmax;
  G->nslots = (int)slots;
  G->st = (hipStream_t)stream;
  for (uint32_t s = 0; s < slots; s++)
    if (G->slot[s].W.alloc(G->C, n, (int)kmax, G->st) != H2HIP_OK) {
      delete G;
      h2hip::set_error("h2hip_gmres_batch_create: device allocation failed");
      return H2HIP_ENOMEM;
    }
  if (G->d_errs.alloc(MAXSLOT * CHUNK) != hipSuccess ||
      hipHostMalloc((void **)&G->h_errs, sizeof(double) * MAXSLOT * CHUNK,
                    hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&G->done, hipEventDisableTiming) !=
          hipSuccess) {
    delete G;
    h2hip::set_error("h2hip_gmres_batch_create: allocation failed");
    return H2HIP_ENOMEM;
  }
  *out = G;
  return H2HIP_OK;
}

void h2hip_gmres_batch_destroy(h2hip_gmres_batch *G) { delete G; }

int h2hip_gmres_batch_combine(h2hip_gmres_batch *G, const double *c1,
                              const h2hip_h2matrix *H2, const double *c2) {
  if (!G || !c1 || !H2 || !c2) {
    h2hip::set_error("h2hip_gmres_batch_combine: invalid arguments");
    return H2HIP_EINVAL;
  }
  for (int s = 0; s < G->nslots; s++)
    if (G->slot[s].active) {
      h2hip::set_error("h2hip_gmres_batch_combine: solves in progress");
      return H2HIP_EINVAL;
    }
  if (G->method == 1) {
    // a combination such as KM - i k V is not Hermitian: CG would not
    // converge
    h2hip::set_error("h2hip_gmres_batch_combine: the solver is CG, which "
                     "needs a Hermitian positive definite operator; switch "
                     "to GMRES first");
    return H2HIP_EINVAL;
  }
  G->c1[0] = c1[0];
  G->c1[1] = c1[1];
  G->H2 = H2;
  G->c2[0] = c2[0];
  G->c2[1] = c2[1];
  return H2HIP_OK;
}

int h2hip_gmres_batch_set_method(h2hip_gmres_batch *G, int method) {
  if (!G || (method != 0 && method != 1) || (method == 1 && G->kmax < 3)) {
    h2hip::set_error("h2hip_gmres_batch_set_method: invalid arguments "
                     "(0: GMRES, 1: CG, which needs kmax >= 3)");
    return H2HIP_EINVAL;
  }
  for (int s = 0; s < G->nslots; s++)
    if (G->slot[s].active || G->inflight) {
      h2hip::set_error("h2hip_gmres_batch_set_method: solves in progress");
      return H2HIP_EINVAL;
    }
  if (method == 1 && G->H2) {
    h2hip::set_error("h2hip_gmres_batch_set_method: CG needs a Hermitian "
                     "positive definite operator, not the combination set "
                     "by h2hip_gmres_batch_combine");
    return H2HIP_EINVAL;
  }
  if (method == G->method) return H2HIP_OK;
  // the flags and block counters of both methods share cnt: start at zero
  for (int s = 0; s < G->nslots; s++)
    H2HIP_CHECK(hipMemsetAsync(G->slot[s].W.cnt.p, 0, sizeof(uint32_t) * 4,
                               G->st));
  H2HIP_CHECK(hipStreamSynchronize(G->st));
  G->method = method;
  return H2HIP_OK;
}

int h2hip_gmres_batch_free_slots(const h2hip_gmres_batch *G) {
  if (!G) return H2HIP_EINVAL;
  int f = 0;
  for (int s = 0; s < G->nslots; s++) f += !G->slot[s].active;
  return f;
}

int h2hip_gmres_batch_submit(h2hip_gmres_batch *G, const void *d_b,
                             void *d_x, double eps, uint32_t maxiter,
                             uint64_t tag) {
  if (!G || !d_b || !d_x) return H2HIP_EINVAL;
  if (G->inflight) {
    h2hip::set_error("h2hip_gmres_batch_submit: a chunk is in flight "
                     "(poll it first)");
    return H2HIP_EINVAL;
  }
  int s = 0;
  while (s < G->nslots && G->slot[s].active) s++;
  if (s == G->nslots) {
    h2hip::set_error("h2hip_gmres_batch_submit: no free slot");
    return H2HIP_EINVAL;
  }
  auto &S = G->slot[s];
  S.b = (const double *)d_b;
  S.x = (double *)d_x;
  S.eps = eps;
  S.maxiter = maxiter;
  S.tag = tag;
  S.k = 0;
  S.iter = 0;
  S.started = false;
  S.need_start = true;
  if (G->method == 1) {  // CG forms |b| on the device
    S.active = true;
    return s;
  }
  // |b| (one small reduction and a wait, once per request)
  const uint32_t gdot = (uint32_t)std::min<uint64_t>(NB, (G->n + NT - 1) / NT);
  if (G->C)
    hipLaunchKernelGGL(multidot<1>, dim3(gdot), dim3(NT), 0, G->st, S.b,
                       (uint64_t)0, 1, S.b, G->n, S.W.part.p, S.W.cnt.p,
                       S.W.nrm);
  else
    hipLaunchKernelGGL(multidot<0>, dim3(gdot), dim3(NT), 0, G->st, S.b,
                       (uint64_t)0, 1, S.b, G->n, S.W.part.p, S.W.cnt.p,
                       S.W.nrm);
  double nb2[2];
  H2HIP_CHECK(hipMemcpyAsync(nb2, S.W.nrm, sizeof(nb2), hipMemcpyDeviceToHost,
                             G->st));
  H2HIP_CHECK(hipStreamSynchronize(G->st));
  S.normb = std::sqrt(nb2[0]);
  S.active = true;
  return s;
}

int h2hip_gmres_batch_run(h2hip_gmres_batch *G, uint32_t *ndone,
                          uint64_t *tags, uint32_t *iterations,
                          double *residuals, uint32_t cap) {
  if (!G || !ndone) return H2HIP_EINVAL;
  *ndone = 0;
  return G->C ? batch_chunk<1>(G, ndone, tags, iterations, residuals, cap)
              : batch_chunk<0>(G, ndone, tags, iterations, residuals, cap);
}

int h2hip_gmres_batch_enqueue(h2hip_gmres_batch *G) {
  if (!G) return H2HIP_EINVAL;
  if (G->inflight) return H2HIP_OK;
  return G->C ? batch_enqueue<1>(G) : batch_enqueue<0>(G);
}

int h2hip_gmres_batch_poll(h2hip_gmres_batch *G, uint32_t *ndone,
                           uint64_t *tags, uint32_t *iterations,
                           double *residuals, uint32_t cap) {
  if (!G || !ndone) return H2HIP_EINVAL;
  int pending = 0;
  int rc = batch_complete(G, &pending, ndone, tags, iterations, residuals,
                          cap);
  if (rc) return rc;
  return pending ? 1 : 0;
}

int h2hip_gmres_batch_busy(const h2hip_gmres_batch *G) {
  if (!G) return H2HIP_EINVAL;
  int a = 0;
  for (int s = 0; s < G->nslots; s++) a += G->slot[s].active;
  return (G->inflight ? 2 : 0) | (a ? 1 : 0);
}

}  // extern "C"

// ---- many systems, one Arnoldi step of all of them per graph launch ---------
//
// Independent systems (typically one per mesh, each with its own H2-matrix)
// step in lock step: they start together and step together, and a system
// leaves when it meets its stopping rule.  The step kernels read the Krylov
// column index k from device memory and the last one increments it, so one
// step of every active system is a fixed sequence of launches that is
// captured once into a HIP graph (one branch per H2-matrix, the branches
// independent) and replayed for every step until the set of active systems
// changes.  The product reads a copy of column k (made by the first kernel
// of the step), so its launch arguments do not depend on k either.

namespace {

__global__ __launch_bounds__(NT) void mstep_prep(const int *dk,
                                                 const double *Q, uint64_t n2,
                                                 double *xb, double *w) {
  const uint64_t k = (uint64_t)*dk;
  const uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (i < n2) {
    xb[i] = Q[n2 * k + i];
    w[i] = 0.0;
  }
}

template <int C>
__global__ __launch_bounds__(NT) void multidot_dk(
    const double *__restrict__ Q, uint64_t ldq, const int *dk,
    const double *__restrict__ w, uint64_t n, double *__restrict__ part,
    uint32_t *cnt, double *out) {
  multidot_body<C>(Q, ldq, *dk + 1, w, n, part, cnt, out);
}

template <int C>
__global__ __launch_bounds__(NT) void multiaxpy_dk(
    const double *__restrict__ Q, uint64_t ldq, const int *dk,
    const double *__restrict__ h, double *__restrict__ w, uint64_t n,
    double *__restrict__ part, uint32_t *cnt, double *nrm) {
  multiaxpy_body<C>(Q, ldq, *dk + 1, h, 1.0, w, n, part, cnt, nrm);
}

// scale_givens for column k = *dk; the residual estimate goes to err[k]
__global__ __launch_bounds__(NT) void scale_givens_dk(
    const double *__restrict__ w, double *__restrict__ Q, uint64_t n2,
    const double *nrm2, const double *hA, const double *hB, LSQ s,
    const int *dk, double *err) {
  const int k = *dk;
  scale_givens_body(w, Q + n2 * (uint64_t)(k + 1), n2, nrm2, hA, hB, s, k,
                    err + k);
}

__global__ void bump_k(int *dk) { *dk += 1; }
__global__ void set_k(int *dk, int v) { *dk = v; }

}  // namespace

struct h2hip_gmres_multi {
  int C = 0, kmax = 0;
  hipStream_t st = nullptr;
  struct Sys {
    const h2hip_h2matrix *H = nullptr;
    uint64_t n = 0, n2 = 0;
    Workspace W;
    h2hip::DevBuf<double> xb;
    int *dk = nullptr;       // device: Krylov column of the next step
    double *derr = nullptr;  // device: kmax step estimates, then |r_0|
    bool active = false;
    int k = 0;
    uint32_t iter = 0;
    double normb = 0.0, error = 0.0;
    const double *b = nullptr;
    double *x = nullptr;
  };
  std::deque<Sys> sys;  // Sys holds device buffers and is not movable
  h2hip::DevBuf<int> d_k;
  h2hip::DevBuf<double> d_err;  // count x (kmax + 1)
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint32_t> exec_set;  // active systems the graph was made for
  uint64_t graph_builds = 0;
  ~h2hip_gmres_multi() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (auto e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (auto b : branch) (void)hipStreamDestroy(b);
  }
};

namespace {

using Multi = h2hip_gmres_multi;

template <int C>
int multi_init(Multi *G, Multi::Sys &S, hipStream_t st) {
  const double mone[2] = {-1.0, 0.0};
  const dim3 gdot((uint32_t)std::min<uint64_t>(NB, (S.n + NT - 1) / NT));
  double *q0 = S.W.col(0);
  hipLaunchKernelGGL(copy_kernel, S.W.gv, dim3(NT), 0, st, S.b, q0, S.n2);
  int r = h2hip_addeval_h2matrix(mone, S.H, S.x, q0, st);
  if (r) return r;
  if (h2hip::reproducible())  // |r|^2 as the other solvers form it
    hipLaunchKernelGGL(multiaxpy<C>, dim3((uint32_t)((S.n + NT - 1) / NT)),
                       dim3(NT), 0, st, (const double *)S.W.Q.p, S.n, 0,
                       (const double *)S.W.hB, 1.0, q0, S.n, S.W.part.p,
                       S.W.cnt.p + 1, S.W.nrm);
  else
    hipLaunchKernelGGL(multidot<C>, gdot, dim3(NT), 0, st,
                       (const double *)q0, (uint64_t)0, 1,
                       (const double *)q0, S.n, S.W.part.p, S.W.cnt.p,
                       S.W.nrm);
  hipLaunchKernelGGL(start_cycle, S.W.gv, dim3(NT), 0, st, q0, S.n2,
                     (const double *)S.W.nrm, S.W.lsq(),
                     S.derr + G->kmax);
  hipLaunchKernelGGL(set_k, dim3(1), dim3(1), 0, st, S.dk, 0);
  return H2HIP_OK;
}

template <int C>
int multi_finish(Multi::Sys &S, int k, hipStream_t st) {
  if (k == 0) return H2HIP_OK;
  const dim3 gax((uint32_t)((S.n + NT - 1) / NT));
  hipLaunchKernelGGL(back_substitute, dim3(1), dim3(NT), 0, st, S.W.lsq(), k,
                     S.W.y);
  hipLaunchKernelGGL(multiaxpy<C>, gax, dim3(NT), 0, st,
                     (const double *)S.W.Q.p, S.n, k, (const double *)S.W.y,
                     1.0, S.x, S.n, S.W.part.p, S.W.cnt.p + 1,
                     (double *)nullptr);
  return H2HIP_OK;
}

// one Arnoldi step of system S on stream st (column index on the device)
template <int C>
int multi_step(Multi::Sys &S, hipStream_t st) {
  const double one[2] = {1.0, 0.0};
  const dim3 gax((uint32_t)((S.n + NT - 1) / NT));
  const dim3 gdot((uint32_t)std::min<uint64_t>(NB, (S.n + NT - 1) / NT));
  hipLaunchKernelGGL(mstep_prep, S.W.gv, dim3(NT), 0, st,
                     (const int *)S.dk, (const double *)S.W.Q.p, S.n2,
                     S.xb.p, S.W.w.p);
  int r = h2hip_addeval_h2matrix(one, S.H, S.xb.p, S.W.w.p, st);
  if (r) return r;
  hipLaunchKernelGGL(multidot_dk<C>, gdot, dim3(NT), 0, st,
                     (const double *)S.W.Q.p, S.n, (const int *)S.dk,
                     (const double *)S.W.w.p, S.n, S.W.part.p, S.W.cnt.p + 2,
                     S.W.hA);
  hipLaunchKernelGGL(multiaxpy_dk<C>, gax, dim3(NT), 0, st,
                     (const double *)S.W.Q.p, S.n, (const int *)S.dk,
                     (const double *)S.W.hA, S.W.w.p, S.n, S.W.part.p,
                     S.W.cnt.p + 1, (double *)nullptr);
  hipLaunchKernelGGL(multidot_dk<C>, gdot, dim3(NT), 0, st,
                     (const double *)S.W.Q.p, S.n, (const int *)S.dk,
                     (const double *)S.W.w.p, S.n, S.W.part.p, S.W.cnt.p + 2,
                     S.W.hB);
  hipLaunchKernelGGL(multiaxpy_dk<C>, gax, dim3(NT), 0, st,
                     (const double *)S.W.Q.p, S.n, (const int *)S.dk,
                     (const double *)S.W.hB, S.W.w.p, S.n, S.W.part.p,
                     S.W.cnt.p + 1, S.W.nrm);
  hipLaunchKernelGGL(scale_givens_dk, S.W.gv, dim3(NT), 0, st,
                     (const double *)S.W.w.p, S.W.Q.p, S.n2,
                     (const double *)S.W.nrm, (const double *)S.W.hA,
                     (const double *)S.W.hB, S.W.lsq(), (const int *)S.dk,
                     S.derr);
  hipLaunchKernelGGL(bump_k, dim3(1), dim3(1), 0, st, S.dk);
  return H2HIP_OK;
}

// Captures one step of the active systems into G->exec: one branch per
// distinct H2-matrix (systems sharing a matrix share its coefficient
// buffers, so they step one after the other on their branch).
template <int C>
int multi_capture(Multi *G, const std::vector<uint32_t> &act) {
  if (G->exec) {
    H2HIP_CHECK(hipGraphExecDestroy(G->exec));
    G->exec = nullptr;
  }
  std::vector<const h2hip_h2matrix *> mats;
  std::vector<int> br(act.size());
  for (size_t j = 0; j < act.size(); j++) {
    const h2hip_h2matrix *H = G->sys[act[j]].H;
    size_t b = 0;
    while (b < mats.size() && mats[b] != H) b++;
    if (b == mats.size()) mats.push_back(H);
    br[j] = (int)b;
  }
  while (G->branch.size() < mats.size()) {
    hipStream_t s;
    hipEvent_t e;
    H2HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    G->branch.push_back(s);
    H2HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    G->join.push_back(e);
  }
  H2HIP_CHECK(hipStreamBeginCapture(G->st, hipStreamCaptureModeThreadLocal));
  int rc = H2HIP_OK;
  hipError_t e = hipEventRecord(G->fork, G->st);
  for (size_t b = 0; b < mats.size() && e == hipSuccess && !rc; b++) {
    e = hipStreamWaitEvent(G->branch[b], G->fork, 0);
    for (size_t j = 0; j < act.size() && e == hipSuccess && !rc; j++)
      if (br[j] == (int)b) rc = multi_step<C>(G->sys[act[j]], G->branch[b]);
    if (e == hipSuccess && !rc) e = hipEventRecord(G->join[b], G->branch[b]);
    if (e == hipSuccess && !rc) e = hipStreamWaitEvent(G->st, G->join[b], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(G->st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    h2hip::set_error(std::string("GMRES step capture failed: ") +
                     hipGetErrorString(e != hipSuccess ? e : e2));
    return H2HIP_EDEVICE;
  }
  const hipError_t e3 = hipGraphInstantiate(&G->exec, graph, nullptr,
                                            nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e3 != hipSuccess) {
    G->exec = nullptr;
    h2hip::set_error(std::string("GMRES step graph: ") +
                     hipGetErrorString(e3));
    return H2HIP_EDEVICE;
  }
  G->exec_set = act;
  G->graph_builds++;
  return H2HIP_OK;
}

template <int C>
int multi_solve(Multi *G, double eps, uint32_t maxiter, int use_graph,
                uint32_t *iters, double *resid) {
  const uint32_t count = (uint32_t)G->sys.size();
  const int kmax = G->kmax;
  const uint64_t stride = (uint64_t)kmax + 1;
  // |b| of every system
  std::vector<double> nb2(2 * count);
  for (uint32_t s = 0; s < count; s++) {
    auto &S = G->sys[s];
    const dim3 gdot((uint32_t)std::min<uint64_t>(NB, (S.n + NT - 1) / NT));
    hipLaunchKernelGGL(multidot<C>, gdot, dim3(NT), 0, G->st, S.b,
                       (uint64_t)0, 1, S.b, S.n, S.W.part.p, S.W.cnt.p,
                       S.W.nrm);
    H2HIP_CHECK(hipMemcpyAsync(&nb2[2 * s], S.W.nrm, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, G->st));
  }
  for (uint32_t s = 0; s < count; s++) {
    int rc = multi_init<C>(G, G->sys[s], G->st);
    if (rc) return rc;
  }
  std::vector<double> errs(count * stride);
  H2HIP_CHECK(hipMemcpyAsync(errs.data(), G->d_err.p,
                             sizeof(double) * errs.size(),
                             hipMemcpyDeviceToHost, G->st));
  H2HIP_CHECK(hipStreamSynchronize(G->st));
  H2HIP_CHECK(hipGetLastError());
  std::vector<uint32_t> act;
  for (uint32_t s = 0; s < count; s++) {
    auto &S = G->sys[s];
    S.normb = std::sqrt(nb2[2 * s]);
    S.k = 0;
    S.iter = 0;
    S.error = errs[s * stride + kmax];
    // the check before the first step (Library/krylovsolvers.c:183-227)
    S.active = S.error > eps * S.normb && S.iter + 1 != maxiter;
    if (S.active) act.push_back(s);
  }
  int k = 0;
  uint32_t iter = 0;
  while (!act.empty()) {
    if (k + 1 >= kmax) {
      for (uint32_t s : act) {
        int rc = multi_finish<C>(G->sys[s], k, G->st);
        if (!rc) rc = multi_init<C>(G, G->sys[s], G->st);
        if (rc) return rc;
      }
      k = 0;
    }
    uint32_t len = (uint32_t)(kmax - 1 - k);
    if (len > (uint32_t)CHUNK) len = CHUNK;
    if (maxiter != 0 && len > maxiter - 1 - iter) len = maxiter - 1 - iter;
    if (use_graph && G->exec_set != act) {
      int rc = multi_capture<C>(G, act);
      if (rc) return rc;
    }
    for (uint32_t j = 0; j < len; j++) {
      if (use_graph) {
        H2HIP_CHECK(hipGraphLaunch(G->exec, G->st));
      } else {
        for (uint32_t s : act) {
          int rc = multi_step<C>(G->sys[s], G->st);
          if (rc) return rc;
        }
      }
    }
    H2HIP_CHECK(hipGetLastError());
    H2HIP_CHECK(hipMemcpyAsync(errs.data(), G->d_err.p,
                               sizeof(double) * errs.size(),
                               hipMemcpyDeviceToHost, G->st));
    H2HIP_CHECK(hipStreamSynchronize(G->st));
    // replay the reference's loop per system over the steps just done
    std::vector<uint32_t> still;
    for (uint32_t s : act) {
      auto &S = G->sys[s];
      bool done = false;
      for (uint32_t j = 0; j < len && !done; j++) {
        S.k = k + (int)j + 1;
        S.iter = iter + j + 1;
        S.error = errs[s * stride + k + j];
        done = !(S.error > eps * S.normb && S.iter + 1 != maxiter);
      }
      if (done) {
        int rc = multi_finish<C>(S, S.k, G->st);
        if (rc) return rc;
      } else {
        still.push_back(s);
      }
    }
    act.swap(still);
    k += (int)len;
    iter += len;
  }
  H2HIP_CHECK(hipGetLastError());
  H2HIP_CHECK(hipStreamSynchronize(G->st));
  for (uint32_t s = 0; s < count; s++) {
    if (iters) iters[s] = G->sys[s].iter;
    if (resid) resid[s] = G->sys[s].error;
  }
  return H2HIP_OK;
}

}  // namespace

extern "C" {

int h2hip_gmres_multi_create(h2hip_gmres_multi **out, uint32_t count,
                             const h2hip_h2matrix *const *H,
                             const uint64_t *n, int is_complex, uint32_t kmax,
                             void *stream) {
  if (!out || count == 0 || !H || !n || !stream || kmax < 2 || kmax > MAXK) {
    h2hip::set_error("h2hip_gmres_multi_create: invalid arguments (count > "
                     "0, a non-default stream, kmax 2..64)");
    return H2HIP_EINVAL;
  }
  for (uint32_t s = 0; s < count; s++)
    if (!H[s] || n[s] == 0) {
      h2hip::set_error("h2hip_gmres_multi_create: missing matrix or size");
      return H2HIP_EINVAL;
    }
  Multi *G = new Multi;
  G->C = is_complex ? 1 : 0;
  G->kmax = (int)kmax;
  G->st = (hipStream_t)stream;
  for (uint32_t s = 0; s < count; s++) G->sys.emplace_back();
  bool ok = G->d_k.alloc(count) == hipSuccess &&
            G->d_err.alloc((uint64_t)count * (kmax + 1)) == hipSuccess &&
            hipEventCreateWithFlags(&G->fork, hipEventDisableTiming) ==
                hipSuccess;
  for (uint32_t s = 0; s < count && ok; s++) {
    auto &S = G->sys[s];
    S.H = H[s];
    S.n = n[s];
    S.n2 = (G->C ? 2 : 1) * n[s];
    S.dk = G->d_k.p + s;
    S.derr = G->d_err.p + (uint64_t)s * (kmax + 1);
    ok = S.W.alloc(G->C, S.n, (int)kmax, G->st) == H2HIP_OK &&
         S.xb.alloc(S.n2) == hipSuccess;
  }
  if (!ok) {
    delete G;
    h2hip::set_error("h2hip_gmres_multi_create: device allocation failed");
    return H2HIP_ENOMEM;
  }
  *out = G;
  return H2HIP_OK;
}

void h2hip_gmres_multi_destroy(h2hip_gmres_multi *G) { delete G; }

int h2hip_gmres_multi_solve(h2hip_gmres_multi *G, const void *const *d_b,
                            void *const *d_x, double eps, uint32_t maxiter,
                            int use_graph, uint32_t *iterations,
                            double *residuals) {
  if (!G || !d_b || !d_x) return H2HIP_EINVAL;
  for (size_t s = 0; s < G->sys.size(); s++) {
    if (!d_b[s] || !d_x[s]) return H2HIP_EINVAL;
    G->sys[s].b = (const double *)d_b[s];
    G->sys[s].x = (double *)d_x[s];
  }
  return G->C ? multi_solve<1>(G, eps, maxiter, use_graph, iterations,
                               residuals)
              : multi_solve<0>(G, eps, maxiter, use_graph, iterations,
                               residuals);
}

uint64_t h2hip_gmres_multi_graph_builds(const h2hip_gmres_multi *G) {
  return G ? G->graph_builds : 0;
}

}  // extern "C"

extern "C" int h2hip_solve_gmres(h2hip_addeval_t addeval, void *A,
                                 int is_complex, uint64_t n, const void *d_b,
                                 void *d_x, double eps, uint32_t maxiter,
                                 uint32_t kmax, void *stream,
                                 uint32_t *iterations, double *residual) {
  if (!addeval || !d_b || !d_x || n == 0 || kmax < 2 || kmax > MAXK) {
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

extern "C" int h2hip_solve_cg(h2hip_addeval_t addeval, void *A, int is_complex,
                              uint64_t n, const void *d_b, void *d_x,
                              double eps, uint32_t maxiter, int use_graph,
                              void *stream, uint32_t *iterations,
                              double *residual) {
  if (!addeval || !d_b || !d_x || n == 0) {
    h2hip::set_error("h2hip_solve_cg: invalid arguments");
    return H2HIP_EINVAL;
  }
  return is_complex
             ? cg<1>(addeval, A, n, (const double *)d_b, (double *)d_x, eps,
                     maxiter, use_graph, (hipStream_t)stream, iterations,
                     residual)
             : cg<0>(addeval, A, n, (const double *)d_b, (double *)d_x, eps,
                     maxiter, use_graph, (hipStream_t)stream, iterations,
                     residual);
}

extern "C" int h2hip_solve_cg_noalloc(h2hip_addeval_t addeval, void *A,
                                      int is_complex, uint64_t n,
                                      const void *d_b, void *d_x, double eps,
                                      uint32_t maxiter, void *stream,
                                      uint32_t *iterations,
                                      double *residual) {
  if (!addeval || !d_b || !d_x || n == 0) {
    h2hip::set_error("h2hip_solve_cg_noalloc: invalid arguments");
    return H2HIP_EINVAL;
  }
  return is_complex
             ? cg_noalloc<1>(addeval, A, n, (const double *)d_b,
                             (double *)d_x, eps, maxiter, (hipStream_t)stream,
                             iterations, residual)
             : cg_noalloc<0>(addeval, A, n, (const double *)d_b,
                             (double *)d_x, eps, maxiter, (hipStream_t)stream,
                             iterations, residual);
}
