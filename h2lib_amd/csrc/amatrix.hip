// Dense products y += alpha A x and y += alpha A^* x on gfx950
// (addeval_amatrix_avector / addevaltrans_amatrix_avector,
// Library/amatrix.c:862-925), column-major A with leading dimension ld, real
// or complex (interleaved pairs).
//
// A x: one thread per row, so every column step of a wave reads 64
// consecutive entries of A.  The columns are split into chunks, one
// workgroup per (256 rows, chunk), so that a matrix of a few thousand rows
// still fills the device; each chunk leaves its partial sums in a scratch
// array and a second launch adds them in chunk order (results do not depend
// on the schedule).  With one chunk the partial sums go to y directly.
// A^* x: one wave per column (a contiguous strip of A), lanes over the
// rows, the wave's sums reduced by a fixed butterfly.
// Both are bound by HBM bandwidth: each entry of A is read once and used in
// one fused multiply-add (two for complex).
#include <algorithm>

#include "h2hip_internal.h"

namespace {

constexpr int NT = 256;
constexpr uint32_t TARGET_BLOCKS = 1024;  // 4 workgroups per CU
constexpr uint32_t MIN_CHUNK = 32;        // columns per chunk at least

struct GemvArgs {
  const double *A;
  uint64_t ld;
  uint32_t rows, cols;
  const double *x;
  double *y;
  double *part;  // nchunks x rows partial sums (x 2 for complex)
  uint32_t cchunk, nchunks;
  double are, aim;
};

template <int C>
__device__ __forceinline__ void madd(const double *A, uint64_t e,
                                     const double *x, uint32_t j, double &sr,
                                     double &si) {
  if constexpr (C) {
    const double ar = A[2 * e], ai = A[2 * e + 1];
    const double xr = x[2 * j], xi = x[2 * j + 1];
    sr = fma(ar, xr, fma(-ai, xi, sr));
    si = fma(ar, xi, fma(ai, xr, si));
  } else {
    sr = fma(A[e], x[j], sr);
  }
}

template <int C>
__device__ __forceinline__ void add_scaled(double *y, uint64_t i, double sr,
                                           double si, double are,
                                           double aim) {
  if constexpr (C) {
    y[2 * i] += are * sr - aim * si;
    y[2 * i + 1] += are * si + aim * sr;
  } else {
    y[i] += are * sr;
  }
}

template <int C>
__global__ __launch_bounds__(NT) void gemv_part(GemvArgs a) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.rows) return;
  const uint32_t j0 = blockIdx.y * a.cchunk;
  const uint32_t j1 = min(a.cols, j0 + a.cchunk);
  // four independent sums, eight loads in flight per lane
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t j = j0;
  for (; j + 8 <= j1; j += 8) {
#pragma unroll
    for (int u = 0; u < 8; u++)
      madd<C>(a.A, i + (uint64_t)(j + u) * a.ld, a.x, j + u, sr[u & 3],
              si[u & 3]);
  }
  for (; j < j1; j++) madd<C>(a.A, i + (uint64_t)j * a.ld, a.x, j, sr[0], si[0]);
  const double tr = (sr[0] + sr[1]) + (sr[2] + sr[3]);
  const double ti = (si[0] + si[1]) + (si[2] + si[3]);
  if (a.nchunks == 1) {
    add_scaled<C>(a.y, i, tr, ti, a.are, a.aim);
    return;
  }
  const uint64_t o = (uint64_t)blockIdx.y * a.rows + i;
  if constexpr (C) {
    a.part[2 * o] = tr;
    a.part[2 * o + 1] = ti;
  } else {
    a.part[o] = tr;
  }
}

template <int C>
__global__ __launch_bounds__(NT) void gemv_reduce(GemvArgs a) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.rows) return;
  double sr = 0.0, si = 0.0;
  for (uint32_t c = 0; c < a.nchunks; c++) {
    const uint64_t o = (uint64_t)c * a.rows + i;
    if constexpr (C) {
      sr += a.part[2 * o];
      si += a.part[2 * o + 1];
    } else {
      sr += a.part[o];
    }
  }
  add_scaled<C>(a.y, i, sr, si, a.are, a.aim);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// y_j += alpha sum_i conj(A_ij) x_i, one wave per column
template <int C>
__global__ __launch_bounds__(NT) void gemv_trans(GemvArgs a) {
  const uint32_t j = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (j >= a.cols) return;  // uniform over the wave
  const double *col = a.A + (C ? 2 : 1) * (uint64_t)j * a.ld;
  double sr[2] = {0.0, 0.0}, si[2] = {0.0, 0.0};
  uint32_t i = lane, u = 0;
  for (; i < a.rows; i += 64, u ^= 1) {
    if constexpr (C) {
      const double ar = col[2 * i], ai = col[2 * i + 1];
      const double xr = a.x[2 * i], xi = a.x[2 * i + 1];
      // conj(a) x
      sr[u] = fma(ar, xr, fma(ai, xi, sr[u]));
      si[u] = fma(ar, xi, fma(-ai, xr, si[u]));
    } else {
      sr[u] = fma(col[i], a.x[i], sr[u]);
    }
  }
  const double tr = wave_sum(sr[0] + sr[1]);
  const double ti = C ? wave_sum(si[0] + si[1]) : 0.0;
  if (lane == 0) add_scaled<C>(a.y, j, tr, ti, a.are, a.aim);
}

// Column chunks of the product A x: enough workgroups to fill the device,
// at least MIN_CHUNK columns each.
void plan_chunks(uint32_t rows, uint32_t cols, uint32_t &cchunk,
                 uint32_t &nchunks) {
  const uint32_t rblk = (rows + NT - 1) / NT;
  uint32_t want = std::max<uint32_t>(1, TARGET_BLOCKS / std::max(1u, rblk));
  want = std::min(want, std::max<uint32_t>(1, (cols + MIN_CHUNK - 1) /
                                                  MIN_CHUNK));
  cchunk = (cols + want - 1) / want;
  nchunks = (cols + cchunk - 1) / cchunk;
}

template <int C>
void launch_gemv(const GemvArgs &a, bool trans, hipStream_t st) {
  if (trans) {
    hipLaunchKernelGGL(gemv_trans<C>, dim3((a.cols + 3) / 4), dim3(NT), 0,
                       st, a);
    return;
  }
  const uint32_t rblk = (a.rows + NT - 1) / NT;
  hipLaunchKernelGGL(gemv_part<C>, dim3(rblk, a.nchunks), dim3(NT), 0, st,
                     a);
  if (a.nchunks > 1)
    hipLaunchKernelGGL(gemv_reduce<C>, dim3(rblk), dim3(NT), 0, st, a);
}

}  // namespace

using h2hip::set_error;

extern "C" int h2hip_addeval_amatrix(const double *alpha, int is_complex,
                                     int trans, uint32_t rows, uint32_t cols,
                                     const void *A, uint64_t ld,
                                     const void *x, void *y, void *stream) {
  if (!alpha || (rows && cols && (!A || !x || !y)) ||
      (cols > 0 && ld < rows)) {
    set_error("h2hip_addeval_amatrix: invalid arguments");
    return H2HIP_EINVAL;
  }
  if (rows == 0 || cols == 0) return H2HIP_OK;
  const hipStream_t st = (hipStream_t)stream;
  GemvArgs a;
  a.A = (const double *)A;
  a.ld = ld;
  a.rows = rows;
  a.cols = cols;
  a.x = (const double *)x;
  a.y = (double *)y;
  a.part = nullptr;
  a.are = alpha[0];
  a.aim = alpha[1];
  a.cchunk = cols;
  a.nchunks = 1;
  if (!trans) {
    plan_chunks(rows, cols, a.cchunk, a.nchunks);
    if (a.nchunks > 1) {
      // partial sums in stream-ordered device memory, released after the
      // reduction on the same stream
      const size_t bytes =
          sizeof(double) * (is_complex ? 2 : 1) * a.nchunks * (size_t)rows;
      if (hipMallocAsync((void **)&a.part, bytes, st) != hipSuccess) {
        set_error("h2hip_addeval_amatrix: device allocation failed");
        return H2HIP_ENOMEM;
      }
    }
  }
  if (is_complex)
    launch_gemv<1>(a, trans != 0, st);
  else
    launch_gemv<0>(a, trans != 0, st);
  const hipError_t le = hipGetLastError();
  if (a.part) H2HIP_CHECK(hipFreeAsync(a.part, st));
  H2HIP_CHECK(le);
  return H2HIP_OK;
}
