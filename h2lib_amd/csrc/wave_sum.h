// Wave-level sums shared by the Krylov and H2-product kernels (gfx950,
// 64-lane waves).
#pragma once

#include <hip/hip_runtime.h>

namespace h2hip {

// One butterfly step over 2 S values per lane: the lane keeps one half of
// v[0 .. 2 S) and adds its partner's (lane ^ S) copy of that half.
template <int S, int N>
__device__ __forceinline__ void bfly_step(double (&v)[N], int lane) {
  const bool up = (lane & S) != 0;
#pragma unroll
  for (int t = 0; t < S; t++) {
    const double send = up ? v[t] : v[t + S];
    const double keep = up ? v[t + S] : v[t];
    v[t] = keep + __shfl_xor(send, S, 64);
  }
}

// Sums of 32 values over the 64 lanes of a wave by a butterfly that halves
// the values a lane holds at each step: lane l returns the sum of value
// l & 31 over all lanes.  32 exchanges instead of 6 for each value.  v is
// clobbered.
__device__ __forceinline__ double wave_sum32(double (&v)[32], int lane) {
  bfly_step<16>(v, lane);
  bfly_step<8>(v, lane);
  bfly_step<4>(v, lane);
  bfly_step<2>(v, lane);
  bfly_step<1>(v, lane);
  return v[0] + __shfl_xor(v[0], 32, 64);
}

// The same for N = 1, 2, 4, .., 32 values: the butterfly over the values,
// then the sums over the lanes that hold the same one; lane l returns the
// sum of value l & (N - 1).
template <int N>
__device__ __forceinline__ double wave_sum_n(double (&v)[N], int lane) {
  if constexpr (N >= 32) bfly_step<16>(v, lane);
  if constexpr (N >= 16) bfly_step<8>(v, lane);
  if constexpr (N >= 8) bfly_step<4>(v, lane);
  if constexpr (N >= 4) bfly_step<2>(v, lane);
  if constexpr (N >= 2) bfly_step<1>(v, lane);
  double s = v[0];
#pragma unroll
  for (int off = N; off < 64; off <<= 1) s += __shfl_xor(s, off, 64);
  return s;
}

}  // namespace h2hip
