// Wave64 / workgroup scan helpers shared by the scan and radix-sort kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hm {

// Lanes (among `valid`) of this wave whose 8-bit digit equals this lane's digit: eight ballots,
// one per digit bit.  Every lane of the wave must execute it.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t peers = valid;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread across a block of NT threads; returns the block total
// in `total`.  `sh` must hold NT/64 + 1 elements.  Ends with a barrier so `sh` can be reused.
template <typename T, int NT>
__device__ __forceinline__ T block_exclusive_scan(T v, T* sh, T& total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(v);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int i = 0; i < NW; ++i) {
      T t = sh[i];
      sh[i] = run;
      run += t;
    }
    sh[NW] = run;
  }
  __syncthreads();
  total = sh[NW];
  T r = inc - v + sh[w];
  __syncthreads();
  return r;
}

}  // namespace hm
