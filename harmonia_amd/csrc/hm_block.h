// Wave64 / workgroup scan helpers shared by the scan and radix-sort kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hm {

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
