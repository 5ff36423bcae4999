// Wave64 / workgroup scan helpers shared by the scan and radix-sort kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hm {

// Make this wave's LDS writes visible to its own later LDS reads (and keep the compiler from moving
// LDS accesses across): the LDS executes one wave's operations in order, so no block barrier is
// needed for a wave-private buffer.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lanes (among `valid`) of this wave whose 8-bit digit equals this lane's digit: eight ballots,
// one per digit bit.  Every lane of the wave must execute it.
// Per bit: s = 0 or all ones (the bit sign-extended), m = the wave's ballot of that bit, and the lane
// keeps the peers whose bit agrees with its own: peers &= ~(m ^ s), one 3-input logic op per half.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint32_t lo = (uint32_t)valid, hi = (uint32_t)(valid >> 32);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t s = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
    const uint64_t m = __ballot(s != 0u);
    lo &= ~((uint32_t)m ^ s);
    hi &= ~((uint32_t)(m >> 32) ^ s);
  }
  return ((uint64_t)hi << 32) | lo;
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
