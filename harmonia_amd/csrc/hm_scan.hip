// Device exclusive scan (reduce-then-scan).  Used for the prefix-sum child-index array, the
// radix-sort digit offsets, range output offsets and the leaf re-layout after updates.
// The prefix array semantics follow H/regular_bpt/traversal.cpp:40,58 and H/src/traversal.cpp:180-195
// (prefix[0] = 1, prefix[k+1] = prefix[k] + children_k), computed here as 1 + exclusive scan.
#include "hm_block.h"
#include "hm_internal.h"

namespace hm {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;

template <typename Tin, typename T>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const Tin* __restrict__ in, size_t n,
                                                               T* __restrict__ partials) {
  __shared__ T sh[kWaves + 1];
  const size_t base = (size_t)blockIdx.x * kTile;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    size_t idx = base + (size_t)i * kThreads + threadIdx.x;
    if (idx < n) s += (T)in[idx];
  }
  T total;
  (void)block_exclusive_scan<T, kThreads>(s, sh, total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Single block: exclusive scan of the per-tile partials, offset by init.
template <typename T>
__global__ __launch_bounds__(kThreads) void scan_partials_kernel(T* __restrict__ partials, size_t m, T init) {
  __shared__ T sh[kWaves + 1];
  T carry = init;
  for (size_t base = 0; base < m; base += kThreads) {
    size_t idx = base + threadIdx.x;
    T v = idx < m ? partials[idx] : (T)0;
    T total;
    T ex = block_exclusive_scan<T, kThreads>(v, sh, total);
    if (idx < m) partials[idx] = ex + carry;
    carry += total;
  }
}

template <typename Tin, typename T>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(const Tin* __restrict__ in, T* __restrict__ out,
                                                             size_t n, const T* __restrict__ partials) {
  __shared__ T sh[kWaves + 1];
  const size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
  T v[kItems];
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    size_t idx = base + i;
    v[i] = idx < n ? (T)in[idx] : (T)0;
    s += v[i];
  }
  T total;
  T ex = block_exclusive_scan<T, kThreads>(s, sh, total) + partials[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    size_t idx = base + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
}

// Small inputs (n <= kSingleMax): one block of 1024 threads, each scanning a contiguous chunk --
// one launch instead of three (the radix sorts of update batches and range batches scan a few
// thousand counters at a time, where launches dominate).
constexpr int kSingleThreads = 1024;
constexpr size_t kSingleMax = 16384;

template <typename Tin, typename T>
__global__ __launch_bounds__(kSingleThreads) void scan_single_kernel(const Tin* __restrict__ in, T* __restrict__ out,
                                                                     size_t n, T init) {
  __shared__ T sh[kSingleThreads / 64 + 1];
  const size_t chunk = (n + kSingleThreads - 1) / kSingleThreads;
  const size_t b = (size_t)threadIdx.x * chunk;
  const size_t e = b + chunk < n ? b + chunk : n;
  T s = 0;
  for (size_t i = b; i < e; ++i) s += (T)in[i];
  T total;
  T run = block_exclusive_scan<T, kSingleThreads>(s, sh, total) + init;
  for (size_t i = b; i < e; ++i) {
    const T v = (T)in[i];  // read before the write: in and out may be the same array
    out[i] = run;
    run += v;
  }
}

template <typename Tin, typename T>
hipError_t scan_impl(const Tin* in, T* out, size_t n, T init, void* tmp, hipStream_t s) {
  if (n == 0) return hipSuccess;
  ProfScope prof("scan", s);
  if (n <= kSingleMax) {
    hipLaunchKernelGGL((scan_single_kernel<Tin, T>), dim3(1), dim3(kSingleThreads), 0, s, in, out, n, init);
    return hipGetLastError();
  }
  const size_t nb = (n + kTile - 1) / kTile;
  T* partials = reinterpret_cast<T*>(tmp);
  hipLaunchKernelGGL((scan_reduce_kernel<Tin, T>), dim3((unsigned)nb), dim3(kThreads), 0, s, in, n, partials);
  hipLaunchKernelGGL((scan_partials_kernel<T>), dim3(1), dim3(kThreads), 0, s, partials, nb, init);
  hipLaunchKernelGGL((scan_down_kernel<Tin, T>), dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n,
                     (const T*)partials);
  return hipGetLastError();
}

}  // namespace

size_t scan_tmp_bytes(size_t n) { return ((n + kTile - 1) / kTile + 1) * sizeof(uint64_t) + 256; }

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t init, void* tmp,
                              hipStream_t s) {
  return scan_impl<uint32_t, uint32_t>(in, out, n, init, tmp, s);
}
hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t init, void* tmp,
                              hipStream_t s) {
  return scan_impl<uint64_t, uint64_t>(in, out, n, init, tmp, s);
}
hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, void* tmp, hipStream_t s) {
  return scan_impl<uint32_t, uint64_t>(in, out, n, (uint64_t)0, tmp, s);
}

}  // namespace hm
