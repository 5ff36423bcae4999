// Leaf-chain range scan.
//
// Replaces range_prefix_sort_1thread_search_1_thread_scan_kernel (H/src/range.cu:101-227): descend
// to the leaf of `start`, then walk the leaves in key order (the leaf pool is contiguous in key
// order, leaf i -> i + 1, as the reference's BLeaf pool, H/src/range.cu:188-221), skipping padding.
// The reference emits at most Range_Size = 8 records per range (H/src/range.cu:32); here a range
// emits up to its own length, or every key of a closed interval [lo, hi] as
// find_range(start, end) does (H/regular_bpt/regular_bpt.cpp:320-338).
// One group of 8 lanes per range: lane s owns slot s of each leaf, and a ballot over the group
// packs the qualifying keys into the output in order.  The walk loads up to kWalk consecutive leaves
// per step (their addresses are known: i, i + 1, ...), so each group keeps several loads in flight
// instead of one dependent load per leaf; the leaves are then consumed in order exactly as one at a
// time.  A step loads no more leaves than the range can still use (ceil(rest / 8) + 1: a leaf holds
// at most 8 keys, and the first one may hold keys below start).  The leaves go to LDS by LDS-direct
// loads, one 16-byte piece per lane (a group's 8 lanes bring its whole 128-B leaf in one instruction,
// where an 8-byte key load and an 8-byte value load per lane took two), and the lanes read their key
// and value from there: 8 KiB of LDS per wave.
// Ranges run in input order: consecutive groups then write consecutive output regions.  (Walking
// them in PSA order of their starts, as the reference sorts range starts, H/src/range.cu:280-285,
// was measured: the leaf reads gain locality, the output writes lose it, and the kernel time did
// not change.)
#include "hm_device.h"

namespace hm {
namespace {

constexpr int kThreads = 256;
constexpr int G = 8;
constexpr int kWalk = 8;

template <int F>
__global__ __launch_bounds__(kThreads) void range_kernel(TreeView tv, const uint64_t* __restrict__ start,
                                                         const uint64_t* __restrict__ hi,
                                                         const uint32_t* __restrict__ len, uint32_t cap_const,
                                                         uint32_t n, const uint64_t* __restrict__ offsets,
                                                         uint64_t offset_stride, uint64_t* __restrict__ out_keys,
                                                         uint64_t* __restrict__ out_vals,
                                                         uint32_t* __restrict__ out_counts, uint64_t capacity) {
  const int gl = threadIdx.x % G;
  const int gbase = (threadIdx.x & 63) - gl;
  const uint32_t gid = threadIdx.x / G;
  constexpr uint32_t gpb = kThreads / G;
  const uint64_t lanes_below = (1ull << gl) - 1ull;
  typedef __attribute__((address_space(1))) void global_void_t;
  typedef __attribute__((address_space(3))) void lds_void_t;
  // the wave's leaves of one step: step u holds its 8 groups' leaves, lane l's 16-byte piece at byte 16 l
  __shared__ __align__(16) uint64_t lbuf[kThreads / 64][kWalk][128];
  uint64_t* wb = &lbuf[threadIdx.x >> 6][0][0];
  const int glane = threadIdx.x & 63;
  for (uint32_t rb = blockIdx.x * gpb; rb < n; rb += gridDim.x * gpb) {  // uniform across the block
    const uint32_t ri = rb + gid;
    const bool valid = ri < n;
    const uint64_t s = valid ? start[ri] : 0ull;
    const uint64_t h = (valid && hi) ? hi[ri] : KEY_MAX;
    const uint32_t L = valid ? (len ? len[ri] : cap_const) : 0u;
    const uint64_t off = valid ? (offsets ? offsets[ri] : (uint64_t)ri * offset_stride) : 0ull;
    const uint64_t sd = s == KEY_MAX ? KEY_MAX - 1 : s;
    uint32_t leaf = descend<F, G>(tv, 0, nullptr, sd, gl, gbase);
    uint32_t got = 0;
    bool done = !valid || L == 0 || s > h || s == KEY_MAX;
    while (!done) {  // uniform inside the group
      uint64_t k[kWalk], v[kWalk];
      const uint32_t need = (L - got + LEAF_KEYS - 1) / LEAF_KEYS + 1;  // got < L here
      const uint32_t step = need < (uint32_t)kWalk ? need : (uint32_t)kWalk;
#pragma unroll
      for (int u = 0; u < kWalk; ++u) {
        if ((uint32_t)u < step) {  // group-uniform; the load writes LDS only
          const uint32_t lf = leaf + u < tv.n_leaves ? leaf + u : tv.n_leaves - 1;
          __builtin_amdgcn_global_load_lds((const global_void_t*)(reinterpret_cast<const ulonglong2*>(tv.leaves + lf) + gl),
                                           (lds_void_t*)(wb + u * 128), 16, 0, 0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's leaves are in LDS
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int u = 0; u < kWalk; ++u) {
        const bool in = (uint32_t)u < step && leaf + u < tv.n_leaves;
        const uint64_t* lp = wb + u * 128 + (glane - gl) * 2;  // this group's leaf
        k[u] = in ? lp[gl] : KEY_MAX;
        v[u] = in ? lp[8 + gl] : 0ull;
      }
      __builtin_amdgcn_wave_barrier();  // the next step's loads overwrite the buffer after these reads
#pragma unroll
      for (int u = 0; u < kWalk; ++u) {
        if (!done && (uint32_t)u < step) {  // group-uniform
          const bool take = k[u] != KEY_MAX && k[u] >= s && k[u] <= h;
          const uint64_t m = (__ballot(take) >> gbase) & 0xFFull;
          const uint64_t beyond = (__ballot(k[u] != KEY_MAX && k[u] > h) >> gbase) & 0xFFull;
          const uint32_t rank = (uint32_t)__popcll(m & lanes_below);
          if (take && got + rank < L) {
            const uint64_t o = off + got + rank;
            if (o < capacity) {
              out_keys[o] = k[u];
              out_vals[o] = v[u];
            }
          }
          got += (uint32_t)__popcll(m);
          done = got >= L || beyond != 0 || leaf + u + 1 >= tv.n_leaves;
        }
      }
      leaf += step;
    }
    if (valid && gl == 0) out_counts[ri] = got < L ? got : L;
  }
}

__global__ void offsets_tail_kernel(const uint32_t* len, uint64_t* offsets, uint32_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) offsets[n] = n ? offsets[n - 1] + len[n - 1] : 0ull;
}

}  // namespace

hipError_t launch_range(const TreeView& tv, const uint64_t* start, const uint64_t* hi, const uint32_t* len,
                        uint32_t cap_const, size_t n, const uint64_t* offsets, uint64_t offset_stride,
                        uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  ProfScope prof("range", s);
  constexpr uint32_t gpb = kThreads / G;
  size_t nb = (n + gpb - 1) / gpb;
  const size_t max_nb = (size_t)launch_cfg().num_cus * 8;
  if (nb > max_nb) nb = max_nb;
  switch (tv.fanout) {
    case 32:
      hipLaunchKernelGGL(range_kernel<32>, dim3((unsigned)nb), dim3(kThreads), 0, s, tv, start, hi, len, cap_const,
                         (uint32_t)n, offsets, offset_stride, out_keys, out_vals, out_counts, capacity);
      break;
    case 64:
      hipLaunchKernelGGL(range_kernel<64>, dim3((unsigned)nb), dim3(kThreads), 0, s, tv, start, hi, len, cap_const,
                         (uint32_t)n, offsets, offset_stride, out_keys, out_vals, out_counts, capacity);
      break;
    case 128:
      hipLaunchKernelGGL(range_kernel<128>, dim3((unsigned)nb), dim3(kThreads), 0, s, tv, start, hi, len, cap_const,
                         (uint32_t)n, offsets, offset_stride, out_keys, out_vals, out_counts, capacity);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t range_offsets_tail(const uint32_t* len, uint64_t* offsets, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(offsets_tail_kernel, dim3(1), dim3(64), 0, s, len, offsets, (uint32_t)n);
  return hipGetLastError();
}

}  // namespace hm
