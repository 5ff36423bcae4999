// Comparator baselines: point search over a pointer-child tree.
//
// The reference's comparators keep an explicit child pointer per slot instead of Harmonia's
// prefix-sum array: the regular tree's CFB layout with one child offset array per node
// (cArr_t, H/regular_bpt/cfb.h:16-20) searched by multi_noprefix_nosort_kernel
// (H/regular_bpt/noprefix-nosort.cu:89-237: narrowed groups scan the node's keys granule by granule
// with an early exit, then child[slot]), and the HB+-tree kernel bpt_search_kernel
// (H/src/gbpt.cu:34-113: index keys, then key group, then child[idx]).  Both are rebuilt here over
// the same key region and leaves as the Harmonia search, plus a u32 child array
// child[node * F + slot], so a comparison isolates the layout: the next node comes from a
// dependent read of child[] (one more cache line per level) instead of prefix[node] + slot.
//
// The kernels are deliberately plain: one query per group, no LDS-resident levels, no PSA, no
// persistent grid, results in input order.  The same plain kernel over the prefix-sum array
// (KIND_PREFIX) is the control that isolates the layout from the production kernel's other changes.
#include "hm_device.h"

namespace hm {
namespace {

constexpr int kThreads = 256;

__global__ void child_array_kernel(const uint32_t* __restrict__ prefix, uint32_t n_inner, int F,
                                   uint32_t* __restrict__ child) {
  const size_t total = (size_t)n_inner * F;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t node = (uint32_t)(i / F);
    child[i] = prefix[node] + (uint32_t)(i % F);
  }
}

enum { KIND_POINTER = 0, KIND_LINEAR = 1, KIND_PREFIX = 2 };

// KIND_LINEAR: scan the node's keys one 64-byte granule (8 keys) per step and stop at the first
// granule holding a key > t (the noprefix-nosort loop, H/regular_bpt/noprefix-nosort.cu:139-162);
// otherwise rank through the index granule first (the HB+ / Harmonia node, H/src/gbpt.cu:59-111).
// KIND_PREFIX takes the next node from prefix[node] + slot instead of child[].
template <int F, int G, int KIND>
__global__ __launch_bounds__(kThreads) void search_ptr_kernel(TreeView tv, const uint32_t* __restrict__ child,
                                                              const uint64_t* __restrict__ q,
                                                              uint64_t* __restrict__ out_vals,
                                                              uint8_t* __restrict__ out_found, uint32_t n) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  const int gl = threadIdx.x % G;
  const int gbase = (threadIdx.x & 63) - gl;
  const uint32_t qi = (uint32_t)(((size_t)blockIdx.x * kThreads + threadIdx.x) / G);
  // groups past the end repeat the last query (every lane of a wave takes part in the ballots) and
  // write nothing
  const uint64_t t = q[qi < n ? qi : n - 1];
  const uint64_t td = t == KEY_MAX ? KEY_MAX - 1 : t;
  uint32_t node = 0;
  for (int l = 0; l < tv.height; ++l) {
    const uint64_t* nd = tv.inner + (size_t)node * NS;
    uint32_t slot;
    if constexpr (KIND == KIND_LINEAR) {
      // slot F - 1 holds KEY_MAX > td, so the scan stops by the last granule at the latest; the
      // exit is uniform across a group (its lanes share the count)
      slot = 0;
      for (int s = 0; s < F / 8; ++s) {
        uint64_t k[W];
        load_keys<W>(nd + IS + 8 * s + gl * W, k);
        const int c = group_count_le<G, W>(k, td, gbase);
        slot += (uint32_t)c;
        if (c < 8) break;
      }
    } else {
      slot = node_rank<F, G>(nd, td, gl, gbase);
    }
    if constexpr (KIND == KIND_PREFIX)
      node = tv.prefix[node] + slot;
    else
      node = child[(size_t)node * F + slot];
  }
  const Leaf* lf = tv.leaves + (node - tv.n_inner);
  uint64_t k[W], v[W];
  load_keys<W>(lf->key + gl * W, k);
  if constexpr (G > 1) load_keys<W>(lf->val + gl * W, v);
  const int lt = group_count_lt<G, W>(k, t, gbase);
  const bool eq = group_any_eq<G, W>(k, t, gbase) && t != KEY_MAX;
  const int writer = eq ? lt / W : 0;
  if (qi < n && gl == writer) {
    uint64_t val = 0ull;
    if constexpr (G > 1) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (eq && lt % W == w) val = v[w];
    } else {
      if (eq) val = lf->val[lt];
    }
    out_vals[qi] = val;
    out_found[qi] = eq ? 1 : 0;
  }
}

template <int F, int KIND>
hipError_t launch_ptr_f(const TreeView& tv, const uint32_t* child, const uint64_t* q, uint64_t* out_vals,
                        uint8_t* out_found, uint32_t n, int group, hipStream_t s) {
  const size_t blocks = ((size_t)n * group + kThreads - 1) / kThreads;
  const dim3 grid((unsigned)blocks), block(kThreads);
  switch (group) {
    case 1: hipLaunchKernelGGL((search_ptr_kernel<F, 1, KIND>), grid, block, 0, s, tv, child, q, out_vals, out_found, n); break;
    case 2: hipLaunchKernelGGL((search_ptr_kernel<F, 2, KIND>), grid, block, 0, s, tv, child, q, out_vals, out_found, n); break;
    case 4: hipLaunchKernelGGL((search_ptr_kernel<F, 4, KIND>), grid, block, 0, s, tv, child, q, out_vals, out_found, n); break;
    case 8: hipLaunchKernelGGL((search_ptr_kernel<F, 8, KIND>), grid, block, 0, s, tv, child, q, out_vals, out_found, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int KIND>
hipError_t launch_ptr(const TreeView& tv, const uint32_t* child, const uint64_t* q, uint64_t* out_vals,
                      uint8_t* out_found, uint32_t n, int group, hipStream_t s) {
  switch (tv.fanout) {
    case 32: return launch_ptr_f<32, KIND>(tv, child, q, out_vals, out_found, n, group, s);
    case 64: return launch_ptr_f<64, KIND>(tv, child, q, out_vals, out_found, n, group, s);
    case 128: return launch_ptr_f<128, KIND>(tv, child, q, out_vals, out_found, n, group, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t build_child_array(const uint32_t* prefix, uint32_t n_inner, int F, uint32_t* child, hipStream_t s) {
  if (!n_inner) return hipSuccess;
  const size_t total = (size_t)n_inner * F;
  size_t nb = (total + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(child_array_kernel, dim3((unsigned)nb), dim3(256), 0, s, prefix, n_inner, F, child);
  return hipGetLastError();
}

hipError_t launch_search_baseline(const TreeView& tv, const uint32_t* child, const uint64_t* q, uint64_t* out_vals,
                                  uint8_t* out_found, uint32_t n, int group, int kind, hipStream_t s) {
  if (!n) return hipSuccess;
  switch (kind) {
    case KIND_POINTER: return launch_ptr<KIND_POINTER>(tv, child, q, out_vals, out_found, n, group, s);
    case KIND_LINEAR: return launch_ptr<KIND_LINEAR>(tv, child, q, out_vals, out_found, n, group, s);
    case KIND_PREFIX: return launch_ptr<KIND_PREFIX>(tv, child, q, out_vals, out_found, n, group, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace hm
