// Device-side building blocks of the descent, shared by the search, range and update kernels.
//
// A query is handled by a narrowed thread group (NTG, Harmonia paper p.140-141; the reference's
// Thread_Per_Query = DEFAULT_ORDER / Narrow_Rate lanes, H/regular_bpt/prefix-sort.cu:31) of G lanes
// inside a 64-lane wavefront.  The group reads one 64-byte granule (8 keys) per step, G lanes x W
// keys each, and ranks the target inside it with one ballot per key slot and a popcount of the
// group's bits of the 64-bit lane mask.  The reference does the same with __ballot_sync over
// 32-lane warps plus shared-memory flag arrays (H/regular_bpt/prefix-sort.cu:147-173).
#pragma once
#include "hm_internal.h"

namespace hm {

template <int W>
__device__ __forceinline__ void load_keys(const uint64_t* p, uint64_t (&k)[W]) {
  if constexpr (W == 1) {
    k[0] = p[0];
  } else {
    static_assert(W % 2 == 0, "W must be 1 or even");
#pragma unroll
    for (int i = 0; i < W / 2; ++i) {
      const ulonglong2 v = reinterpret_cast<const ulonglong2*>(p)[i];
      k[2 * i] = v.x;
      k[2 * i + 1] = v.y;
    }
  }
}

// Number of keys <= t among the group's G*W keys (all lanes of the group get the same count).
template <int G, int W>
__device__ __forceinline__ int group_count_le(const uint64_t (&k)[W], uint64_t t, int gbase) {
  int c = 0;
  if constexpr (G == 1) {
#pragma unroll
    for (int w = 0; w < W; ++w) c += (k[w] <= t) ? 1 : 0;
  } else {
    constexpr uint64_t gm = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll((__ballot(k[w] <= t) >> gbase) & gm);
  }
  return c;
}

template <int G, int W>
__device__ __forceinline__ int group_count_lt(const uint64_t (&k)[W], uint64_t t, int gbase) {
  int c = 0;
  if constexpr (G == 1) {
#pragma unroll
    for (int w = 0; w < W; ++w) c += (k[w] < t) ? 1 : 0;
  } else {
    constexpr uint64_t gm = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll((__ballot(k[w] < t) >> gbase) & gm);
  }
  return c;
}

template <int G, int W>
__device__ __forceinline__ bool group_any_eq(const uint64_t (&k)[W], uint64_t t, int gbase) {
  bool e = false;
  if constexpr (G == 1) {
#pragma unroll
    for (int w = 0; w < W; ++w) e |= (k[w] == t);
  } else {
    constexpr uint64_t gm = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    uint64_t m = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) m |= __ballot(k[w] == t);
    e = ((m >> gbase) & gm) != 0;
  }
  return e;
}

// upper_bound rank of t inside one inner node: the child slot (H/regular_bpt/cfb.cpp:15-19,
// H/regular_bpt/regular_bpt.cpp:361-365).  Two granule reads: the index granule picks the group
// of 8 keys (Harmonia's inner_index step, H/src/bpt.h:94-120), the group gives the slot.
template <int F, int G>
__device__ __forceinline__ uint32_t node_rank(const uint64_t* nd, uint64_t t, int gl, int gbase) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  uint64_t k[W];
  int j;
  if constexpr (F / 8 <= 8) {
    load_keys<W>(nd + gl * W, k);
    j = group_count_le<G, W>(k, t, gbase);
  } else {
    uint64_t k2[W];
    load_keys<W>(nd + gl * W, k);
    load_keys<W>(nd + 8 + gl * W, k2);
    j = group_count_le<G, W>(k, t, gbase) + group_count_le<G, W>(k2, t, gbase);
  }
  load_keys<W>(nd + IS + 8 * j + gl * W, k);
  return (uint32_t)(8 * j + group_count_le<G, W>(k, t, gbase));
}

// Descend from the root to a leaf id.  Levels [0, lds_levels) are read from LDS (nodes at
// `snodes`, prefix entries at `spre`), deeper levels from global memory.
// t must be < KEY_MAX (callers clamp), so every node holds a separator > t.
template <int F, int G>
__device__ __forceinline__ uint32_t descend(const TreeView& tv, int lds_levels, const uint64_t* snodes,
                                            const uint32_t* spre, uint64_t t, int gl, int gbase) {
  constexpr int NS = node_slots(F);
  uint32_t node = 0;
  int l = 0;
  for (; l < lds_levels; ++l) node = spre[node] + node_rank<F, G>(snodes + (size_t)node * NS, t, gl, gbase);
  for (; l < tv.height; ++l) {
    const uint32_t base = tv.prefix[node];
    node = base + node_rank<F, G>(tv.inner + (size_t)node * NS, t, gl, gbase);
  }
  return node - tv.n_inner;
}

// Q independent descents interleaved per group, so each lane keeps Q loads in flight per level
// (memory-level parallelism: random L2/MALL/HBM accesses are latency bound).
template <int F, int G, int Q>
__device__ __forceinline__ void descend_q(const TreeView& tv, int lds_levels, const uint64_t* snodes,
                                          const uint32_t* spre, const uint64_t (&t)[Q], uint32_t (&node)[Q], int gl,
                                          int gbase) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] = 0;
  int l = 0;
  for (; l < lds_levels; ++l) {
    uint64_t k[Q][W];
    int j[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) load_keys<W>(snodes + (size_t)node[q] * NS + gl * W, k[q]);
#pragma unroll
    for (int q = 0; q < Q; ++q) j[q] = group_count_le<G, W>(k[q], t[q], gbase);
    if constexpr (F / 8 > 8) {
#pragma unroll
      for (int q = 0; q < Q; ++q) load_keys<W>(snodes + (size_t)node[q] * NS + 8 + gl * W, k[q]);
#pragma unroll
      for (int q = 0; q < Q; ++q) j[q] += group_count_le<G, W>(k[q], t[q], gbase);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) load_keys<W>(snodes + (size_t)node[q] * NS + IS + 8 * j[q] + gl * W, k[q]);
#pragma unroll
    for (int q = 0; q < Q; ++q) node[q] = spre[node[q]] + 8 * j[q] + group_count_le<G, W>(k[q], t[q], gbase);
  }
  for (; l < tv.height; ++l) {
    uint64_t k[Q][W];
    uint64_t k2[Q][W];
    uint32_t base[Q];
    int j[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint64_t* nd = tv.inner + (size_t)node[q] * NS;
      base[q] = tv.prefix[node[q]];
      load_keys<W>(nd + gl * W, k[q]);
      if constexpr (F / 8 > 8) load_keys<W>(nd + 8 + gl * W, k2[q]);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      j[q] = group_count_le<G, W>(k[q], t[q], gbase);
      if constexpr (F / 8 > 8) j[q] += group_count_le<G, W>(k2[q], t[q], gbase);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) load_keys<W>(tv.inner + (size_t)node[q] * NS + IS + 8 * j[q] + gl * W, k[q]);
#pragma unroll
    for (int q = 0; q < Q; ++q) node[q] = base[q] + 8 * j[q] + group_count_le<G, W>(k[q], t[q], gbase);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] -= tv.n_inner;
}

// Stage the top `lds_levels` levels (nodes, then their prefix entries) into dynamic LDS.
template <int F>
__device__ __forceinline__ void stage_top_levels(const TreeView& tv, int lds_levels, uint64_t* smem,
                                                 uint32_t*& spre) {
  constexpr int NS = node_slots(F);
  const uint32_t nl = lds_levels > 0 ? tv.level_start[lds_levels] : 0u;
  const uint32_t words = nl * NS;
  for (uint32_t i = threadIdx.x; i < words / 2; i += blockDim.x)
    reinterpret_cast<ulonglong2*>(smem)[i] = reinterpret_cast<const ulonglong2*>(tv.inner)[i];
  spre = reinterpret_cast<uint32_t*>(smem + words);
  for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) spre[i] = tv.prefix[i];
  __syncthreads();
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous share of the
// virtual block range so neighbouring (PSA-sorted) queries meet in the same L2.  nb % 8 == 0.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) { return (b & 7u) * (nb >> 3) + (b >> 3); }

}  // namespace hm
