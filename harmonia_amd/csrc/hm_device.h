// Device-side building blocks of the descent, shared by the search, range and update kernels.
//
// A query is handled by a narrowed thread group (NTG, Harmonia paper p.140-141; the reference's
// Thread_Per_Query = DEFAULT_ORDER / Narrow_Rate lanes, H/regular_bpt/prefix-sort.cu:31) of G lanes
// inside a 64-lane wavefront.  The group reads one 64-byte granule (8 keys) per step, G lanes x W
// keys each, and ranks the target inside it with one ballot per key slot and a popcount of the
// group's bits of the 64-bit lane mask.  The reference does the same with __ballot_sync over
// 32-lane warps plus shared-memory flag arrays (H/regular_bpt/prefix-sort.cu:147-173).
#pragma once
#include "hm_block.h"
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

// The low 32 bits of v from the last lane of each G-lane group, in every lane of the group
// (ds_swizzle in bitmask mode: lane i reads lane (i & and_mask) | or_mask of its 32-lane half).
template <int G>
__device__ __forceinline__ uint32_t group_last_u32(uint32_t v) {
  if constexpr (G == 1) {
    return v;
  } else {
    static_assert(G <= 32 && (G & (G - 1)) == 0, "groups of 2..32 lanes");
    constexpr int kPattern = (0x1F & ~(G - 1)) | ((G - 1) << 5);
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, kPattern);
  }
}

// Index step of a node (Harmonia's inner_index, H/src/bpt.h:94-120): the number of index slots
// <= t, i.e. which 8-key group holds the child slot.  The last index slot holds the node's prefix
// entry (the BFS index of its first child): it is returned in `base` and counts as KEY_MAX.
// F <= 64 has one index granule, F = 128 two.
template <int F, int G>
__device__ __forceinline__ int index_step(const uint64_t* nd, uint64_t t, int gl, int gbase, uint32_t& base) {
  constexpr int W = 8 / G;
  constexpr int NG = F / 8 > 8 ? 2 : 1;
  uint64_t k[NG][W];
#pragma unroll
  for (int g = 0; g < NG; ++g) load_keys<W>(nd + 8 * g + gl * W, k[g]);
  base = group_last_u32<G>((uint32_t)k[NG - 1][W - 1]);
  if (gl == G - 1) k[NG - 1][W - 1] = KEY_MAX;
  int j = 0;
#pragma unroll
  for (int g = 0; g < NG; ++g) j += group_count_le<G, W>(k[g], t, gbase);
  return j;
}

// upper_bound rank of t inside one inner node: the child slot (H/regular_bpt/cfb.cpp:15-19,
// H/regular_bpt/regular_bpt.cpp:361-365).  Two granule reads: the index granule picks the group
// of 8 keys, the group gives the slot.
template <int F, int G>
__device__ __forceinline__ uint32_t node_rank(const uint64_t* nd, uint64_t t, int gl, int gbase) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  uint32_t base;
  const int j = index_step<F, G>(nd, t, gl, gbase, base);
  uint64_t k[W];
  load_keys<W>(nd + IS + 8 * j + gl * W, k);
  return (uint32_t)(8 * j + group_count_le<G, W>(k, t, gbase));
}

// The next node of the descent: child = prefix[node] + slot, the prefix entry read with the index
// granule.
template <int F, int G>
__device__ __forceinline__ uint32_t node_next(const uint64_t* nd, uint64_t t, int gl, int gbase) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  uint32_t base;
  const int j = index_step<F, G>(nd, t, gl, gbase, base);
  uint64_t k[W];
  load_keys<W>(nd + IS + 8 * j + gl * W, k);
  return base + (uint32_t)(8 * j + group_count_le<G, W>(k, t, gbase));
}

// Descend from the root to a leaf id.  Levels [0, lds_levels) are read from the copy of the top
// nodes at `snodes` (LDS), deeper levels from global memory.
// t must be < KEY_MAX (callers clamp), so every node holds a separator > t.
template <int F, int G>
__device__ __forceinline__ uint32_t descend(const TreeView& tv, int lds_levels, const uint64_t* snodes, uint64_t t,
                                            int gl, int gbase) {
  constexpr int NS = node_slots(F);
  uint32_t node = 0;
  int l = 0;
  for (; l < lds_levels; ++l) node = node_next<F, G>(snodes + (size_t)node * NS, t, gl, gbase);
  for (; l < tv.height; ++l) node = node_next<F, G>(tv.inner + (size_t)node * NS, t, gl, gbase);
  return node - tv.n_inner;
}

// One level for Q independent queries of a group, every load of a step issued before the first is
// used (memory-level parallelism: random L2/MALL/HBM accesses are latency bound).
template <int F, int G, int Q>
__device__ __forceinline__ void level_q(const uint64_t* nodes, const uint64_t (&t)[Q], uint32_t (&node)[Q], int gl,
                                        int gbase) {
  constexpr int W = 8 / G;
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  constexpr int NG = F / 8 > 8 ? 2 : 1;
  uint64_t k[Q][NG][W];
  uint32_t base[Q];
  int j[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int g = 0; g < NG; ++g) load_keys<W>(nodes + (size_t)node[q] * NS + 8 * g + gl * W, k[q][g]);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    base[q] = group_last_u32<G>((uint32_t)k[q][NG - 1][W - 1]);
    if (gl == G - 1) k[q][NG - 1][W - 1] = KEY_MAX;
    j[q] = 0;
#pragma unroll
    for (int g = 0; g < NG; ++g) j[q] += group_count_le<G, W>(k[q][g], t[q], gbase);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) load_keys<W>(nodes + (size_t)node[q] * NS + IS + 8 * j[q] + gl * W, k[q][0]);
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] = base[q] + 8 * j[q] + group_count_le<G, W>(k[q][0], t[q], gbase);
}

// Q independent descents interleaved per group.
template <int F, int G, int Q>
__device__ __forceinline__ void descend_q(const TreeView& tv, int lds_levels, const uint64_t* snodes,
                                          const uint64_t (&t)[Q], uint32_t (&node)[Q], int gl, int gbase) {
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] = 0;
  int l = 0;
  for (; l < lds_levels; ++l) level_q<F, G, Q>(snodes, t, node, gl, gbase);
  for (; l < tv.height; ++l) level_q<F, G, Q>(tv.inner, t, node, gl, gbase);
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] -= tv.n_inner;
}


// Wide narrowed groups, G = 16, 32 or 64 lanes: the reference's NTG shape, Thread_Per_Query =
// DEFAULT_ORDER / Narrow_Rate lanes stepping over a node's keys (H/regular_bpt/prefix-sort.cu:30,148-167),
// taken to one step: lane gl compares keys [gl * WN, gl * WN + WN) of the node's F key slots (WN = F / G),
// so the whole node is ranked by one load per lane and the index granule is not read.  The prefix entry
// (last index slot) is one more load, to the same address in every lane of the group.
template <int F, int G>
__device__ __forceinline__ uint32_t node_next_wide(const uint64_t* nd, uint64_t t, int gl, int gbase) {
  constexpr int IS = idx_slots(F);
  constexpr int WN = F / G;
  static_assert(WN >= 1 && G >= 16, "wide groups: 16..F lanes");
  uint64_t k[WN];
  load_keys<WN>(nd + IS + gl * WN, k);
  const uint32_t base = (uint32_t)nd[IS - 1];
  return base + (uint32_t)group_count_le<G, WN>(k, t, gbase);
}

// Q descents of a wide group interleaved, every load of a level issued before the first is used.
template <int F, int G, int Q>
__device__ __forceinline__ void descend_wide_q(const TreeView& tv, int lds_levels, const uint64_t* snodes,
                                               const uint64_t (&t)[Q], uint32_t (&node)[Q], int gl, int gbase) {
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  constexpr int WN = F / G;
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] = 0;
  for (int l = 0; l < tv.height; ++l) {
    const uint64_t* nodes = l < lds_levels ? snodes : tv.inner;
    uint64_t k[Q][WN];
    uint32_t base[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint64_t* nd = nodes + (size_t)node[q] * NS;
      load_keys<WN>(nd + IS + gl * WN, k[q]);
      base[q] = (uint32_t)nd[IS - 1];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) node[q] = base[q] + (uint32_t)group_count_le<G, WN>(k[q], t[q], gbase);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] -= tv.n_inner;
}

// The next node for a node held in LDS (one lane per query).
template <int F>
__device__ __forceinline__ uint32_t node_next_lds(const uint64_t* nd, uint64_t t) {
  constexpr int IS = idx_slots(F);
  int j = 0;
#pragma unroll
  for (int i = 0; i < IS - 1; ++i) j += nd[i] <= t ? 1 : 0;
  const uint64_t* g = nd + IS + 8 * j;
  int c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) c += g[i] <= t ? 1 : 0;
  return (uint32_t)nd[IS - 1] + (uint32_t)(8 * j + c);
}

// One global level of the descent for PSA-sorted queries, one lane per query, Q queries per lane.
// Neighbouring sorted queries mostly share their node, so instead of every lane reading two 64-B
// granules of it, the wave loads each distinct node once (coalesced 16-B pieces) into its private
// LDS slot and the lanes that need it rank from LDS.  Up to K distinct nodes per query slot are
// served that way; lanes left over (unsorted input) read their node directly.
template <int F, int Q, int K>
__device__ __forceinline__ void sorted_level(const uint64_t* __restrict__ inner, uint64_t* wslot,
                                             const uint64_t (&t)[Q], uint32_t (&node)[Q], int lane) {
  constexpr int NS = node_slots(F);
  constexpr int CH = NS / 2;          // 16-byte pieces per node
  constexpr int PL = (CH + 63) / 64;  // pieces per lane
  uint32_t todo = (1u << Q) - 1u;     // this lane's queries without a result yet (bit q)
  uint32_t res[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) res[q] = 0;
#pragma unroll 1
  for (int it = 0; it < K; ++it) {
    uint32_t n0[Q];
    uint32_t live = 0;  // wave-uniform: query slots some lane still needs
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint64_t act = __ballot((todo >> q) & 1u);
      if (act) live |= 1u << q;
      n0[q] = act ? (uint32_t)__builtin_amdgcn_readlane((int)node[q], __ffsll((long long)act) - 1) : 0u;
    }
    if (!live) break;
    // every slot is loaded (a finished slot reloads node n0 = 0, harmless) so all loads of the
    // step are issued before the first is waited for and stay in registers
    ulonglong2 piece[Q][PL];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const ulonglong2* src = reinterpret_cast<const ulonglong2*>(inner + (size_t)n0[q] * NS);
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        const int idx = lane + 64 * k;
        piece[q][k] = src[idx < CH ? idx : CH - 1];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
#pragma unroll
      for (int k = 0; k < PL; ++k)
        if (lane + 64 * k < CH) reinterpret_cast<ulonglong2*>(wslot + q * NS)[lane + 64 * k] = piece[q][k];
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (((todo >> q) & 1u) && node[q] == n0[q]) {
        res[q] = node_next_lds<F>(wslot + q * NS, t[q]);
        todo &= ~(1u << q);
      }
    }
    wave_lds_sync();  // before the next iteration overwrites the slots
  }
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if ((todo >> q) & 1u) res[q] = node_next<F, 1>(inner + (size_t)node[q] * NS, t[q], 0, lane);
#pragma unroll
  for (int q = 0; q < Q; ++q) node[q] = res[q];
}

// Stage the top `lds_levels` levels (their nodes, prefix entries included) into dynamic LDS.  Whole
// block-wide steps of 16-byte pieces go by LDS-direct loads, all in flight at once; the partial last step
// (which must not write past the staged region) by one ordinary load per thread.  blockDim.x % 64 == 0.
template <int F>
__device__ __forceinline__ void stage_top_levels(const TreeView& tv, int lds_levels, uint64_t* smem) {
  typedef __attribute__((address_space(1))) void global_void_t;
  typedef __attribute__((address_space(3))) void lds_void_t;
  constexpr int NS = node_slots(F);
  const uint32_t nl = lds_levels > 0 ? tv.level_start[lds_levels] : 0u;
  const uint32_t pieces = nl * NS / 2;
  const uint32_t nt = blockDim.x;
  const uint32_t full = pieces / nt;
  const uint32_t w0 = threadIdx.x & ~63u;  // the wave's first thread: its LDS-direct destination base
  const ulonglong2* src = reinterpret_cast<const ulonglong2*>(tv.inner);
  for (uint32_t k = 0; k < full; ++k)  // uniform across the block
    __builtin_amdgcn_global_load_lds((const global_void_t*)(src + k * nt + threadIdx.x),
                                     (lds_void_t*)(smem + 2 * (k * nt + w0)), 16, 0, 0);
  const uint32_t i = full * nt + threadIdx.x;
  if (i < pieces) reinterpret_cast<ulonglong2*>(smem)[i] = src[i];
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's LDS-direct loads have landed
  __syncthreads();
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous share of the
// virtual block range so neighbouring (PSA-sorted) queries meet in the same L2.  nb % 8 == 0.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) { return (b & 7u) * (nb >> 3) + (b >> 3); }

}  // namespace hm
