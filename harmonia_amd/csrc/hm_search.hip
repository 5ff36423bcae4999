// Batched point search over the Harmonia layout.
//
// Replaces multi_prefix_sort_kernel (H/regular_bpt/prefix-sort.cu:98-250) and
// full_multi_prefix_sort_8thread_kernel (H/src/full-multi-prefix-sort-8thread.cu:95-263):
// level-synchronous descent child = prefix[node] + slot, then an exact match in the leaf.
// Differences by design: groups of G lanes of a 64-lane wave instead of 32-lane warps with
// shared-memory flag arrays; the top levels are staged in LDS (the reference's v14 kernels keep
// the top-level prefix entries in __constant__ memory, H/src/ppi-bpt-v14-1.cu:128,223);
// the value is read on the device and written with a found flag in the caller's order.
#include "hm_device.h"

namespace hm {
namespace {

constexpr int kThreads = 256;

// node_slots(F) is even, so the wave buffers after the staged nodes stay 16-byte aligned.
size_t sorted_lds_bytes(const TreeView& tv, int levels, int Q) {
  const size_t nl = levels > 0 ? tv.level_start[levels] : 0;
  return (nl + (size_t)(kThreads / 64) * Q) * node_slots(tv.fanout) * sizeof(uint64_t);
}

// Leaf step of a group's Q descents: exact match of t[i] in leaf[i]; the result of query qi[i] goes to
// (ov, of)[qi[i]] (or to pos[qi[i]] with OUT_SCATTER), skipped for qi[i] >= end.  EARLY_POS: the
// OUT_SCATTER positions are loaded with the leaves (a load inside the writer's branch waits a round trip
// per query); off where the Q extra registers would spill.
template <int G, int Q, int OUT, bool EARLY_POS = true>
__device__ __forceinline__ void leaf_step(const TreeView& tv, const uint64_t (&t)[Q], const uint32_t (&leaf)[Q],
                                          const uint32_t (&qi)[Q], uint32_t end, const uint32_t* __restrict__ pos,
                                          uint64_t* __restrict__ ov, uint8_t* __restrict__ of, int gl, int gbase) {
  constexpr int W = 8 / G;
  // G == 1: one lane per query reads its 8 leaf keys, then only the matching value (same 128-B
  // line, an L2 hit); G > 1: the group reads keys and values of the leaf line together.
  constexpr bool kValsWithKeys = G > 1;
  uint64_t k[Q][W], v[Q][W];
  constexpr bool kEarly = OUT != OUT_DIRECT && EARLY_POS;
  uint32_t o[kEarly ? Q : 1];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Leaf* lf = tv.leaves + leaf[i];
    load_keys<W>(lf->key + gl * W, k[i]);
    if constexpr (kValsWithKeys) load_keys<W>(lf->val + gl * W, v[i]);
    if constexpr (kEarly) o[i] = pos[qi[i] < end ? qi[i] : end - 1];
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int lt = group_count_lt<G, W>(k[i], t[i], gbase);  // lower_bound slot
    // exact match (H/regular_bpt/prefix-sort.cu:220-227); KEY_MAX is padding, never a stored key
    const bool eq = group_any_eq<G, W>(k[i], t[i], gbase) && t[i] != KEY_MAX;
    const int writer = eq ? lt / W : 0;
    if (qi[i] < end && gl == writer) {
      uint64_t val = 0ull;
      if constexpr (kValsWithKeys) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (eq && lt % W == w) val = v[i][w];
      } else {
        if (eq) val = tv.leaves[leaf[i]].val[lt];
      }
      if constexpr (OUT == OUT_DIRECT) {
        ov[qi[i]] = val;
        of[qi[i]] = eq ? 1 : 0;
      } else {
        uint32_t p;
        if constexpr (kEarly)
          p = o[i];
        else
          p = pos[qi[i]];
        ov[p] = val;
        of[p] = eq ? 1 : 0;
      }
    }
  }
}

// Queries [begin, end) of qs; result i goes to (ov, of)[i] (or to pos[i] with OUT_SCATTER).
template <int F, int G, int Q, int OUT>
__device__ __forceinline__ void search_run(const TreeView& tv, int lds_levels, const uint64_t* smem,
                                           const uint64_t* __restrict__ qs, const uint32_t* __restrict__ pos,
                                           uint64_t* __restrict__ ov, uint8_t* __restrict__ of, uint32_t begin,
                                           uint32_t end) {
  const int gl = threadIdx.x % G;
  const int gbase = (threadIdx.x & 63) - gl;
  const uint32_t gid = threadIdx.x / G;
  constexpr uint32_t gpb = kThreads / G;
  constexpr uint32_t tile = gpb * Q;
  for (uint32_t qb = begin; qb < end; qb += tile) {  // uniform across the block
    uint64_t t[Q], td[Q];
    uint32_t leaf[Q], qi[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      qi[i] = qb + i * gpb + gid;
      t[i] = qs[qi[i] < end ? qi[i] : end - 1];  // branch-free: lanes past the end repeat the last query
      td[i] = t[i] == KEY_MAX ? KEY_MAX - 1 : t[i];
    }
    descend_q<F, G, Q>(tv, lds_levels, smem, td, leaf, gl, gbase);
    leaf_step<G, Q, OUT>(tv, t, leaf, qi, end, pos, ov, of, gl, gbase);
  }
}

// Persistent grid: each block owns one contiguous chunk of the batch; the count may live on the
// device (graph replays with data-dependent batch sizes).
template <int F, int G, int Q, int OUT>
__device__ __forceinline__ void search_body(const TreeView& tv, int lds_levels, const uint64_t* __restrict__ q,
                                            const uint32_t* __restrict__ pos, uint64_t* __restrict__ out_vals,
                                            uint8_t* __restrict__ out_found, uint32_t n,
                                            const uint32_t* __restrict__ dn) {
  extern __shared__ __align__(16) uint64_t smem[];
  stage_top_levels<F>(tv, lds_levels, smem);
  constexpr uint32_t tile = kThreads / G * Q;
  if (dn) n = *dn;
  uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + tile - 1) / tile * tile;
  const uint32_t vb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = vb * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  search_run<F, G, Q, OUT>(tv, lds_levels, smem, q, pos, out_vals, out_found, begin, end);
}

template <int F, int G, int Q, int OUT>
__global__ __launch_bounds__(kThreads) void search_kernel(TreeView tv, int lds_levels,
                                                          const uint64_t* __restrict__ q,
                                                          const uint32_t* __restrict__ pos,
                                                          uint64_t* __restrict__ out_vals,
                                                          uint8_t* __restrict__ out_found, uint32_t n,
                                                          const uint32_t* __restrict__ dn) {
  search_body<F, G, Q, OUT>(tv, lds_levels, q, pos, out_vals, out_found, n, dn);
}

// search_kernel over a segment list (unsorted queries in several device arrays, results at each
// segment's slots), so a packed batch needs no gather copy before the search.  Blocks split [0, total)
// into chunks exactly as search_kernel does; a block finds the first segment of its chunk through
// gran_first and walks the segments its chunk overlaps.
template <int F, int G, int Q>
__global__ __launch_bounds__(kThreads) void search_seg_kernel(TreeView tv, int lds_levels,
                                                              const QuerySegList* __restrict__ segs,
                                                              uint64_t* __restrict__ out_vals,
                                                              uint8_t* __restrict__ out_found) {
  extern __shared__ __align__(16) uint64_t smem[];
  stage_top_levels<F>(tv, lds_levels, smem);
  constexpr uint32_t tile = kThreads / G * Q;
  const uint32_t n = (uint32_t)segs->total;
  const uint64_t m = segs->n_items;
  const QuerySeg* items = segs->items;
  uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + tile - 1) / tile * tile;
  const uint32_t vb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = vb * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  if (begin >= end) return;  // block-uniform
  for (uint64_t j = segs->gran_first[begin >> kSegGranuleLog2]; j < m; ++j) {  // uniform across the block
    const QuerySeg it = items[j];
    if (it.dst >= end) break;
    const uint64_t lo = it.dst > begin ? it.dst : begin;
    const uint64_t hi = it.dst + it.n < end ? it.dst + it.n : end;
    if (lo < hi)
      search_run<F, G, Q, OUT_DIRECT>(tv, lds_levels, smem, it.src + (lo - it.dst), nullptr, out_vals + lo,
                                      out_found + lo, 0u, (uint32_t)(hi - lo));
  }
}

// Wide groups (G = 16, 32, 64 lanes per query): whole-node ranks (descend_wide_q); at the leaf, lanes
// 0..7 of the group hold its 8 keys and values, the rest compare padding.
template <int F, int G, int Q, int OUT>
__global__ __launch_bounds__(kThreads) void search_wide_kernel(TreeView tv, int lds_levels,
                                                               const uint64_t* __restrict__ q,
                                                               const uint32_t* __restrict__ pos,
                                                               uint64_t* __restrict__ out_vals,
                                                               uint8_t* __restrict__ out_found, uint32_t n,
                                                               const uint32_t* __restrict__ dn) {
  extern __shared__ __align__(16) uint64_t smem[];
  stage_top_levels<F>(tv, lds_levels, smem);
  const int gl = threadIdx.x % G;
  const int gbase = (threadIdx.x & 63) - gl;
  const uint32_t gid = threadIdx.x / G;
  constexpr uint32_t gpb = kThreads / G;
  constexpr uint32_t tile = gpb * Q;
  if (dn) n = *dn;
  uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + tile - 1) / tile * tile;
  const uint32_t vb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = vb * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  for (uint32_t qb = begin; qb < end; qb += tile) {  // uniform across the block
    uint64_t t[Q], td[Q];
    uint32_t leaf[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t qi = qb + i * gpb + gid;
      t[i] = q[qi < end ? qi : end - 1];
      td[i] = t[i] == KEY_MAX ? KEY_MAX - 1 : t[i];
    }
    descend_wide_q<F, G, Q>(tv, lds_levels, smem, td, leaf, gl, gbase);
    uint64_t k[Q], v[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      k[i] = gl < LEAF_KEYS ? tv.leaves[leaf[i]].key[gl] : KEY_MAX;
      v[i] = gl < LEAF_KEYS ? tv.leaves[leaf[i]].val[gl] : 0ull;
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t qi = qb + i * gpb + gid;
      const uint64_t kk[1] = {k[i]};
      const int lt = group_count_lt<G, 1>(kk, t[i], gbase);
      const bool eq = group_any_eq<G, 1>(kk, t[i], gbase) && t[i] != KEY_MAX;
      if (qi < end && gl == (eq ? lt : 0)) {
        const uint64_t val = eq ? v[i] : 0ull;
        const uint32_t o = OUT == OUT_DIRECT ? qi : pos[qi];
        out_vals[o] = val;
        out_found[o] = eq ? 1 : 0;
      }
    }
  }
}

// ---- last-level span search (PSA-sorted batches) ----------------------------------------------
//
// Two lanes per query, four queries per group, and each wave takes 128 consecutive queries of the
// sorted batch.  Those queries lie in a narrow key range, so at the upper global levels they share one
// node (loaded once per wave into its LDS buffer), and at the last inner level they reach a short run
// of consecutive nodes (BFS order keeps a level's nodes in key order): 128 queries sorted by their top
// 16 bits out of 2^24 cover one or two 16-bit buckets, 3-5 nodes of 512 keys on a 2^26-key tree.  The
// wave loads that whole run with one set of coalesced 16-B loads and every group ranks its node from
// LDS, so the last level costs one global round trip instead of two dependent granule reads per
// group.  A wave whose run exceeds the buffer (sparse or unsorted input) reads its nodes directly, and
// a group outside the wave's shared upper-level node reads that level directly: correct for any input
// order.  The node and run loads are LDS-direct (global_load_lds): every load of a run is in flight at
// once and no registers hold the pieces (register-staged loads, each consumed by a conditional LDS
// store, were issued one round trip apart).  Five waves per SIMD (94 VGPRs at fanout 32 and 64, 92 at
// 128, with the key prefetch); with the two staged top levels of the 2^26-key fanout-64 tree a block
// takes 30.6 KiB of LDS, so five blocks share a CU.
typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;

// Nodes of a wave's run held in LDS: 9 of fanout 32, 5 of fanout 64, 3 of fanout 128.  The wave's
// buffer is whole KiB (one 16-byte piece per lane per KiB), so every lane's LDS-direct load lands inside it.
constexpr int span_nodes_of(int F) { return F >= 128 ? 3 : (F >= 64 ? 5 : 9); }
template <int F>
constexpr int span_nodes() {
  return span_nodes_of(F);
}
constexpr int span_steps(int F) { return (span_nodes_of(F) * node_slots(F) / 2 + 63) / 64; }
constexpr int node_steps(int F) { return (node_slots(F) / 2 + 63) / 64; }
constexpr int span_bytes(int F) { return (span_steps(F) > node_steps(F) ? span_steps(F) : node_steps(F)) * 1024; }

size_t span_lds_bytes(const TreeView& tv, int levels) {
  const size_t nl = levels > 0 ? tv.level_start[levels] : 0;
  return nl * node_slots(tv.fanout) * sizeof(uint64_t) + (size_t)(kThreads / 64) * span_bytes(tv.fanout);
}

// Pieces [0, count) (16 bytes each, count wave-uniform) of src into buf by LDS-direct loads, piece j at
// byte 16 j: STEPS loads per lane at most, all issued before the wait (lanes past count reload the last
// piece into the buffer's padding).  The caller has finished every read of buf.
template <int STEPS>
__device__ __forceinline__ void wave_stage(const ulonglong2* src, int count, uint64_t* buf, int lane) {
#pragma unroll
  for (int k = 0; k < STEPS; ++k) {
    if (64 * k < count) {  // wave-uniform
      // min(lane + 64 k, count - 1) written so that only `lane` stays live across calls
      const int lim = count - 1 - 64 * k;  // >= 0 here, uniform
      __builtin_amdgcn_global_load_lds((const global_void*)(src + 64 * k + (lane < lim ? lane : lim)),
                                       (lds_void*)(buf + 128 * k), 16, 0, 0);
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the loads have landed in LDS
  wave_lds_sync();
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = w < v ? w : v;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = w > v ? w : v;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// PF: while a tile's leaves are read, LDS-direct loads bring the next tile's 128 keys of the wave into the
// first KiB of its buffer (four dword loads per lane, no registers held across the leaf step), so a tile
// starts with its keys in LDS instead of a global round trip.
template <int F, int OUT, bool PF>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void search_span_kernel(
    TreeView tv, int lds_levels, const uint64_t* __restrict__ q, const uint32_t* __restrict__ pos,
    uint64_t* __restrict__ out_vals, uint8_t* __restrict__ out_found, uint32_t n, const uint32_t* __restrict__ dn) {
  constexpr int G = 2, Q = 4;
  constexpr int NS = node_slots(F);
  constexpr int CH = NS / 2;                   // 16-byte pieces per node
  constexpr int PL = (CH + 63) / 64;           // pieces per lane of one node
  constexpr int SPAN = span_nodes<F>();
  constexpr int SPL = (SPAN * CH + 63) / 64;   // pieces per lane of a full run
  static_assert(span_bytes(F) >= SPL * 1024 && span_bytes(F) >= PL * 1024, "wave buffer holds every lane's piece");
  constexpr uint32_t per_wave = 64 / G * Q;    // 128 consecutive queries per wave and tile
  constexpr uint32_t tile = kThreads / 64 * per_wave;
  extern __shared__ __align__(16) uint64_t smem[];
  stage_top_levels<F>(tv, lds_levels, smem);
  if (dn) n = *dn;
  uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + tile - 1) / tile * tile;
  const uint32_t vb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = vb * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  const int lane = threadIdx.x & 63;
  const int gl = lane % G;
  const int gbase = lane - gl;
  const uint32_t nl = lds_levels > 0 ? tv.level_start[lds_levels] : 0u;
  uint64_t* wbuf = smem + (size_t)nl * NS + (threadIdx.x >> 6) * (span_bytes(F) / sizeof(uint64_t));
  const uint32_t wq = (threadIdx.x >> 6) * per_wave + lane / G;
  const uint64_t* __restrict__ inner = tv.inner;
  const int last = tv.height - 1;
  // the wave's 128 keys of the tile at b as 256 dwords into wbuf, dword indices clamped to the last one
  // of the batch (lanes past the end read a copy of it; every key below n arrives whole)
  // (u32 dword indices: the launcher takes this path only for n <= 2^31 - 1024, so 2 (n + tile) fits)
  const uint32_t last_dw = 2 * n - 1;  // used only when n > 0
  auto prefetch = [&](uint32_t b) {
    const uint32_t w0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(b + (threadIdx.x >> 6) * per_wave));
    const uint32_t* qd = reinterpret_cast<const uint32_t*>(q);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t d = 2 * w0 + 64 * k + lane;
      __builtin_amdgcn_global_load_lds((const global_void*)(qd + (d < last_dw ? d : last_dw)),
                                       (lds_void*)(reinterpret_cast<uint32_t*>(wbuf) + 64 * k), 4, 0, 0);
    }
  };
  if constexpr (PF) {
    if (begin < end) prefetch(begin);
  }
  for (uint32_t qb = begin; qb < end; qb += tile) {  // uniform across the block
    uint64_t t[Q], td[Q];
    uint32_t node[Q], qi[Q];
    if constexpr (PF) {
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the tile's keys are in LDS
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < Q; ++i) t[i] = wbuf[i * (64 / G) + lane / G];
      wave_lds_sync();  // every lane has its keys before the buffer is reused
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const uint32_t j = qb + wq + i * (64 / G);
        t[i] = q[j < end ? j : end - 1];  // lanes past the end repeat the last query
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      qi[i] = qb + wq + i * (64 / G);
      td[i] = t[i] == KEY_MAX ? KEY_MAX - 1 : t[i];
      node[i] = 0;
    }
    int l = 0;
    for (; l < lds_levels; ++l) {
#pragma unroll
      for (int i = 0; i < Q; ++i) node[i] = node_next<F, G>(smem + (size_t)node[i] * NS, td[i], gl, gbase);
    }
    for (; l < last; ++l) {  // upper global levels: one node for the wave's queries
      const uint32_t n0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)node[0]);
      wave_stage<PL>(reinterpret_cast<const ulonglong2*>(inner + (size_t)n0 * NS), CH, wbuf, lane);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        if (node[i] == n0)  // uniform inside a group
          node[i] = node_next<F, G>(wbuf, td[i], gl, gbase);
        else
          node[i] = node_next<F, G>(inner + (size_t)node[i] * NS, td[i], gl, gbase);
      }
      wave_lds_sync();  // the next level overwrites the buffer
    }
    if (l == last) {  // the last inner level: the run of nodes [lo, hi] the wave reaches
      uint32_t lo = node[0], hi = node[0];
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        lo = node[i] < lo ? node[i] : lo;
        hi = node[i] > hi ? node[i] : hi;
      }
      lo = wave_min_u32(lo);
      hi = wave_max_u32(hi);
      if (hi - lo < (uint32_t)SPAN) {  // wave-uniform
        // nodes lo..hi (consecutive in memory), every load in flight at once
        wave_stage<SPL>(reinterpret_cast<const ulonglong2*>(inner + (size_t)lo * NS), (int)(hi - lo + 1) * CH, wbuf,
                        lane);
#pragma unroll
        for (int i = 0; i < Q; ++i) node[i] = node_next<F, G>(wbuf + (size_t)(node[i] - lo) * NS, td[i], gl, gbase);
        wave_lds_sync();  // the next tile overwrites the buffer
      } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) node[i] = node_next<F, G>(inner + (size_t)node[i] * NS, td[i], gl, gbase);
      }
    }
    if constexpr (PF) {  // every read of wbuf in this tile is done: bring in the next tile's keys
      if (qb + tile < end) prefetch(qb + tile);
    }
    uint32_t leaf[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) leaf[i] = node[i] - tv.n_inner;
    leaf_step<G, Q, OUT, (F < 128)>(tv, t, leaf, qi, end, pos, out_vals, out_found, gl, gbase);
  }
}

// Search over PSA-sorted queries, one lane per query, Q per lane: the global levels go through
// sorted_level (each distinct node of a wave read once into the wave's LDS slot), the leaf is read
// per lane (its 8 keys, then only the matching value).  Wave slots follow the staged top levels.
template <int F, int Q, int OUT>
__global__ __launch_bounds__(kThreads) void search_sorted_kernel(TreeView tv, int lds_levels,
                                                                 const uint64_t* __restrict__ q,
                                                                 const uint32_t* __restrict__ pos,
                                                                 uint64_t* __restrict__ out_vals,
                                                                 uint8_t* __restrict__ out_found, uint32_t n,
                                                                 const uint32_t* __restrict__ dn) {
  extern __shared__ __align__(16) uint64_t smem[];
  stage_top_levels<F>(tv, lds_levels, smem);
  constexpr int NS = node_slots(F);
  const uint32_t nl = lds_levels > 0 ? tv.level_start[lds_levels] : 0u;
  uint64_t* wslot = smem + (nl + (threadIdx.x >> 6) * Q) * NS;
  const int lane = threadIdx.x & 63;
  constexpr uint32_t tile = kThreads * Q;
  if (dn) n = *dn;
  uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + tile - 1) / tile * tile;
  const uint32_t vb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = vb * chunk;
  const uint32_t end = begin + chunk < n ? begin + chunk : n;
  const uint64_t* __restrict__ inner = tv.inner;

  for (uint32_t qb = begin; qb < end; qb += tile) {  // uniform across the block
    uint64_t t[Q], td[Q];
    uint32_t node[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t qi = qb + i * kThreads + threadIdx.x;
      t[i] = q[qi < end ? qi : end - 1];  // lanes past the end repeat the last query (all lanes take part)
      td[i] = t[i] == KEY_MAX ? KEY_MAX - 1 : t[i];
      node[i] = 0;
    }
    for (int l = 0; l < lds_levels; ++l) {
#pragma unroll
      for (int i = 0; i < Q; ++i) node[i] = node_next_lds<F>(smem + (size_t)node[i] * NS, td[i]);
    }
    for (int l = lds_levels; l < tv.height; ++l) sorted_level<F, Q, 4>(inner, wslot, td, node, lane);
    uint64_t k[Q][8];
#pragma unroll
    for (int i = 0; i < Q; ++i) load_keys<8>(tv.leaves[node[i] - tv.n_inner].key, k[i]);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t qi = qb + i * kThreads + threadIdx.x;
      const uint32_t leaf = node[i] - tv.n_inner;
      const int lt = group_count_lt<1, 8>(k[i], t[i], 0);
      const bool eq = group_any_eq<1, 8>(k[i], t[i], 0) && t[i] != KEY_MAX;
      if (qi < end) {
        const uint64_t val = eq ? tv.leaves[leaf].val[lt] : 0ull;
        if constexpr (OUT == OUT_DIRECT) {
          out_vals[qi] = val;
          out_found[qi] = eq ? 1 : 0;
        } else {
          const uint32_t o = pos[qi];
          out_vals[o] = val;
          out_found[o] = eq ? 1 : 0;
        }
      }
    }
  }
}

// ---- leaf-window search (dense PSA-sorted batches) -------------------------------------------
//
// After the PSA sort, a chunk of consecutive queries lies in a narrow key range, and every query's
// leaf lies between the leaves of the chunk's smallest and largest key (the leaf of t is
// nondecreasing in t).  When that range holds at most R leaves, the wave streams the range's leaf
// keys into a wave-private LDS ring (consecutive 16-B pieces: each leaf read once, whole 64-B key
// halves), ranks every query against the ring (binary search over the leaf minima, then the leaf's
// 8 keys), and reads from HBM only the value of a matching key.  Each wave walks one contiguous part
// of the batch, so consecutive chunks share most of their range and the ring keeps it: per chunk
// only the leaves past the previous range are loaded.  A chunk whose range exceeds R leaves (a
// sparse part of the batch) is searched query by query instead.
//
// The leaf of t is the last leaf j of the range with min(j) <= t: a leaf's minimum key is its
// separator in the parent (the bulk build and re-layouts set separators to leaf minima, and an
// insert only adds keys >= the separator), so this is the descent's choice, and no leaf is empty.
constexpr int kWinThreads = 256;
constexpr int kWinWaves = kWinThreads / 64;
constexpr int kWinQ = 4;                    // queries per lane per chunk
constexpr uint32_t kWinChunk = 64 * kWinQ;  // queries per chunk

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Keys of leaves [a, b] (b - a < R) into the ring: leaf x at ring[(x % R) * 8], 4 pieces of 16 B.
template <int R>
__device__ __forceinline__ void ring_load(const Leaf* __restrict__ leaves, uint64_t* ring, uint32_t a, uint32_t b,
                                          int lane) {
  const uint32_t pieces = (b - a + 1) * 4;
  constexpr int kBatch = 8;  // loads in flight per lane before their stores
  for (uint32_t p0 = 0; p0 < pieces; p0 += 64 * kBatch) {  // wave-uniform
    ulonglong2 v[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const uint32_t p = p0 + lane + 64 * k;
      const uint32_t x = a + (p < pieces ? p : pieces - 1) / 4;
      v[k] = reinterpret_cast<const ulonglong2*>(leaves[x].key)[p & 3];
    }
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const uint32_t p = p0 + lane + 64 * k;
      if (p < pieces) reinterpret_cast<ulonglong2*>(ring + ((a + p / 4) % R) * 8)[p & 3] = v[k];
    }
  }
}

template <int F, int R>
__global__ __launch_bounds__(kWinThreads) void search_window_kernel(TreeView tv, const uint64_t* __restrict__ q,
                                                                    uint64_t* __restrict__ out_vals,
                                                                    uint8_t* __restrict__ out_found, uint32_t n,
                                                                    const uint32_t* __restrict__ dn) {
  static_assert((R & (R - 1)) == 0, "ring size is a power of two");
  constexpr int kSteps = __builtin_ctz(R);  // binary-search steps over at most R leaves
  __shared__ __align__(16) uint64_t rings[kWinWaves][R * 8];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t* ring = rings[wv];
  if (dn) n = *dn;
  // one contiguous part of the batch per wave, whole chunks
  const uint32_t nwaves = gridDim.x * kWinWaves;
  const uint32_t vw = xcd_block(blockIdx.x, gridDim.x) * kWinWaves + wv;
  uint32_t per = (n + nwaves - 1) / nwaves;
  per = (per + kWinChunk - 1) / kWinChunk * kWinChunk;
  const uint64_t b64 = (uint64_t)vw * per;
  const uint32_t begin = b64 < n ? (uint32_t)b64 : n;
  const uint32_t end = begin + per < n ? begin + per : n;
  uint32_t wlo = 1, whi = 0;  // leaves held by the ring (empty)
  for (uint32_t cb = begin; cb < end; cb += kWinChunk) {  // wave-uniform
    uint64_t t[kWinQ], td[kWinQ];
#pragma unroll
    for (int i = 0; i < kWinQ; ++i) {
      const uint32_t qi = cb + lane + 64 * i;
      t[i] = q[qi < end ? qi : end - 1];  // lanes past the end repeat the last query
      td[i] = t[i] == KEY_MAX ? KEY_MAX - 1 : t[i];
    }
    uint64_t mn = td[0], mx = td[0];
#pragma unroll
    for (int i = 1; i < kWinQ; ++i) {
      mn = td[i] < mn ? td[i] : mn;
      mx = td[i] > mx ? td[i] : mx;
    }
    mn = uniform_u64(wave_min_u64(mn));
    mx = uniform_u64(wave_max_u64(mx));
    // leaves of the smallest and largest key: 8-lane groups, even groups mn, odd groups mx
    const int gl = lane & 7, gbase = lane & ~7;
    const uint32_t lf = descend<F, 8>(tv, 0, nullptr, ((lane >> 3) & 1) ? mx : mn, gl, gbase);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)lf, 0);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)lf, 8);
    uint32_t leaf[kWinQ];
    int lt[kWinQ];
    bool eq[kWinQ];
    if (hi - lo < (uint32_t)R) {
      // bring the ring to [lo, hi]: what it lacks below and above the leaves it holds
      if (wlo > whi || lo > whi || hi < wlo) {
        ring_load<R>(tv.leaves, ring, lo, hi, lane);
      } else {
        if (lo < wlo) ring_load<R>(tv.leaves, ring, lo, wlo - 1, lane);
        if (hi > whi) ring_load<R>(tv.leaves, ring, whi + 1, hi, lane);
      }
      wlo = lo;
      whi = hi;
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < kWinQ; ++i) {
        uint32_t a = lo, b = hi;  // the leaf of td[i] is in [a, b]
#pragma unroll
        for (int st = 0; st < kSteps; ++st) {
          const uint32_t mid = (a + b + 1) >> 1;
          const bool up = ring[(mid % R) * 8] <= td[i];
          if (a < b) {
            if (up)
              a = mid;
            else
              b = mid - 1;
          }
        }
        leaf[i] = a;
        uint64_t k[8];
        load_keys<8>(ring + (a % R) * 8, k);
        lt[i] = group_count_lt<1, 8>(k, t[i], 0);
        eq[i] = group_any_eq<1, 8>(k, t[i], 0) && t[i] != KEY_MAX;
      }
    } else {
      // sparse chunk: one descent per query
      descend_q<F, 1, kWinQ>(tv, 0, nullptr, td, leaf, 0, lane);
#pragma unroll
      for (int i = 0; i < kWinQ; ++i) {
        uint64_t k[8];
        load_keys<8>(tv.leaves[leaf[i]].key, k);
        lt[i] = group_count_lt<1, 8>(k, t[i], 0);
        eq[i] = group_any_eq<1, 8>(k, t[i], 0) && t[i] != KEY_MAX;
      }
    }
    uint64_t val[kWinQ];
#pragma unroll
    for (int i = 0; i < kWinQ; ++i) val[i] = eq[i] ? tv.leaves[leaf[i]].val[lt[i]] : 0ull;
#pragma unroll
    for (int i = 0; i < kWinQ; ++i) {
      const uint32_t qi = cb + lane + 64 * i;
      if (qi < end) {
        out_vals[qi] = val[i];
        out_found[qi] = eq[i] ? 1 : 0;
      }
    }
    wave_lds_sync();  // the next chunk's ring loads overwrite what this one read
  }
}

template <int F>
hipError_t launch_window(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, uint64_t* out_vals,
                         uint8_t* out_found, size_t n, hipStream_t s, const uint32_t* dn) {
  const LaunchCfg& lc = launch_cfg();
  const size_t per_block = (size_t)kWinThreads * kWinQ;
  const size_t max_blocks = (size_t)lc.num_cus * (size_t)(sl.blocks_per_cu > 0 ? sl.blocks_per_cu : 2);
  size_t nb = (n + per_block - 1) / per_block;
  nb = nb < max_blocks ? nb : max_blocks;
  nb = (nb + 7) / 8 * 8;
  hipLaunchKernelGGL((search_window_kernel<F, kWindowLeaves>), dim3((unsigned)nb), dim3(kWinThreads), 0, s, tv, q,
                     out_vals, out_found, (uint32_t)n, dn);
  return hipGetLastError();
}

// Locate for the update path: leaf id, lower_bound slot and found flag per key (input order).
template <int F, int G>
__global__ __launch_bounds__(kThreads) void locate_kernel(TreeView tv, const uint64_t* __restrict__ keys, uint32_t n,
                                                          uint32_t* __restrict__ leaf_out,
                                                          uint32_t* __restrict__ slot_out,
                                                          uint8_t* __restrict__ found_out,
                                                          const uint32_t* __restrict__ dn) {
  if (dn) n = *dn;  // device-side count (<= the launch size)
  constexpr int W = 8 / G;
  const int gl = threadIdx.x % G;
  const int gbase = (threadIdx.x & 63) - gl;
  constexpr uint32_t gpb = kThreads / G;
  const uint32_t stride = gridDim.x * gpb;
  const uint32_t first = blockIdx.x * gpb;
  for (uint32_t qb = first; qb < n; qb += stride) {
    const uint32_t qi = qb + threadIdx.x / G;
    const bool valid = qi < n;
    const uint64_t t = valid ? keys[qi] : 0ull;
    const uint64_t td = t == KEY_MAX ? KEY_MAX - 1 : t;
    const uint32_t leaf = descend<F, G>(tv, 0, nullptr, td, gl, gbase);
    uint64_t k[W];
    load_keys<W>(tv.leaves[leaf].key + gl * W, k);
    const int lt = group_count_lt<G, W>(k, t, gbase);
    const bool eq = group_any_eq<G, W>(k, t, gbase) && t != KEY_MAX;
    if (valid && gl == 0) {
      leaf_out[qi] = leaf;
      slot_out[qi] = (uint32_t)lt;
      found_out[qi] = eq ? 1 : 0;
    }
  }
}

template <int F, int G, int Q>
hipError_t launch_fgq(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                     uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                     const uint32_t* dn) {
  const LaunchCfg& lc = launch_cfg();
  constexpr uint32_t gpb = kThreads / G * Q;
  size_t max_blocks = (size_t)lc.num_cus * (size_t)(sl.blocks_per_cu > 0 ? sl.blocks_per_cu : 8);
  size_t need = (n + gpb - 1) / gpb;
  size_t nb = need < max_blocks ? need : max_blocks;
  nb = (nb + 7) / 8 * 8;
  // PSA-sorted input, two lanes per query, four queries per group: search_span_kernel.
  if constexpr (G == 2 && Q == 4) {
    if (sl.sorted) {
      int levels = sl.lds_levels;  // the wave buffers come on top of the staged levels: stay within 64 KiB
      while (levels > 0 && span_lds_bytes(tv, levels) > 65536) --levels;
      const size_t lds = span_lds_bytes(tv, levels);
      const bool pf = n <= (1u << 31) - 1024;  // the prefetch's u32 dword indices fit
      // (fanout 128 with scattered output has no prefetch form: its registers would spill)
      constexpr bool kScatterPf = F < 128;
      if (mode == OUT_DIRECT && pf)
        hipLaunchKernelGGL((search_span_kernel<F, OUT_DIRECT, true>), dim3((unsigned)nb), dim3(kThreads), lds, s,
                           tv, levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      else if (mode == OUT_DIRECT)
        hipLaunchKernelGGL((search_span_kernel<F, OUT_DIRECT, false>), dim3((unsigned)nb), dim3(kThreads), lds, s,
                           tv, levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      else if (pf)
        hipLaunchKernelGGL((search_span_kernel<F, OUT_SCATTER, kScatterPf>), dim3((unsigned)nb), dim3(kThreads), lds,
                           s, tv, levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      else
        hipLaunchKernelGGL((search_span_kernel<F, OUT_SCATTER, false>), dim3((unsigned)nb), dim3(kThreads), lds, s,
                           tv, levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      return hipGetLastError();
    }
  }
  const size_t lds = lds_bytes_for(tv, sl.lds_levels);
  if (sl.segs && !sl.sorted && mode == OUT_DIRECT) {
    hipLaunchKernelGGL((search_seg_kernel<F, G, Q>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv, sl.lds_levels,
                       sl.segs, out_vals, out_found);
    return hipGetLastError();
  }
  switch (mode) {
    case OUT_DIRECT:
      hipLaunchKernelGGL((search_kernel<F, G, Q, OUT_DIRECT>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                         sl.lds_levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      break;
    default:
      hipLaunchKernelGGL((search_kernel<F, G, Q, OUT_SCATTER>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                         sl.lds_levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
      break;
  }
  return hipGetLastError();
}

template <int F, int Q>
hipError_t launch_sorted_q(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                           uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                           const uint32_t* dn) {
  const LaunchCfg& lc = launch_cfg();
  constexpr uint32_t per_block = kThreads * Q;
  size_t max_blocks = (size_t)lc.num_cus * (size_t)(sl.blocks_per_cu > 0 ? sl.blocks_per_cu : 8);
  size_t need = (n + per_block - 1) / per_block;
  size_t nb = need < max_blocks ? need : max_blocks;
  nb = (nb + 7) / 8 * 8;
  int levels = sl.lds_levels;  // the wave slots come on top of the staged levels: stay within 64 KiB
  while (levels > 0 && sorted_lds_bytes(tv, levels, Q) > 65536) --levels;
  const size_t lds = sorted_lds_bytes(tv, levels, Q);
  if (mode == OUT_DIRECT)
    hipLaunchKernelGGL((search_sorted_kernel<F, Q, OUT_DIRECT>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                       levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
  else
    hipLaunchKernelGGL((search_sorted_kernel<F, Q, OUT_SCATTER>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                       levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
  return hipGetLastError();
}

template <int F, int G, int Q>
hipError_t launch_wide_q(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                         uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                         const uint32_t* dn) {
  const LaunchCfg& lc = launch_cfg();
  constexpr uint32_t gpb = kThreads / G * Q;
  const size_t max_blocks = (size_t)lc.num_cus * (size_t)(sl.blocks_per_cu > 0 ? sl.blocks_per_cu : 8);
  const size_t need = (n + gpb - 1) / gpb;
  size_t nb = need < max_blocks ? need : max_blocks;
  nb = (nb + 7) / 8 * 8;
  const size_t lds = lds_bytes_for(tv, sl.lds_levels);
  if (mode == OUT_DIRECT)
    hipLaunchKernelGGL((search_wide_kernel<F, G, Q, OUT_DIRECT>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                       sl.lds_levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
  else
    hipLaunchKernelGGL((search_wide_kernel<F, G, Q, OUT_SCATTER>), dim3((unsigned)nb), dim3(kThreads), lds, s, tv,
                       sl.lds_levels, q, pos, out_vals, out_found, (uint32_t)n, dn);
  return hipGetLastError();
}

template <int F, int G>
hipError_t launch_wide(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                       uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                       const uint32_t* dn) {
  if constexpr (G > F) {
    return hipErrorInvalidValue;
  } else {
    switch (sl.queries_per_group) {
      case 1: return launch_wide_q<F, G, 1>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      case 2: return launch_wide_q<F, G, 2>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      case 4: return launch_wide_q<F, G, 4>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      default: return hipErrorInvalidValue;
    }
  }
}

template <int F, int G>
hipError_t launch_fg(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                     uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                     const uint32_t* dn) {
  switch (sl.queries_per_group) {
    case 1: return launch_fgq<F, G, 1>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 2: return launch_fgq<F, G, 2>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 4: return launch_fgq<F, G, 4>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    default: return hipErrorInvalidValue;
  }
}

template <int F>
hipError_t launch_f(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                    uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                     const uint32_t* dn) {
  if (sl.sorted && sl.window && mode == OUT_DIRECT) {
    SearchLaunch w = sl;
    w.blocks_per_cu = sl.window_blocks_per_cu;
    return launch_window<F>(tv, w, q, out_vals, out_found, n, s, dn);
  }
  if (sl.sorted && sl.group == 1) {
    switch (sl.queries_per_group) {
      case 1: return launch_sorted_q<F, 1>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      case 2: return launch_sorted_q<F, 2>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      case 4: return launch_sorted_q<F, 4>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
      default: return hipErrorInvalidValue;
    }
  }
  switch (sl.group) {
    case 1: return launch_fg<F, 1>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 2: return launch_fg<F, 2>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 4: return launch_fg<F, 4>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 8: return launch_fg<F, 8>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 16: return launch_wide<F, 16>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 32: return launch_wide<F, 32>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 64: return launch_wide<F, 64>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

size_t lds_bytes_for(const TreeView& tv, int levels) {
  if (levels <= 0) return 0;
  const size_t nl = tv.level_start[levels];
  return nl * (size_t)node_slots(tv.fanout) * sizeof(uint64_t);
}

hipError_t launch_search(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                         uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                     const uint32_t* dn) {
  if (n == 0) return hipSuccess;
  ProfScope prof("search", s);
  switch (tv.fanout) {
    case 32: return launch_f<32>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 64: return launch_f<64>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    case 128: return launch_f<128>(tv, sl, q, pos, out_vals, out_found, n, mode, s, dn);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_locate(const TreeView& tv, const uint64_t* keys, size_t n, uint32_t* leaf_out, uint32_t* slot_out,
                         uint8_t* found_out, hipStream_t s, const uint32_t* dn) {
  if (n == 0) return hipSuccess;
  ProfScope prof("update", s);
  constexpr int G = 8;
  constexpr uint32_t gpb = kThreads / G;
  size_t nb = (n + gpb - 1) / gpb;
  if (nb > 8192) nb = 8192;
  switch (tv.fanout) {
    case 32:
      hipLaunchKernelGGL((locate_kernel<32, G>), dim3((unsigned)nb), dim3(kThreads), 0, s, tv, keys, (uint32_t)n,
                         leaf_out, slot_out, found_out, dn);
      break;
    case 64:
      hipLaunchKernelGGL((locate_kernel<64, G>), dim3((unsigned)nb), dim3(kThreads), 0, s, tv, keys, (uint32_t)n,
                         leaf_out, slot_out, found_out, dn);
      break;
    case 128:
      hipLaunchKernelGGL((locate_kernel<128, G>), dim3((unsigned)nb), dim3(kThreads), 0, s, tv, keys, (uint32_t)n,
                         leaf_out, slot_out, found_out, dn);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hm

namespace hm {
namespace {
__global__ void set_u32_kernel(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
}  // namespace

hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(set_u32_kernel, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

namespace {
template <typename T>
__global__ void copy_elems_kernel(T* __restrict__ dst, const T* __restrict__ src, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
hipError_t copy_elems(T* dst, const T* src, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(copy_elems_kernel<T>, dim3((unsigned)nb), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}
}  // namespace

hipError_t copy_words(uint64_t* dst, const uint64_t* src, size_t n, hipStream_t s) {
  return copy_elems(dst, src, n, s);
}
hipError_t copy_u32s(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s) {
  return copy_elems(dst, src, n, s);
}
}  // namespace hm
