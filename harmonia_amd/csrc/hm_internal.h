// Internal declarations shared by the HIP translation units of libharmonia_amd.
//
// Layout (all device memory, built on the device):
//   inner   : n_inner nodes in breadth-first order, node_slots(F) u64 each.
//             slots [0, IS)      index granule: idx[j] = key[8j+7] for j < F/8, else KEY_MAX
//             slots [IS, IS+F)   separators key[0..F-2] ascending, KEY_MAX padding; key[F-1] = KEY_MAX
//   prefix  : n_inner u32, prefix[g] = BFS index of the first child of node g
//             (= 1 + exclusive scan of child counts, H/regular_bpt/traversal.cpp:40,58).
//             BFS indices >= n_inner name leaves: leaf = index - n_inner.
//   leaves  : 128-byte records {key[8], val[8]}, keys ascending then KEY_MAX padding.
// On the device the last index slot, idx[IS-1] (KEY_MAX in the layout above: it would be
// key[F-1], or an unused slot), holds prefix[node] instead, so a descent step reads the next
// node's base with the index granule rather than as a separate load; rank computations treat that
// slot as KEY_MAX, and hm_tree_export restores KEY_MAX.
// Child choice is upper_bound (number of separators <= target), leaf match is exact equality
// (H/regular_bpt/prefix-sort.cu:147-173, 202-228; H/regular_bpt/regular_bpt.cpp:346-393).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace hm {

constexpr uint64_t KEY_MAX = ~0ull;
constexpr int LEAF_KEYS = 8;
constexpr int MAX_LEVELS = 16;

struct __attribute__((aligned(128))) Leaf {
  uint64_t key[8];
  uint64_t val[8];
};
static_assert(sizeof(Leaf) == 128, "leaf record must be 128 bytes");

__host__ __device__ constexpr int idx_slots(int F) { return F / 8 < 8 ? 8 : F / 8; }
__host__ __device__ constexpr int node_slots(int F) { return idx_slots(F) + F; }

// Everything a search kernel needs, passed by value.
struct TreeView {
  const uint64_t* inner;
  const uint32_t* prefix;
  const Leaf* leaves;
  uint32_t n_inner;
  uint32_t n_leaves;
  int32_t height;  // inner levels, >= 1
  int32_t fanout;
  uint32_t level_start[MAX_LEVELS + 1];
  // Host side: one past the highest bit in which the tree's keys differ (64 for fewer than two keys).
  // Every key agrees with the others above it, so a PSA window that ends there orders queries in the
  // key range as finely as its width allows.
  int32_t psa_top;
};

// ---- launch helpers (host) -------------------------------------------------------------
struct LaunchCfg {
  int num_cus = 256;
  int max_lds = 160 * 1024;
};
const LaunchCfg& launch_cfg();

// Stage timing (HIP events on the launch stream); no-ops unless profiling is enabled.
struct ProfScope {
  ProfScope(const char* stage, hipStream_t s);
  ~ProfScope();
  int slot;
  hipStream_t stream;
};
// While one is alive, ProfScope records nothing on this thread (the allocation-free entry point:
// profiling events come from hipEventCreate).
struct ProfMute {
  ProfMute();
  ~ProfMute();
};

// ---- scan (hm_scan.hip) ----------------------------------------------------------------
// Exclusive scan of n u32 / u64 values; out may alias in.  tmp must hold scan_tmp_bytes(n).
size_t scan_tmp_bytes(size_t n);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t init, void* tmp,
                              hipStream_t s);
hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t init, void* tmp,
                              hipStream_t s);
// u32 input widened into a u64 exclusive scan (range offsets).
hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, void* tmp, hipStream_t s);

// ---- radix sort (hm_sort.hip) ----------------------------------------------------------
// Stable LSD radix sort of u64 keys with a payload over bits [begin_bit, end_bit).
// Ping-pong between (k0,v0) and (k1,v1); returns which buffer (0/1) holds the result in *which.
size_t radix_tmp_bytes(size_t n);
// payload = position 0..n-1 generated on the first pass (vals_in unused)
hipError_t radix_sort_pos(const uint64_t* keys_in, size_t n, int begin_bit, int end_bit, uint64_t* k0, uint32_t* v0,
                          uint64_t* k1, uint32_t* v1, void* tmp, hipStream_t s, int* which,
                          const uint32_t* dn = nullptr);
hipError_t radix_sort_u64pay(const uint64_t* keys_in, const uint64_t* vals_in, size_t n, int begin_bit, int end_bit,
                             uint64_t* k0, uint64_t* v0, uint64_t* k1, uint64_t* v1, void* tmp, hipStream_t s,
                             int* which);
// Keys-only stable LSD sort over bits [begin_bit, end_bit) that records how to undo it: per pass,
// the per-digit scanned (digit, tile) offsets with the digit totals, and each element's index in its
// tile's digit-ordered staging (16 bits).  `replay` must hold replay_bytes(n, end_bit - begin_bit).
size_t replay_bytes(size_t n, int bits);
hipError_t radix_sort_keys_replay(const uint64_t* keys_in, size_t n, int begin_bit, int end_bit, uint64_t* k0,
                                  uint64_t* k1, void* replay, hipStream_t s, int* which,
                                  const uint32_t* dn = nullptr);
// Route per-element results (vals, found) of the sorted order back to input order by running the
// recorded passes backwards: each pass gathers its tile's digit runs (coalesced) into LDS and writes
// the tile in its input order.  (v_tmp, f_tmp) hold n entries; (vals, found) are clobbered when
// there are three or more passes.
hipError_t unsort_results(uint64_t* vals, uint8_t* found, size_t n, int bits, const void* replay, uint64_t* v_tmp,
                          uint8_t* f_tmp, uint64_t* out_vals, uint8_t* out_found, hipStream_t s,
                          const uint32_t* dn = nullptr);

// ---- build (hm_build.hip) --------------------------------------------------------------
// flags |= 1 if a key is KEY_MAX, |= 2 if keys are not strictly ascending.
hipError_t validate_sorted(const uint64_t* keys, size_t n, uint32_t* d_flags, hipStream_t s);
// OR / AND over keys (to skip radix passes whose digit is constant).
hipError_t key_bit_summary(const uint64_t* keys, size_t n, uint64_t* d_orand /*2 u64*/, hipStream_t s);
hipError_t fill_leaves(const uint64_t* keys, const uint64_t* vals, size_t n, int fill, Leaf* leaves, size_t n_leaves,
                       hipStream_t s);
// Leaf minimum keys (KEY_MAX for an empty leaf) -> child_min.
hipError_t leaf_min_keys(const Leaf* leaves, size_t n_leaves, uint64_t* child_min, hipStream_t s);
// One inner level: m nodes over m_child children whose minimum keys are child_min.
hipError_t build_level(const uint64_t* child_min, size_t m_child, int F, uint64_t* nodes, uint64_t* node_min,
                       uint32_t* nchild, size_t m, hipStream_t s);
// inner[node][IS-1] = prefix[node] for every node (after the prefix scan of a build or rebuild).
hipError_t fold_prefix(uint64_t* inner, const uint32_t* prefix, size_t n_inner, int F, hipStream_t s);

// ---- search (hm_search.hip) ------------------------------------------------------------
// Query segments: segment j holds n queries at src (device memory) whose results go to slots
// [dst, dst + n) of the output; the segments are sorted by dst and tile [0, total).  The list may live
// in mapped pinned host memory (the batcher writes it there before each graph replay).
// gran_first[g] = the first segment that reaches slot g * kSegGranule (a search block finds the
// segments of its chunk with one lookup).
struct QuerySeg {
  const uint64_t* src;
  uint64_t dst;
  uint64_t n;
};
struct QuerySegList {
  const QuerySeg* items;
  uint64_t n_items;
  uint64_t total;  // sum of n
  const uint32_t* gran_first;
};
// Queries per segment: the batcher cuts batches into pieces of at most this many.
constexpr uint64_t kSegPiece = 1u << 13;
constexpr int kSegGranuleLog2 = 12;

struct SearchLaunch {
  int group = 4;        // lanes per query
  int lds_levels = 2;   // levels staged in LDS
  int blocks_per_cu = 8;
  int queries_per_group = 2;  // independent descents interleaved per group
  bool sorted = false;        // queries are PSA-sorted: waves share node reads (group 1, or group 2 x 4)
  bool window = false;        // with sorted, OUT_DIRECT: the leaf-window kernel (dense batches)
  int window_blocks_per_cu = 2;
  // unsorted OUT_DIRECT searches with 1..8 lanes per query: read the queries from these segments
  // (blocks take whole segments) instead of the packed array
  const QuerySegList* segs = nullptr;
};
// Leaves a wave's LDS ring holds in the leaf-window search (64 B of keys each).
constexpr int kWindowLeaves = 256;
// Where a search writes result i of its (possibly PSA-sorted) query array:
enum SearchOut {
  OUT_DIRECT = 0,   // out_vals[i], out_found[i]
  OUT_SCATTER = 1,  // out_vals[pos[i]], out_found[pos[i]]
};
// n sizes the grid; with dn != nullptr the kernel reads the actual count (<= n) from device memory.
hipError_t launch_search(const TreeView& tv, const SearchLaunch& sl, const uint64_t* q, const uint32_t* pos,
                         uint64_t* out_vals, uint8_t* out_found, size_t n, SearchOut mode, hipStream_t s,
                         const uint32_t* dn = nullptr);
// Bytes of LDS the top `levels` need.
size_t lds_bytes_for(const TreeView& tv, int levels);
// *p = v as a one-thread kernel (a graph-safe way to hand a count to the next replay).
hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
// dst[i] = src[i] for n u64 words as a kernel: either side may be device memory or mapped pinned
// host memory, so moving small buffers between host and device takes no DMA-engine copy.
hipError_t copy_words(uint64_t* dst, const uint64_t* src, size_t n, hipStream_t s);
hipError_t copy_u32s(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s);

// Comparator baselines (pointer-child trees): child[node * F + slot] = prefix[node] + slot, and a
// plain search kernel (one query per group) that follows child[] (kind 0: index + key granule,
// kind 1: granule-by-granule node scan) or, as the control, prefix[node] + slot (kind 2).
hipError_t build_child_array(const uint32_t* prefix, uint32_t n_inner, int F, uint32_t* child, hipStream_t s);
hipError_t launch_search_baseline(const TreeView& tv, const uint32_t* child, const uint64_t* q, uint64_t* out_vals,
                                  uint8_t* out_found, uint32_t n, int group, int kind, hipStream_t s);
// Locate (leaf, slot, found) per key (update path); n sizes the grid, dn (device) may hold the
// actual count.
hipError_t launch_locate(const TreeView& tv, const uint64_t* keys, size_t n, uint32_t* leaf_out, uint32_t* slot_out,
                         uint8_t* found_out, hipStream_t s, const uint32_t* dn = nullptr);

// ---- range (hm_range.hip) --------------------------------------------------------------
hipError_t launch_range(const TreeView& tv, const uint64_t* start, const uint64_t* hi, const uint32_t* len,
                        uint32_t cap_const, size_t n, const uint64_t* offsets, uint64_t offset_stride,
                        uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, hipStream_t s);

// ---- update (hm_update.hip) ------------------------------------------------------------
// Marks run ends of a sorted key array (keep the last of equal keys): keep[i] = 1 if last.
hipError_t mark_last_of_runs(const uint64_t* sorted_keys, size_t n, uint32_t* keep, hipStream_t s);
hipError_t compact_pairs(const uint64_t* k, const uint64_t* v, const uint32_t* keep, const uint32_t* keep_scan,
                         size_t n, uint64_t* ko, uint64_t* vo, hipStream_t s);
// Per-leaf insert counts for the not-found keys (keys sorted => leaf ids nondecreasing).
hipError_t count_inserts(const uint32_t* leaf, const uint8_t* found, size_t n, uint32_t* ins_per_leaf,
                         hipStream_t s);
// Small batches (n <= kSmallUpdate): one block sorts the staged (key, value) pairs by key in LDS,
// keeps the last write of each key and writes the m unique pairs to (uk, uv), m to *dm.
constexpr uint32_t kSmallUpdate = 4096;
hipError_t small_sort_dedupe(const uint64_t* pk, const uint64_t* pv, uint32_t n, uint64_t* uk, uint64_t* uv,
                             uint32_t* dm, hipStream_t s);
// Mid-size batches (kSmallUpdate < n <= kMergeUpdate): stable sort of the staged pairs by key, as a
// merge sort -- chunks of kSmallUpdate sorted in LDS, then log2(n / kSmallUpdate) merge passes over
// (key, staging index) in (k0, i0) / (k1, i1), then the values gathered into vout.  *sorted_keys is
// k0 or k1.  1 + log2(n / 4096) + 1 launches in place of the radix sort's 24.
constexpr uint32_t kMergeUpdate = 1u << 22;
hipError_t merge_sort_pairs(const uint64_t* pk, const uint64_t* pv, uint32_t n, uint64_t* k0, uint64_t* k1,
                            uint32_t* i0, uint32_t* i1, uint64_t* vout, hipStream_t s, const uint64_t** sorted_keys);
// *dm = keep_scan[n - 1] + keep[n - 1] (unique count after compact_pairs).
hipError_t unique_count(const uint32_t* keep, const uint32_t* keep_scan, size_t n, uint32_t* dm, hipStream_t s);
// Over the m located unique keys (sorted, so each leaf's keys form one run and each last-level
// node's keys one range): summary[1] += number of new keys; summary[0] |= 1 if some leaf would exceed
// 8 keys.  Such a node goes on the re-spread list (nlist: pairs of node index within the last level and
// first batch entry; summary[4] = their count) when its leaves can hold all its keys, else
// summary[0] |= 2 (a full re-layout is needed).  n sizes the grid; nlist holds 2 n entries.
hipError_t plan_runs(const Leaf* leaves, uint32_t n_leaves, int F, const uint32_t* leaf, const uint8_t* found, size_t n,
                     const uint32_t* dm, uint32_t* summary, uint32_t* nlist, hipStream_t s);
// edge[0] = uk[0] and edge[1] = uk[*dm - 1], the smallest and largest of the *dm sorted unique keys
// (edge = {KEY_MAX, 0} when *dm == 0).  edge may be mapped pinned host memory.
hipError_t edge_keys(const uint64_t* uk, const uint32_t* dm, uint64_t* edge, hipStream_t s);
// found[j] |= 2 for the batch entries of the listed nodes (the in-place merge then skips them).
hipError_t mark_nodes(const uint32_t* nlist, uint32_t n_nodes, int F, const uint32_t* leaf, uint8_t* found,
                      const uint32_t* dm, hipStream_t s);
// Re-spread each listed node in place: its keys, the batch's values and new keys merged, spread evenly
// over its own leaves, separators and index slots rewritten.  last_level = the first node of the last
// inner level.
hipError_t node_respread(Leaf* leaves, uint64_t* last_level, uint32_t n_leaves, int F, const uint32_t* nlist,
                         uint32_t n_nodes, const uint32_t* leaf, const uint32_t* slot, const uint8_t* found,
                         const uint64_t* uk, const uint64_t* uv, const uint32_t* dm, hipStream_t s);
// No leaf overflows: per run, overwrite the values of existing keys and merge the new keys in place.
hipError_t apply_runs(Leaf* leaves, const uint32_t* leaf, const uint32_t* slot, const uint8_t* found,
                      const uint64_t* uk, const uint64_t* uv, size_t n, const uint32_t* dm, hipStream_t s);
// Out of place: new leaf count per old leaf (1 if it fits, else ceil(total / fill)).
hipError_t new_leaf_counts(const Leaf* leaves, const uint32_t* ins_per_leaf, size_t n_leaves, int fill,
                           uint32_t* cnt, hipStream_t s);
// first[leaf[j]] = j for the first batch index j of each leaf's run (first pre-filled with ~0).
hipError_t leaf_first_keys(const uint32_t* leaf, size_t n, uint32_t* first, hipStream_t s);
// Re-layout into new_leaves (old leaves are only read); both parts also write each new leaf's
// minimum key to new_mins[new leaf].
//   relayout_copy: the old leaves in [lo, hi) that the batch does not touch, moved to new_start[i];
//   relayout_touched: the touched leaves -- batch values of existing keys applied, new keys merged,
//   overflowing leaves split evenly.
hipError_t relayout_copy(const Leaf* old_leaves, size_t lo, size_t hi, const uint32_t* first, const uint32_t* new_start,
                         Leaf* new_leaves, uint64_t* new_mins, hipStream_t s);
hipError_t relayout_touched(const Leaf* old_leaves, const uint32_t* ins_per_leaf, const uint32_t* bleaf,
                            const uint32_t* bslot, const uint64_t* bkeys, const uint64_t* bvals, const uint8_t* found,
                            size_t m, const uint32_t* new_start, int fill, Leaf* new_leaves, uint64_t* new_mins,
                            hipStream_t s);

}  // namespace hm
