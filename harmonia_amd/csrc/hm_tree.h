// Host-side tree object and helpers shared by the C ABI translation units (hm_api.cpp,
// hm_batch.cpp, hm_server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/harmonia.h"
#include "hm_internal.h"

namespace hm {

int set_err(int code, const char* fmt, ...);
size_t align256(size_t x);
bool is_device_ptr(const void* p);
// Pinned host memory that kernels read and write at its host address (mapped, coherent).
int host_alloc_mapped(void** p, size_t bytes);

#define HM_TRY(expr)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) {                                                                                \
      int code_ = (e_ == hipErrorOutOfMemory) ? HM_ENOMEM : HM_EDEVICE;                                    \
      return ::hm::set_err(code_, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    }                                                                                                      \
  } while (0)

// Launch configuration of one search path (tuned per tree).
struct SearchCfg {
  int group;
  int queries_per_group;
  int blocks_per_cu;
  int output_mode;
};

struct Workspace {
  void* buf = nullptr;
  size_t bytes = 0;
  hipEvent_t last_use = nullptr;  // the latest search/range issued on this stream
};

}  // namespace hm

struct hm_tree {
  hm_params p{};
  int F = 64, fill = 8, NS = 72;
  int device = 0;
  uint64_t n_keys = 0;
  uint64_t n_leaves = 0, leaf_cap = 0;
  uint32_t n_inner = 0, inner_cap = 0;
  int height = 1;
  uint32_t level_start[hm::MAX_LEVELS + 1] = {0};
  uint64_t* d_inner = nullptr;
  uint32_t* d_prefix = nullptr;
  hm::Leaf* d_leaves = nullptr;
  int lds_levels = 0;
  uint64_t layout_version = 0;  // bumped whenever a device array of the layout is reallocated
  hm::SearchCfg cfg_direct{4, 4, 4, HM_OUT_UNPERMUTE};  // unsorted batches: memory-transaction bound
  // PSA-ordered batches: neighbours share nodes.  10 blocks per CU = two full rounds of the five
  // blocks of this shape that fit on a CU (96 VGPRs): a partial last round of the persistent grid
  // (8 per CU: 5 + 3) measured 3% slower.
  hm::SearchCfg cfg_psa{2, 4, 10, HM_OUT_UNPERMUTE};
  std::mutex mu;
  std::unordered_map<hipStream_t, hm::Workspace> ws;
  hipStream_t own = nullptr;
  // Host-resident batches larger than one chunk stream through two streams of the tree's own (default
  // priority), one chunk in flight on each; pageable caller memory goes through pinned staging.
  hipStream_t hs[2] = {nullptr, nullptr};
  hipEvent_t hs_ev[2] = {nullptr, nullptr};   // the latest chunk on hs[k] completed
  void* hs_stage[2] = {nullptr, nullptr};     // mapped pinned: chunk queries (8 B), values (8 B), found (1 B)
  size_t hs_stage_n = 0;                      // queries per staging slot
  hipEvent_t publish_ev = nullptr;
  bool publish_valid = false;
  // the latest allocation-free search (hm_search_noalloc), so a publish waits for its reads; created
  // with the tree because that entry point creates nothing
  hipEvent_t na_ev = nullptr;
  bool na_valid = false;
  uint64_t epoch = 0;
  // smallest and largest key ever stored (updates only add keys) and the PSA window top they give
  uint64_t key_lo = ~0ull, key_hi = 0;
  int psa_top = 64;
  uint64_t relayouts = 0;
  uint64_t respreads = 0;  // last-level nodes re-spread in place by update epochs
  uint64_t* pend_k = nullptr;
  uint64_t* pend_v = nullptr;
  size_t pend_n = 0, pend_cap = 0;
  uint64_t* pin_stage = nullptr;  // pinned bounce for host update batches: keys [0, pin_cap), values after
  size_t pin_cap = 0;
  hipEvent_t stage_ev = nullptr;  // the latest staging copies (a publish on another stream waits for it)
  bool stage_valid = false;
  void* upd_buf = nullptr;       // grow-only device scratch of hm_update / hm_epoch_publish
  size_t upd_bytes = 0;
  uint32_t* upd_host = nullptr;  // pinned: the publish plan summary read back by the host
  // Standby layout: a re-layout is built here while searches keep reading the active arrays,
  // then the two are swapped.  Kept (grow-only) for the next re-layout.
  hm::Leaf* sb_leaves = nullptr;
  uint64_t sb_leaf_cap = 0;
  uint64_t* sb_inner = nullptr;
  uint32_t* sb_prefix = nullptr;
  uint32_t sb_inner_cap = 0;
  void* rl_buf = nullptr;  // grow-only scratch of re-layouts and inner rebuilds
  size_t rl_bytes = 0;
  std::mutex pub_mu;  // serializes update staging and publishes; taken before mu
  // comparator baselines: explicit child index per inner slot, rebuilt when the layout or epoch changes
  uint32_t* bl_child = nullptr;
  size_t bl_cap = 0;  // u32 entries
  uint64_t bl_version = ~0ull, bl_epoch = ~0ull;

  hm::TreeView view() const {
    hm::TreeView tv{};
    tv.inner = d_inner;
    tv.prefix = d_prefix;
    tv.leaves = d_leaves;
    tv.n_inner = n_inner;
    tv.n_leaves = (uint32_t)n_leaves;
    tv.height = height;
    tv.fanout = F;
    for (int i = 0; i <= hm::MAX_LEVELS; ++i) tv.level_start[i] = level_start[i];
    tv.psa_top = psa_top;
    return tv;
  }
  // Widen the key range [key_lo, key_hi] by [lo, hi] (lo > hi: nothing) and recompute psa_top.
  void widen_key_range(uint64_t lo, uint64_t hi) {
    if (lo > hi) return;
    key_lo = lo < key_lo ? lo : key_lo;
    key_hi = hi > key_hi ? hi : key_hi;
    const uint64_t d = key_lo ^ key_hi;
    psa_top = d ? 64 - __builtin_clzll(d) : 64;
  }
  uint64_t device_bytes() const {
    return (uint64_t)(inner_cap + sb_inner_cap) * (NS * 8 + 4) + (leaf_cap + sb_leaf_cap) * sizeof(hm::Leaf) +
           pend_cap * 16 + upd_bytes + rl_bytes + bl_cap * 4;
  }
};

namespace hm {
// Levels of a layout.
struct Levels {
  int height = 0;
  uint32_t n_inner = 0;
  uint32_t level_start[MAX_LEVELS + 1] = {0};
};

// One epoch being published: what publish_prepare found and built, for publish_commit.
struct PublishPlan {
  size_t n = 0;          // staged updates
  size_t m = 0;          // unique keys among them
  uint64_t n_new = 0;    // keys not in the tree before
  bool relayout = false; // the standby layout holds the new epoch
  uint32_t respread = 0;  // last-level nodes re-spread in place (their list in the update scratch)
  uint64_t new_leaves = 0;
  uint64_t key_lo = ~0ull, key_hi = 0;  // smallest and largest staged key (lo > hi: none)
  Levels lv;
};
// Staging and the two publish phases; the caller holds t->pub_mu across them.  Searches may run
// during staging and publish_prepare; publish_commit is short (an enqueue or a pointer swap).
int stage_updates(hm_tree* t, const uint64_t* keys, const uint64_t* vals, uint64_t n, hipStream_t s);
// hm_search_ex / hm_range on stream s.  mapped = every buffer is device memory or mapped pinned host
// memory (host_alloc_mapped): kernels use the pointers as given, nothing is staged and the call
// returns without waiting for the stream.
int search_impl(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                const hm_search_opts* o, hipStream_t s, bool mapped);
// Stream workspace of a range batch of n ranges with device (or mapped) buffers.
size_t range_ws_bytes(size_t n);
int range_impl(hm_tree* t, const uint64_t* start, const uint32_t* len, uint64_t n, uint64_t* out_offsets,
               uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, hipStream_t s,
               bool mapped);
int publish_prepare(hm_tree* t, hipStream_t s, PublishPlan* plan);
int publish_commit(hm_tree* t, hipStream_t s, const PublishPlan& plan, int* rebuilt);
// Drop the staged batch after a failed prepare or commit (caller holds t->pub_mu).
void discard_staged(hm_tree* t);
// Allocate the standby layout and re-layout scratch now (best effort under the memory budget).
int reserve_relayout(hm_tree* t, hipStream_t s);
// Size the update staging and publish scratch for batches of up to n keys now.
int reserve_updates(hm_tree* t, size_t n, hipStream_t s);

int ensure_ws(hm_tree* t, hipStream_t s, size_t bytes, Workspace** out);
// Grow stream s's workspace to at least `bytes` now (takes t->mu).
int reserve_workspace(hm_tree* t, hipStream_t s, size_t bytes);
int choose_lds_levels(const hm_tree* t, int req);
// Resolve the launch shape of one search call (tree defaults overridden by opts); psa_bits = 0
// means no PSA (then out_mode = 0, results in input order).
int resolve_search(const hm_tree* t, uint64_t n, const hm_search_opts* o, SearchLaunch* sl, int* psa_bits,
                   int* out_mode);

// Device scratch of one sort -> search -> (un-sort) pipeline over up to n queries.
struct PipeBufs {
  uint64_t *k0, *k1;  // PSA-sorted keys (ping-pong)
  uint32_t *p0, *p1;  // positions (ping-pong; scatter and PSA-order outputs)
  void* tmp;          // radix counts of the position sort (scatter and PSA-order outputs)
  void* replay;       // per-pass offsets and staging indices (input-order output)
  uint64_t *rv, *vt;  // PSA-order result values and the un-sort ping-pong buffer
  uint8_t *rf, *ft;   // PSA-order found flags and their ping-pong buffer
};
size_t pipe_bytes(size_t n, int psa_bits, int out_mode);
PipeBufs carve_pipe(char*& cur, size_t n, int psa_bits, int out_mode);
// Enqueue sort -> search -> un-sort on s.  n sizes every launch; dn (device) may hold a smaller
// actual count, so the same enqueue can be captured once into a graph and replayed.  The PSA-order
// mode copies positions to dpos with a device-to-device copy of n entries.
int enqueue_pipeline(const TreeView& tv, const SearchLaunch& sl, int psa_bits, int out_mode, const uint64_t* dq,
                     size_t n, const uint32_t* dn, const PipeBufs& b, uint64_t* dvals, uint8_t* dfound,
                     uint32_t* dpos, hipStream_t s);
}  // namespace hm
