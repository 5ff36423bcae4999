/*
 * harmonia.h -- C ABI of the MI355X-native Harmonia B+tree.
 *
 * The tree keeps Harmonia's layout: inner nodes in one breadth-first key
 * region (every node = an 8-key index granule followed by F key slots) plus a
 * prefix-sum array holding each inner node's first-child index, so child i of
 * node p is node prefix[p] + i.  Leaves are 128-byte records of 8 keys and
 * their 8 values, contiguous in key order (the leaf chain is leaf i -> i + 1).
 * Unused key slots hold HM_KEY_MAX, which is therefore not a storable key.
 *
 * Every entry point returns an int status (HM_OK or a negative HM_E* code);
 * nothing here calls exit().  hm_last_error() returns the calling thread's
 * last message.  Pointers to keys, values and results may be device memory or
 * host memory (host memory is staged through the library's buffers).
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *
 * Reference interfaces these entry points replace are cited per function
 * (paths relative to the reference repository root, H = open-harmonia/).
 */
#ifndef HARMONIA_H
#define HARMONIA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_OK 0
#define HM_EINVAL (-1)  /* invalid argument (bad size, null pointer, duplicate or reserved key) */
#define HM_ENOMEM (-2)  /* device allocation failed */
#define HM_EBUDGET (-3) /* request exceeds the tree's device-memory budget */
#define HM_EDEVICE (-4) /* HIP runtime / kernel error */

#define HM_KEY_MAX 0xFFFFFFFFFFFFFFFFull /* padding value; never a valid key */

/* hm_params.flags */
#define HM_FLAG_KEYS_SORTED 1u /* build input is already strictly ascending */

typedef struct hm_tree hm_tree;

typedef struct hm_params {
  uint32_t fanout;         /* inner-node fanout F: 32, 64 or 128 (H/CMakeLists.txt:55 DEFAULT_ORDER) */
  uint32_t leaf_fill;      /* keys per leaf at bulk build, 1..8 (slack = 8 - fill for inserts) */
  int32_t lds_levels;      /* top inner levels staged in LDS, 0..3; -1 = pick automatically */
  uint32_t group_size;     /* lanes cooperating on one query: 1, 2, 4, 8 (index granule, then one key
                            * granule per node), 16, 32, 64 (the whole node in one step; at most F);
                            * 0 = automatic */
  uint32_t psa_bits;       /* partial-sort bits for batched search (reference: 16); 0 = PSA off */
  uint32_t reserved0;
  uint64_t sort_threshold; /* PSA applies only to batches of at least this many queries */
  uint64_t mem_budget;     /* device bytes the tree may hold (tree + growth); 0 = unlimited */
  uint32_t flags;          /* HM_FLAG_* */
  uint32_t reserved1;
} hm_params;

typedef struct hm_stats {
  uint64_t n_keys;
  uint64_t n_leaves;
  uint64_t n_inner;        /* inner nodes over all levels */
  uint32_t height;         /* number of inner levels (leaf level not counted) */
  uint32_t fanout;
  uint32_t node_slots;     /* u64 slots per inner node (index granule + F keys) */
  uint32_t lds_levels;     /* LDS-resident levels the search uses by default */
  uint64_t level_start[17];/* BFS index of the first node of each inner level; [height] = n_inner */
  uint64_t bytes_inner;
  uint64_t bytes_prefix;
  uint64_t bytes_leaves;
  uint64_t bytes_device;   /* everything the tree holds on the device */
  uint64_t epoch;          /* number of published update epochs */
  uint64_t relayouts;      /* epochs that re-laid out every leaf (a last-level node was full) */
  uint64_t respreads;      /* last-level nodes whose keys an epoch re-spread over their own leaves in
                            * place (a leaf overflowed, the node had room) */
  uint64_t key_min;        /* smallest stored key (HM_KEY_MAX for an empty tree) */
  uint64_t key_max;        /* largest stored key (0 for an empty tree) */
  uint32_t psa_window_top; /* input-order PSA searches sort by bits [top - psa_bits, top): one past the
                            * highest bit in which key_min and key_max differ (64 for fewer than two
                            * keys); PSA-order results (HM_OUT_PSA_ORDER) always use [64 - psa_bits, 64) */
  uint32_t reserved;
} hm_stats;

/* hm_search_opts.output_mode (with PSA on; without PSA results are always in input order) */
#define HM_OUT_DEFAULT 0 /* the tree's choice */
#define HM_OUT_UNPERMUTE 1 /* input order: the sort passes replayed backwards over the results (coalesced) */
#define HM_OUT_SCATTER 2 /* input order: each result written straight to its position */
#define HM_OUT_PSA_ORDER 3 /* PSA order plus out_pos, as the reference returns them (H/regular_bpt/prefix-sort.cu:296-297) */

/* Per-call overrides for batched search (used by the autotuner); 0 / -1 = tree default. */
typedef struct hm_search_opts {
  uint32_t group_size;     /* lanes per query: 1, 2, 4, 8, 16, 32, 64 (at most the fanout) */
  int32_t lds_levels;      /* -1 = default */
  int32_t psa_bits;        /* -1 = default (PSA for batches >= sort_threshold), 0 = off, 1..32 = on */
  uint32_t blocks_per_cu;  /* 0 = default */
  uint32_t output_mode;    /* HM_OUT_* */
  uint32_t queries_per_group; /* independent queries each group keeps in flight: 1, 2, 4; 0 = default */
  uint32_t* out_pos;       /* with HM_OUT_PSA_ORDER: out_pos[i] = input index of result i */
  uint32_t sorted_kernel;  /* search kernel for PSA-sorted batches: HM_SORTED_* */
  uint32_t reserved;
} hm_search_opts;

/* hm_search_opts.sorted_kernel */
#define HM_SORTED_AUTO 0    /* the tree's choice (one descent per query) */
#define HM_SORTED_QUERY 1   /* one descent per query (each wave shares its upper-level node reads and
                             * loads the run of last-level nodes its queries reach) */
#define HM_SORTED_WINDOW 2  /* leaf window: each wave streams the leaves its chunk of sorted queries
                             * spans into LDS and ranks the chunk there (chunks spanning more
                             * than 256 leaves fall back to one descent per query) */

/* Fill *p with defaults: F=64, fill 8, LDS auto, group auto, PSA 16 bits, threshold 2^20. */
int hm_params_init(hm_params* p);

/* Bulk build.  Replaces the insert loop + relayout of the reference:
 *   BPlusTree::insert + BPlusTree::sort + prepareArray (H/src/bpt.cpp:357-409,
 *   H/src/traversal.cpp:35-195), insert + generateCFB (H/regular_bpt/regular_bpt.cpp:759,
 *   H/regular_bpt/traversal.cpp:36-99) and the upload prepareGPU (H/regular_bpt/helper.cu:12-28).
 * keys need not be sorted; duplicates and HM_KEY_MAX are rejected with HM_EINVAL.
 * n == 0 builds an empty tree. */
int hm_tree_build(const uint64_t* keys, const uint64_t* vals, uint64_t n, const hm_params* p, hm_tree** out);

int hm_tree_destroy(hm_tree* t);

/* Batched point search, results in input order: out_found[i] = 1 and out_vals[i] = value when
 * q[i] is present, else out_found[i] = 0 and out_vals[i] = 0.
 * Replaces multi_prefix_sort / full_multi_prefix_sort_8thread (H/regular_bpt/prefix-sort.cu:311,
 * H/src/full-multi-prefix-sort-8thread.cu:329) and BPlusTree::search (H/src/bpt.cpp:673-680).
 * Buffers may be device or host memory.  With host buffers the call returns when the results are in
 * them; when q, out_vals and out_found are all host memory and n > 2^22, the batch is streamed through
 * the device in 2^22-query chunks on two streams of the tree's own, copies of one chunk overlapping the
 * other's kernels (the reference's two-stream bucket loop, H/regular_bpt/prefix-sort.cu:261-309);
 * pinned memory (hipHostMalloc) is copied in place, pageable memory through pinned staging. */
int hm_search(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found, void* stream);
int hm_search_ex(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                 const hm_search_opts* opts, void* stream);

/* Allocation-free batched search.  No counterpart in the reference (its drivers cudaMalloc and free
 * every buffer per call, H/src/full-multi-prefix-sort-8thread.cu:338-342,478-491): this is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same results and pipeline as hm_search with the tree's defaults (results in
 * input order), but the sort/un-sort scratch comes from a fixed 44 MiB array in the library's device
 * code (one per GPU, held until the process ends), so the call makes no hipMalloc, hipFree or event
 * creation.  Device buffers only, at most hm_search_noalloc_max() (2^20) queries per call, the tree's
 * device current; larger batches or host pointers fail with HM_EINVAL.  All calls on one GPU share
 * the array: calls that overlap in time race, so serialize them (one stream, or wait for the previous
 * call).  Async on stream. */
int hm_search_noalloc(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                      void* stream);
uint64_t hm_search_noalloc_max(void);

/* Comparator baselines for layout ablations (not a serving path): the reference's pointer-child
 * trees over the same keys and leaves, so only the inner layout differs.  The next node is a
 * dependent read of an explicit child index, child[node * F + slot], instead of prefix[node] + slot.
 *   HM_BASELINE_POINTER_CHILD : index granule, then key granule, then child[] -- the HB+-tree
 *                               search bpt_search_kernel (H/src/gbpt.cu:34-113)
 *   HM_BASELINE_POINTER_LINEAR: the node's keys scanned one 64-B granule per step with an early
 *                               exit, then child[] -- multi_noprefix_nosort_kernel
 *                               (H/regular_bpt/noprefix-nosort.cu:89-237) over the CFB child
 *                               arrays (H/regular_bpt/cfb.h:16-24)
 *   HM_BASELINE_PREFIX_PLAIN  : the control -- the same plain kernel as POINTER_CHILD, next node
 *                               from the prefix-sum array (isolates the layout)
 * One query per group of group_size lanes (1, 2, 4, 8), no LDS levels, no PSA; results in input
 * order, identical to hm_search.  Device buffers only; async on stream.  The child array
 * (n_inner * F u32) is built on first use and again after a publish. */
#define HM_BASELINE_POINTER_CHILD 1
#define HM_BASELINE_POINTER_LINEAR 2
#define HM_BASELINE_PREFIX_PLAIN 3
int hm_baseline_search(hm_tree* t, int kind, uint32_t group_size, const uint64_t* q, uint64_t n, uint64_t* out_vals,
                       uint8_t* out_found, void* stream);

/* Range scan: for each i, the first min(len[i], #keys >= start[i]) keys >= start[i] in ascending
 * order.  out_offsets[0..n] = exclusive scan of len (capacity per range); range i writes its pairs at
 * out_keys/out_vals[out_offsets[i] ..] and its count to out_counts[i].  Writes never pass
 * `capacity` entries.  Replaces range_prefix_sort_1thread_search_and_1thread_scan (H/src/range.cu:312)
 * and BPlusTree::range_search / find_range (H/src/bpt.cpp:695-769, H/regular_bpt/regular_bpt.cpp:320-338). */
int hm_range(hm_tree* t, const uint64_t* start, const uint32_t* len, uint64_t n, uint64_t* out_offsets,
             uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, void* stream);

/* Closed-interval form of the same scan: keys in [lo[i], hi[i]], at most `cap` per range
 * (find_range(start, end), H/regular_bpt/regular_bpt.cpp:320-338). out_offsets[i] = i * cap. */
int hm_range_between(hm_tree* t, const uint64_t* lo, const uint64_t* hi, uint64_t n, uint32_t cap,
                     uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, void* stream);

/* Batched upsert, staged: the batch becomes visible at the next hm_epoch_publish.  Within one
 * epoch a later write of the same key wins.  Replaces BPlusTree::insert_parallel* and
 * rebuild_without_leaf (H/src/bpt.cpp:518-670, H/src/bpt-ppi-update-batch.cpp:86-200,
 * H/src/bpt-ppi-update.cpp:320-463). */
int hm_update(hm_tree* t, const uint64_t* keys, const uint64_t* vals, uint64_t n, void* stream);

/* Apply all staged updates on `stream` after every search/range issued so far on any stream,
 * then make later searches wait for the new snapshot.  out_rebuilt (may be NULL) reports
 * whether leaves had to be re-laid-out (leaf overflow) and the prefix array re-scanned. */
int hm_epoch_publish(hm_tree* t, void* stream, int* out_rebuilt);

/* Batched pipeline for many uneven query batches.  Batches are packed back to back into the
 * batcher's device buffer (hm_batcher_add); hm_batcher_flush then serves all of them with one
 * hipGraph launch: PSA radix sort -> one persistent search launch -> un-sort, every kernel reading
 * the packed query count from device memory, so the graph is captured once and replayed for any
 * count <= capacity.  Batch i's results sit at its offset in the batcher's result buffers.
 * Replaces the reference's fixed-bucket, two-stream host loop (H/regular_bpt/prefix-sort.cu:261-309,
 * H/src/full-multi-prefix-sort-8thread.cu:276-327).  opts (may be NULL) fixes the launch shape;
 * output_mode may be HM_OUT_UNPERMUTE or HM_OUT_SCATTER.  A batcher is used from one thread. */
typedef struct hm_batcher hm_batcher;
int hm_batcher_create(hm_tree* t, uint64_t capacity, const hm_search_opts* opts, void* stream, hm_batcher** out);
/* Add one batch; *out_offset = its first slot.  HM_EBUDGET if full.  A host batch is copied before
 * the call returns; a device batch is only recorded and read by the next flush's work on the
 * batcher's stream, so keep it valid and unchanged until then (hm_batcher_sync). */
int hm_batcher_add(hm_batcher* b, const uint64_t* q, uint64_t n, uint64_t* out_offset);
/* Add batches qs[0..count) (ns[i] queries each) in order until the next one does not fit; *n_added =
 * how many were added (flush, then add the rest), out_offsets[i] (may be NULL) = slot of batch i.
 * flags HM_BATCH_DEVICE: every qs[i] is device memory (the per-batch pointer query is skipped).
 * One call per flush in place of the reference driver's per-bucket host loop
 * (H/regular_bpt/prefix-sort.cu:261-309). */
#define HM_BATCH_DEVICE 1u
int hm_batcher_add_many(hm_batcher* b, const uint64_t* const* qs, const uint64_t* ns, uint64_t count, uint32_t flags,
                        uint64_t* out_offsets, uint64_t* n_added);
/* Launch the pipeline over everything added since the last flush (asynchronous). */
int hm_batcher_flush(hm_batcher* b);
/* Device pointers of the result buffers (valid in stream order after a flush). */
int hm_batcher_results(hm_batcher* b, const uint64_t** vals, const uint8_t** found);
int hm_batcher_copy_out(hm_batcher* b, uint64_t offset, uint64_t n, uint64_t* out_vals, uint8_t* out_found);
int hm_batcher_sync(hm_batcher* b);
int hm_batcher_destroy(hm_batcher* b);

/* Serving loop.  The tree stays resident; lookup, range and update requests are admitted as they
 * arrive.  Worker threads (one HIP stream each, max_inflight of them) repeatedly take every pending
 * read request (continuous batching: whatever arrived while the other workers' batches were on the
 * GPU), coalesce the lookups into one batched search and the ranges into one batched scan, subject
 * to the device-memory budget for in-flight batches.  When the oldest pending request is an update,
 * every pending update is applied as one epoch: staged and prepared while reads continue on the
 * current epoch, then committed under a short exclusive lock, so a read sees the tree before or
 * after an epoch, never in between.  The reference has no serving layer: its drivers run one fixed
 * bucket after another and report no latency (H/regular_bpt/prefix-sort.cu:261-349). */
typedef struct hm_server hm_server;
typedef struct hm_server_params {
  uint64_t mem_budget;   /* device bytes for in-flight request batches (0 = 1 GiB) */
  uint32_t max_batch;    /* queries (or range keys) coalesced into one launch (0 = 1 << 20) */
  uint32_t max_inflight; /* batches on the GPU at once, one worker + stream each (0 = 4) */
  uint32_t coalesce;     /* 1 = continuous batching; 0 = one request per launch (the plain path) */
  uint32_t reserved;
} hm_server_params;
typedef struct hm_server_stats {
  uint64_t completed, lookups, ranges, updates, queries, rejected, launches, epochs;
  double p50_us, p99_us, mean_us, max_us; /* request latency, submit -> results written */
  double span_s;                          /* first submit -> last completion */
  /* worker-side durations (longest since the last reset): one read batch from taking it to its
   * results written; one update epoch, its prepare phase, and the wait for the exclusive commit */
  double max_read_batch_us, max_update_us, max_prepare_us, max_commit_wait_us;
  double p99_lookup_us, p99_range_us, p99_update_us; /* request latency p99 by kind */
} hm_server_stats;
int hm_server_create(hm_tree* t, const hm_server_params* p, hm_server** out);
/* Requests: every buffer (inputs and outputs, host memory) must stay valid and unchanged until the
 * request completes (hm_server_wait returns or hm_server_drain returns); inputs are read once, when
 * the request joins a batch, and outputs are written before completion.  Requests pending at the
 * same time may be ordered either way (reads may run before or after a pending update); a request
 * submitted after another one completed sees its effect. */
int hm_server_lookup(hm_server* s, const uint64_t* keys, uint32_t n, uint64_t* out_vals, uint8_t* out_found,
                     uint64_t* ticket);
int hm_server_range(hm_server* s, const uint64_t* starts, const uint32_t* lens, uint32_t n, uint64_t* out_keys,
                    uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, uint64_t* ticket);
int hm_server_update(hm_server* s, const uint64_t* keys, const uint64_t* vals, uint32_t n, uint64_t* ticket);
/* Block until the request completes; *status = its HM_* result, *epoch = the tree epoch it saw. */
int hm_server_wait(hm_server* s, uint64_t ticket, int* status, uint64_t* epoch);
int hm_server_drain(hm_server* s);
int hm_server_get_stats(hm_server* s, hm_server_stats* st);
int hm_server_reset_stats(hm_server* s);
int hm_server_destroy(hm_server* s);

/* Autotuner.  The reference fixes fanout / narrowing / PSA bits at compile time
 * (H/CMakeLists.txt:55-56, H/regular_bpt/prefix-sort.cu:288-293) and the paper sizes the thread
 * group with a static model (Harmonia.pdf p.140-141); here candidates are timed on the device with a
 * sample batch: lanes per query {1,2,...,64} x queries per group {1,2,4} x blocks per CU {4,5,8,10} x
 * LDS-resident levels for the unsorted path; PSA bits {12,16,20,24} x output mode x shape for the
 * sorted path; then the sort threshold (the smallest batch on a sqrt(2) grid ending at the sample size
 * at which PSA wins).  The winners
 * become the tree's defaults. */
#define HM_TUNE_QUICK 1u /* fewer timing repetitions */
typedef struct hm_tune_result {
  uint32_t fanout;
  uint32_t direct_group, direct_queries_per_group, direct_blocks_per_cu;
  int32_t lds_levels;
  uint32_t psa_group, psa_queries_per_group, psa_blocks_per_cu;
  uint32_t psa_bits, output_mode;
  uint64_t sort_threshold; /* UINT64_MAX: PSA never won up to the sample size */
  double direct_ms, psa_ms, best_ms; /* per call at the sample size */
  uint32_t trials;
  uint32_t reserved;
} hm_tune_result;
int hm_autotune(hm_tree* t, const uint64_t* sample, uint64_t n, uint32_t flags, hm_tune_result* out);
/* Build with fanout 32, 64 and 128, autotune each on the sample, keep the fastest. */
int hm_tree_build_tuned(const uint64_t* keys, const uint64_t* vals, uint64_t n, const hm_params* p,
                        const uint64_t* sample, uint64_t m, uint32_t flags, hm_tree** out, hm_tune_result* res);

int hm_tree_stats(const hm_tree* t, hm_stats* s);

/* Datasets in the reference's text format: one decimal key per line, read by the reference with
 * getline + sscanf("%lld") (H/src/main.cpp:50-68, H/src/bpt.h:32-35).  Every non-empty line must
 * hold one integer in [0, 2^63); anything else fails with the line number.  keys = NULL, cap = 0
 * counts only. */
int hm_read_keys_text(const char* path, uint64_t* keys, uint64_t cap, uint64_t* n);
/* Range files: a start line and an end line per range (H/src/range.cu:348-357); *n = ranges. */
int hm_read_ranges_text(const char* path, uint64_t* starts, uint64_t* ends, uint64_t cap, uint64_t* n);

/* Binary snapshot for a fast reload (no counterpart in the reference, which rebuilds the tree from
 * the text file on every run, H/src/main.cpp:50-68).  Save needs no staged updates.  Load takes
 * fanout and leaf fill from the file and the rest of p (NULL = defaults); it uploads the leaves,
 * rebuilds the inner levels and prefix array on the device and rejects a file whose stored inner
 * sections differ from them. */
int hm_tree_save(const hm_tree* t, const char* path);
int hm_tree_load(const char* path, const hm_params* p, hm_tree** out);

/* Copy the device layout to host buffers (any may be NULL): inner (n_inner * node_slots u64),
 * prefix (n_inner u32), leaves (n_leaves * 16 u64: 8 keys then 8 values per leaf). */
int hm_tree_export(const hm_tree* t, uint64_t* inner, uint32_t* prefix, uint64_t* leaves);

/* PSA partial sort: stable LSD radix sort of (key, position) over key bits [begin_bit, end_bit),
 * the semantics of cub::DeviceRadixSort::SortPairs(keys, pos, n, begin_bit, end_bit) with
 * pos = 0..n-1 (H/regular_bpt/prefix-sort.cu:288-293, H/cub/cub/device/device_radix_sort.cuh:148-175). */
int hm_psa_sort(const uint64_t* keys, uint64_t n, int begin_bit, int end_bit, uint64_t* keys_out,
                uint32_t* pos_out, void* stream);

/* Per-stage kernel timing with HIP events on the launch stream (off by default). */
int hm_profile_enable(int on);
int hm_profile_reset(void);
/* stage: "search", "psa_hist", "psa_scatter", "scan", "range", "build", "update" ... */
int hm_profile_get(const char* stage, double* total_ms, uint64_t* launches);

int hm_device_synchronize(void);
int hm_last_error(char* buf, uint64_t len);

#ifdef __cplusplus
}
#endif
#endif /* HARMONIA_H */
