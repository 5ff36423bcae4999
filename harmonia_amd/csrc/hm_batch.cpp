// Batched pipeline: many uneven query batches packed into one buffer, served by one hipGraph
// replay of sort -> persistent search -> un-sort with the query count read from device memory.
//
// The reference drives fixed-size buckets from a host thread over two streams, one
// H2D -> SortPairs -> kernel -> D2H sequence per bucket (H/regular_bpt/prefix-sort.cu:261-309).
// Here the graph is captured once per (tree layout, path) and replayed for any packed count up to
// the capacity; small flushes skip PSA (the tree's sort threshold), large ones sort first.
//
// Adding a batch costs no device call: a device batch is recorded (pointer, count, slot) and a host
// batch is copied into pinned staging.  The flush copies the staged host part with one copy kernel (if
// there is one) and replays the graph: on the direct path the search kernel reads every batch in place
// from the segment list; on the PSA path a first kernel gathers the batches into the packed buffer and
// stores the count that the sort and search read.  One graph launch per flush either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hm_tree.h"

using namespace hm;

namespace {

// The batches of one flush as a segment list (QuerySeg: source, slot, count), cut into pieces of
// kSegPiece = 8192 queries (256 threads x 32), in mapped pinned memory written by the host before the
// launch: the graph of list slot k captures the address of header k, the header points at the items.
// Host batches are copied into the packed buffer first and listed as segments of that buffer.
using GatherItem = QuerySeg;
using GatherHdr = QuerySegList;
constexpr uint64_t kGatherPiece = kSegPiece;
constexpr int kGatherUnroll = 8;
constexpr int kGatherBlocks = 2048;  // fixed grid (captured); blocks stride over the items

// The PSA path's packing: items b, b + grid, ... copied into the packed buffer (each thread issues
// kGatherUnroll loads before their stores, so a wave keeps eight 512-byte loads in flight).  Block 0
// stores the flush's count for the kernels after it.  (The direct path reads the segments in place.)
__global__ __launch_bounds__(256) void gather_kernel(const GatherHdr* __restrict__ hdr, uint64_t* __restrict__ dst,
                                                     uint32_t* __restrict__ d_n) {
  const uint64_t m = hdr->n_items;
  const GatherItem* items = hdr->items;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_n = (uint32_t)hdr->total;
  for (uint64_t j = blockIdx.x; j < m; j += gridDim.x) {
    const GatherItem it = items[j];
    uint64_t* d = dst + it.dst;
    for (uint64_t i0 = 0; i0 < it.n; i0 += 256 * kGatherUnroll) {
      uint64_t v[kGatherUnroll];
#pragma unroll
      for (int r = 0; r < kGatherUnroll; ++r) {
        const uint64_t i = i0 + (uint64_t)r * 256 + threadIdx.x;
        v[r] = i < it.n ? it.src[i] : 0ull;
      }
#pragma unroll
      for (int r = 0; r < kGatherUnroll; ++r) {
        const uint64_t i = i0 + (uint64_t)r * 256 + threadIdx.x;
        if (i < it.n) d[i] = v[r];
      }
    }
  }
}

struct Graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t version = ~0ull;  // tree layout version it was captured for
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    version = ~0ull;
  }
};

}  // namespace

struct hm_batcher {
  hm_tree* t = nullptr;
  uint64_t cap = 0;
  hipStream_t s = nullptr;
  bool own_stream = false;
  hm_search_opts opts{};
  bool have_opts = false;
  uint64_t* d_q = nullptr;
  uint64_t* d_vals = nullptr;
  uint8_t* d_found = nullptr;
  uint32_t* d_n = nullptr;
  void* scratch = nullptr;
  Graph g_psa[2], g_direct[2];  // per gather-list slot
  uint64_t fill = 0;
  // batches added since the last flush
  std::vector<GatherItem> items;     // device batches, split into pieces of at most kGatherPiece
  uint64_t* h_stage = nullptr;       // mapped pinned copy of host batches (capacity u64, on first use)
  uint64_t host_lo = ~0ull, host_hi = 0;
  hipEvent_t stage_ev = nullptr;     // the last upload of h_stage (host adds wait for it)
  bool stage_busy = false;
  // gather lists: two mapped pinned buffers that the gather kernel reads in place, used alternately,
  // each reused once its flush is done; h_hdr[k] (mapped) describes list k
  GatherHdr* h_hdr = nullptr;
  GatherItem* h_items[2] = {nullptr, nullptr};
  uint32_t* h_gran[2] = {nullptr, nullptr};  // first segment of each kSegGranule slots (mapped)
  size_t items_cap[2] = {0, 0};
  hipEvent_t items_ev[2] = {nullptr, nullptr};
  bool items_busy[2] = {false, false};
  int slot = 0;
};

namespace {

void free_batcher(hm_batcher* b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  for (int k = 0; k < 2; ++k) {
    b->g_psa[k].reset();
    b->g_direct[k].reset();
  }
  for (void* p : {(void*)b->d_q, (void*)b->d_vals, (void*)b->d_found, (void*)b->d_n, b->scratch})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->h_stage, (void*)b->h_items[0], (void*)b->h_items[1], (void*)b->h_hdr, (void*)b->h_gran[0],
                  (void*)b->h_gran[1]})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t e : {b->stage_ev, b->items_ev[0], b->items_ev[1]})
    if (e) (void)hipEventDestroy(e);
  if (b->own_stream && b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

// Capture gather (list slot k) + the pipeline for `psa` (true: tree's PSA path at full capacity) into g.
int capture(hm_batcher* b, bool psa, int k, Graph& g) {
  hm_tree* t = b->t;
  hm_search_opts o = b->have_opts ? b->opts : hm_search_opts{0, -1, -1, 0, 0, 0, nullptr, 0, 0};
  if (!psa) o.psa_bits = 0;
  // force the PSA decision for the captured capacity regardless of the threshold
  if (psa && o.psa_bits <= 0) o.psa_bits = (int32_t)(t->p.psa_bits ? t->p.psa_bits : 16);
  SearchLaunch sl;
  int bits = 0, mode = 0;
  int rc = resolve_search(t, b->cap, &o, &sl, &bits, &mode);
  if (rc) return rc;
  if (mode == HM_OUT_PSA_ORDER) return set_err(HM_EINVAL, "a batcher returns results in input order");
  char* cur = (char*)b->scratch;
  PipeBufs pb = carve_pipe(cur, b->cap, bits, mode);
  // the direct path with 1..8 lanes per query reads the segments in place; PSA (and wide groups) gather
  const bool in_place = bits == 0 && sl.group <= 8;
  if (in_place) sl.segs = b->h_hdr + k;
  g.reset();
  HM_TRY(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  if (!in_place) {
    hipLaunchKernelGGL(gather_kernel, dim3(kGatherBlocks), dim3(256), 0, b->s, (const GatherHdr*)(b->h_hdr + k), b->d_q,
                       b->d_n);
    rc = hipGetLastError() == hipSuccess ? HM_OK : set_err(HM_EDEVICE, "gather launch failed");
  }
  if (!rc)
    rc = enqueue_pipeline(t->view(), sl, bits, mode, b->d_q, b->cap, b->d_n, pb, b->d_vals, b->d_found, nullptr, b->s);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(b->s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HM_TRY(e);
  g.graph = graph;
  HM_TRY(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  g.version = t->layout_version;
  return HM_OK;
}

}  // namespace

extern "C" {

int hm_batcher_create(hm_tree* t, uint64_t capacity, const hm_search_opts* opts, void* stream, hm_batcher** out) {
  if (!t || !out) return set_err(HM_EINVAL, "NULL argument");
  *out = nullptr;
  if (capacity == 0 || capacity > (1ull << 31) - 1) return set_err(HM_EINVAL, "capacity must be in 1..2^31-1");
  if (opts && opts->output_mode == HM_OUT_PSA_ORDER)
    return set_err(HM_EINVAL, "a batcher returns results in input order");
  hm_batcher* b = new hm_batcher();
  b->t = t;
  b->cap = capacity;
  if (opts) {
    b->opts = *opts;
    b->opts.out_pos = nullptr;
    b->have_opts = true;
  }
  auto fail = [&](int rc) {
    free_batcher(b);
    return rc;
  };
  if (stream) {
    b->s = (hipStream_t)stream;
  } else {
    // graphs cannot be captured on the null stream: use a private stream
    if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess)
      return fail(set_err(HM_EDEVICE, "hipStreamCreate failed"));
    b->own_stream = true;
  }
  const size_t need = pipe_bytes(capacity, 32, HM_OUT_UNPERMUTE);  // the PSA path needs the most scratch
  if (t->p.mem_budget && t->device_bytes() + need + capacity * 17 > t->p.mem_budget)
    return fail(set_err(HM_EBUDGET, "batcher of %llu queries exceeds the device-memory budget",
                        (unsigned long long)capacity));
  if (hipMalloc(&b->d_q, capacity * 8) != hipSuccess || hipMalloc(&b->d_vals, capacity * 8) != hipSuccess ||
      hipMalloc(&b->d_found, capacity) != hipSuccess || hipMalloc(&b->d_n, 256) != hipSuccess ||
      hipMalloc(&b->scratch, need) != hipSuccess)
    return fail(set_err(HM_ENOMEM, "batcher allocation failed"));
  if (int rc = host_alloc_mapped((void**)&b->h_hdr, 2 * sizeof(GatherHdr))) return fail(rc);
  for (int k = 0; k < 2; ++k)
    if (int rc = host_alloc_mapped((void**)&b->h_gran[k], ((capacity >> kSegGranuleLog2) + 2) * sizeof(uint32_t)))
      return fail(rc);
  for (int k = 0; k < 2; ++k)
    if (hipEventCreateWithFlags(&b->items_ev[k], hipEventDisableTiming) != hipSuccess)
      return fail(set_err(HM_EDEVICE, "hipEventCreate failed"));
  *out = b;
  return HM_OK;
}

int hm_batcher_add(hm_batcher* b, const uint64_t* q, uint64_t n, uint64_t* out_offset) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  if (n == 0) {
    if (out_offset) *out_offset = b->fill;
    return HM_OK;
  }
  if (!q) return set_err(HM_EINVAL, "NULL buffer");
  if (b->fill + n > b->cap) return set_err(HM_EBUDGET, "batch of %llu does not fit (%llu of %llu used)",
                                           (unsigned long long)n, (unsigned long long)b->fill,
                                           (unsigned long long)b->cap);
  if (is_device_ptr(q)) {
    for (uint64_t o = 0; o < n; o += kGatherPiece)
      b->items.push_back(GatherItem{q + o, b->fill + o, std::min<uint64_t>(kGatherPiece, n - o)});
  } else {
    if (!b->h_stage) {
      int rc = host_alloc_mapped((void**)&b->h_stage, b->cap * 8);
      if (rc) return rc;
      HM_TRY(hipEventCreateWithFlags(&b->stage_ev, hipEventDisableTiming));
    }
    if (b->stage_busy) {  // the previous flush may still be reading the staging buffer
      HM_TRY(hipEventSynchronize(b->stage_ev));
      b->stage_busy = false;
    }
    std::memcpy(b->h_stage + b->fill, q, n * 8);
    b->host_lo = std::min(b->host_lo, b->fill);
    b->host_hi = b->fill + n;
    // listed as a segment of the packed buffer, which the flush's copy kernel fills first
    for (uint64_t o = 0; o < n; o += kGatherPiece)
      b->items.push_back(GatherItem{b->d_q + b->fill + o, b->fill + o, std::min<uint64_t>(kGatherPiece, n - o)});
  }
  if (out_offset) *out_offset = b->fill;
  b->fill += n;
  return HM_OK;
}

int hm_batcher_add_many(hm_batcher* b, const uint64_t* const* qs, const uint64_t* ns, uint64_t count, uint32_t flags,
                        uint64_t* out_offsets, uint64_t* n_added) {
  if (!b || !n_added) return set_err(HM_EINVAL, "NULL argument");
  *n_added = 0;
  if (count && (!qs || !ns)) return set_err(HM_EINVAL, "NULL batch arrays");
  if (flags & ~HM_BATCH_DEVICE) return set_err(HM_EINVAL, "unknown flags 0x%x", flags);
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t n = ns[i];
    if (b->fill + n > b->cap) break;
    if (n && !qs[i]) return set_err(HM_EINVAL, "NULL buffer for batch %llu", (unsigned long long)i);
    if ((flags & HM_BATCH_DEVICE) && n) {
      for (uint64_t o = 0; o < n; o += kGatherPiece)
        b->items.push_back(GatherItem{qs[i] + o, b->fill + o, std::min<uint64_t>(kGatherPiece, n - o)});
      if (out_offsets) out_offsets[i] = b->fill;
      b->fill += n;
    } else {
      int rc = hm_batcher_add(b, qs[i], n, out_offsets ? out_offsets + i : nullptr);
      if (rc) return rc;
    }
    *n_added = i + 1;
  }
  return HM_OK;
}

namespace {
// Copy the staged host batches into d_q (stream order on b->s) and fill gather-list slot k with the
// device batches for the graph's gather kernel (after that slot's previous flush has finished).
int pack(hm_batcher* b, int k) {
  if (b->host_hi > b->host_lo) {  // host batches first: the gather in the graph writes only device batches
    // a copy kernel reads the mapped staging buffer: no DMA-engine copy in the flush
    HM_TRY(copy_words(b->d_q + b->host_lo, b->h_stage + b->host_lo, b->host_hi - b->host_lo, b->s));
    HM_TRY(hipEventRecord(b->stage_ev, b->s));
    b->stage_busy = true;
  }
  if (b->items_busy[k]) {  // slot k was last used two flushes ago
    HM_TRY(hipEventSynchronize(b->items_ev[k]));
    b->items_busy[k] = false;
  }
  const size_t m = b->items.size();
  if (m > b->items_cap[k]) {
    if (b->h_items[k]) HM_TRY(hipHostFree(b->h_items[k]));  // its last use has completed (event above)
    b->h_items[k] = nullptr;
    b->items_cap[k] = 0;
    const size_t c = std::max<size_t>(2 * m, 1024);
    int rc = host_alloc_mapped((void**)&b->h_items[k], c * sizeof(GatherItem));
    if (rc) return rc;
    b->items_cap[k] = c;
  }
  if (m) std::memcpy(b->h_items[k], b->items.data(), m * sizeof(GatherItem));
  // segments are in slot order (batches are added at the fill position): index them per granule
  uint32_t* gf = b->h_gran[k];
  const uint64_t ng = (b->fill + (1ull << kSegGranuleLog2) - 1) >> kSegGranuleLog2;
  size_t j = 0;
  for (uint64_t gi = 0; gi < ng; ++gi) {
    const uint64_t slot = gi << kSegGranuleLog2;
    while (j + 1 < m && b->items[j].dst + b->items[j].n <= slot) ++j;
    gf[gi] = (uint32_t)j;
  }
  b->h_hdr[k] = GatherHdr{b->h_items[k], (uint64_t)m, b->fill, gf};
  b->items.clear();
  b->host_lo = ~0ull;
  b->host_hi = 0;
  return HM_OK;
}
}  // namespace

int hm_batcher_flush(hm_batcher* b) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  if (b->fill == 0) return HM_OK;
  hm_tree* t = b->t;
  std::lock_guard<std::mutex> g(t->mu);
  bool psa = b->have_opts && b->opts.psa_bits >= 0 ? b->opts.psa_bits > 0
                                                   : (t->p.psa_bits > 0 && b->fill >= t->p.sort_threshold);
  const int k = b->slot;
  Graph& gr = psa ? b->g_psa[k] : b->g_direct[k];
  if (!gr.exec || gr.version != t->layout_version) {
    int rc = capture(b, psa, k, gr);
    if (rc) return rc;
  }
  Workspace* w = nullptr;
  int rc = ensure_ws(t, b->s, 0, &w);  // registers the stream so epoch publishes wait for it
  if (rc) return rc;
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(b->s, t->publish_ev, 0));
  rc = pack(b, k);
  if (rc) return rc;
  {
    ProfScope prof("batcher_graph", b->s);
    HM_TRY(hipGraphLaunch(gr.exec, b->s));
  }
  HM_TRY(hipEventRecord(b->items_ev[k], b->s));
  b->items_busy[k] = true;
  b->slot = k ^ 1;
  HM_TRY(hipEventRecord(w->last_use, b->s));
  b->fill = 0;
  return HM_OK;
}

int hm_batcher_results(hm_batcher* b, const uint64_t** vals, const uint8_t** found) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  if (vals) *vals = b->d_vals;
  if (found) *found = b->d_found;
  return HM_OK;
}

int hm_batcher_copy_out(hm_batcher* b, uint64_t offset, uint64_t n, uint64_t* out_vals, uint8_t* out_found) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  if (offset + n > b->cap) return set_err(HM_EINVAL, "range past the batcher capacity");
  if (n == 0) return HM_OK;
  bool host = false;
  if (out_vals) {
    const bool d = is_device_ptr(out_vals);
    host |= !d;
    HM_TRY(hipMemcpyAsync(out_vals, b->d_vals + offset, n * 8, d ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          b->s));
  }
  if (out_found) {
    const bool d = is_device_ptr(out_found);
    host |= !d;
    HM_TRY(hipMemcpyAsync(out_found, b->d_found + offset, n, d ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          b->s));
  }
  if (host) HM_TRY(hipStreamSynchronize(b->s));
  return HM_OK;
}

int hm_batcher_sync(hm_batcher* b) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  HM_TRY(hipStreamSynchronize(b->s));
  return HM_OK;
}

int hm_batcher_destroy(hm_batcher* b) {
  if (!b) return set_err(HM_EINVAL, "batcher is NULL");
  free_batcher(b);
  return HM_OK;
}

}  // extern "C"
