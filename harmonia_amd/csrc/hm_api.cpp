// C ABI of libharmonia_amd: tree ownership, device buffers, epochs and the host side of every
// pipeline (sort -> search, range, staged upsert).  The reference keeps every device buffer inside
// one driver function per variant (allocate, run, free; H/src/full-multi-prefix-sort-8thread.cu:329-495);
// here a tree stays resident on the device and is searched any number of times from any stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/harmonia.h"
#include "hm_internal.h"
#include "hm_tree.h"

namespace hm {
hipError_t range_offsets_tail(const uint32_t* len, uint64_t* offsets, size_t n, hipStream_t s);
}

using namespace hm;

// ---------------------------------------------------------------------------------------------
// errors and small helpers
namespace hm {
thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

int host_alloc_mapped(void** p, size_t bytes) {
  *p = nullptr;
  HM_TRY(hipHostMalloc(p, bytes, hipHostMallocMapped | hipHostMallocCoherent));
  void* dp = nullptr;
  hipError_t e = hipHostGetDevicePointer(&dp, *p, 0);
  if (e != hipSuccess || dp != *p) {
    (void)hipGetLastError();
    (void)hipHostFree(*p);
    *p = nullptr;
    return set_err(HM_EDEVICE, "pinned host memory is not mapped at its host address");
  }
  // touch every page now: the CPU mapping is filled in on first access, one fault per 4-KiB page,
  // which made the first batches through a fresh serving buffer copy at ~2 GB/s
  std::memset(*p, 0, bytes);
  return HM_OK;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}
}  // namespace hm

// ---------------------------------------------------------------------------------------------
// launch configuration and profiling
namespace hm {

const LaunchCfg& launch_cfg() {
  static LaunchCfg cfg;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
      cfg.num_cus = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
      cfg.max_lds = (int)(p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : 65536);
    }
  });
  return cfg;
}

namespace {
struct ProfRec {
  int stage;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<std::string> g_stage_names;
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_ev_pool;

hipEvent_t prof_event() {
  if (!g_ev_pool.empty()) {
    hipEvent_t e = g_ev_pool.back();
    g_ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

thread_local int g_prof_muted = 0;
ProfMute::ProfMute() { ++g_prof_muted; }
ProfMute::~ProfMute() { --g_prof_muted; }

ProfScope::ProfScope(const char* stage, hipStream_t s) : slot(-1), stream(s) {
  if (!g_prof_on || g_prof_muted) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;  // no events in graphs
  std::lock_guard<std::mutex> g(g_prof_mu);
  int id = -1;
  for (size_t i = 0; i < g_stage_names.size(); ++i)
    if (g_stage_names[i] == stage) id = (int)i;
  if (id < 0) {
    g_stage_names.emplace_back(stage);
    id = (int)g_stage_names.size() - 1;
  }
  ProfRec r{id, prof_event(), prof_event()};
  if (!r.a || !r.b) return;
  if (hipEventRecord(r.a, s) != hipSuccess) return;
  g_recs.push_back(r);
  slot = (int)g_recs.size() - 1;
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  if (slot < (int)g_recs.size()) (void)hipEventRecord(g_recs[slot].b, stream);
}

}  // namespace hm

// ---------------------------------------------------------------------------------------------
// tree object helpers
namespace hm {

int ensure_ws(hm_tree* t, hipStream_t s, size_t bytes, Workspace** out) {
  Workspace& w = t->ws[s];
  if (!w.last_use) HM_TRY(hipEventCreateWithFlags(&w.last_use, hipEventDisableTiming));
  if (w.bytes < bytes) {
    const size_t old = w.bytes;
    if (w.buf) {
      HM_TRY(hipStreamSynchronize(s));
      HM_TRY(hipFree(w.buf));
      w.buf = nullptr;
      w.bytes = 0;
    }
    // geometric growth: a hipFree synchronizes the whole device
    size_t want = std::max(std::max(bytes, 2 * old), (size_t)1 << 20);
    HM_TRY(hipMalloc(&w.buf, want));
    w.bytes = want;
  }
  *out = &w;
  return HM_OK;
}

int reserve_workspace(hm_tree* t, hipStream_t s, size_t bytes) {
  std::lock_guard<std::mutex> g(t->mu);
  Workspace* w = nullptr;
  return ensure_ws(t, s, bytes, &w);
}

int choose_lds_levels(const hm_tree* t, int req) {
  const TreeView tv = t->view();
  int best = 0;
  const int maxl = std::min(3, t->height);
  // an explicit request is honoured up to 64 KiB of LDS per block; the default keeps it under 48 KiB
  // so several blocks share a CU
  const size_t cap = req >= 0 ? 64 * 1024 : 48 * 1024;
  const int top = req >= 0 ? std::min(req, maxl) : maxl;
  for (int l = 1; l <= top; ++l)
    if (lds_bytes_for(tv, l) <= cap) best = l;
  return best;
}

int resolve_search(const hm_tree* t, uint64_t n, const hm_search_opts* o, SearchLaunch* sl, int* psa_bits,
                   int* out_mode) {
  int bits = (o && o->psa_bits >= 0) ? o->psa_bits : (int)t->p.psa_bits;
  if (bits < 0 || bits > 32) return set_err(HM_EINVAL, "psa_bits %d not in 0..32", bits);
  const bool psa = bits > 0 && (n >= t->p.sort_threshold || (o && o->psa_bits > 0));
  const SearchCfg& base = psa ? t->cfg_psa : t->cfg_direct;
  sl->group = (o && o->group_size) ? (int)o->group_size : base.group;
  if (sl->group < 1 || sl->group > 64 || (sl->group & (sl->group - 1)) || sl->group > t->F)
    return set_err(HM_EINVAL, "group_size %d not a power of two in 1..min(64, fanout %d)", sl->group, t->F);
  sl->lds_levels = (o && o->lds_levels >= 0) ? choose_lds_levels(t, o->lds_levels) : t->lds_levels;
  sl->blocks_per_cu = (o && o->blocks_per_cu) ? (int)o->blocks_per_cu : base.blocks_per_cu;
  sl->queries_per_group = (o && o->queries_per_group) ? (int)o->queries_per_group : base.queries_per_group;
  if (sl->queries_per_group != 1 && sl->queries_per_group != 2 && sl->queries_per_group != 4)
    return set_err(HM_EINVAL, "queries_per_group %d not in {1,2,4}", sl->queries_per_group);
  int mode = (o && o->output_mode) ? (int)o->output_mode : base.output_mode;
  if (mode < HM_OUT_UNPERMUTE || mode > HM_OUT_PSA_ORDER)
    return set_err(HM_EINVAL, "output_mode %d not in 0..3", mode);
  if (mode == HM_OUT_PSA_ORDER && (!psa || !o || !o->out_pos))
    return set_err(HM_EINVAL, "PSA-order output needs PSA on and out_pos");
  const uint32_t sk = o ? o->sorted_kernel : (uint32_t)HM_SORTED_AUTO;
  if (sk > HM_SORTED_WINDOW) return set_err(HM_EINVAL, "sorted_kernel %u not in 0..2", sk);
  // leaf window: only when asked for (measured slower than one descent per query on the 2^26-key tree)
  sl->window = psa && mode != HM_OUT_SCATTER && sk == HM_SORTED_WINDOW;
  sl->window_blocks_per_cu = (o && o->blocks_per_cu && sk == HM_SORTED_WINDOW) ? (int)o->blocks_per_cu : 2;
  *psa_bits = psa ? bits : 0;
  *out_mode = psa ? mode : 0;
  return HM_OK;
}

size_t pipe_bytes(size_t n, int psa_bits, int out_mode) {
  if (!psa_bits) return 0;
  const size_t kq = align256(n * 8), kp = align256(n * 4), kf = align256(n);
  size_t b = 2 * kq;
  if (out_mode == HM_OUT_UNPERMUTE)
    b += align256(replay_bytes(n, psa_bits)) + 2 * (kq + kf);  // the counts live in the replay records
  else
    b += align256(radix_tmp_bytes(n)) + 2 * kp;
  return b;
}

PipeBufs carve_pipe(char*& cur, size_t n, int psa_bits, int out_mode) {
  PipeBufs b{};
  if (!psa_bits) return b;
  const size_t kq = align256(n * 8), kp = align256(n * 4), kf = align256(n);
  b.k0 = (uint64_t*)cur, cur += kq;
  b.k1 = (uint64_t*)cur, cur += kq;
  if (out_mode == HM_OUT_UNPERMUTE) {
    b.replay = cur, cur += align256(replay_bytes(n, psa_bits));
    b.rv = (uint64_t*)cur, cur += kq;
    b.rf = (uint8_t*)cur, cur += kf;
    b.vt = (uint64_t*)cur, cur += kq;
    b.ft = (uint8_t*)cur, cur += kf;
  } else {
    b.tmp = cur, cur += align256(radix_tmp_bytes(n));
    b.p0 = (uint32_t*)cur, cur += kp;
    b.p1 = (uint32_t*)cur, cur += kp;
  }
  return b;
}

int enqueue_pipeline(const TreeView& tv, const SearchLaunch& sl, int psa_bits, int out_mode, const uint64_t* dq,
                     size_t n, const uint32_t* dn, const PipeBufs& b, uint64_t* dvals, uint8_t* dfound,
                     uint32_t* dpos, hipStream_t s) {
  if (!psa_bits) {
    HM_TRY(launch_search(tv, sl, dq, nullptr, dvals, dfound, n, OUT_DIRECT, s, dn));
    return HM_OK;
  }
  int which = 0;
  // Results in input order may come from any bit window: sort by the psa_bits bits below the tree's
  // psa_top (the bits above it are equal for every key, so they would order nothing but misses outside
  // the key range).  PSA-order results keep the reference's window [64 - psa_bits, 64).
  const int top = out_mode == HM_OUT_PSA_ORDER ? 64 : std::max(std::min((int)tv.psa_top, 64), psa_bits);
  if (out_mode == HM_OUT_UNPERMUTE) {
    // keys-only sort recording its passes; search in PSA order (coalesced result writes); then the
    // passes run backwards to put the results in input order
    HM_TRY(radix_sort_keys_replay(dq, n, top - psa_bits, top, b.k0, b.k1, b.replay, s, &which, dn));
    const uint64_t* sk = which ? b.k1 : b.k0;
    SearchLaunch ss = sl;
    ss.sorted = true;
    HM_TRY(launch_search(tv, ss, sk, nullptr, b.rv, b.rf, n, OUT_DIRECT, s, dn));
    HM_TRY(unsort_results(b.rv, b.rf, n, psa_bits, b.replay, b.vt, b.ft, dvals, dfound, s, dn));
    return HM_OK;
  }
  HM_TRY(radix_sort_pos(dq, n, top - psa_bits, top, b.k0, b.p0, b.k1, b.p1, b.tmp, s, &which, dn));
  const uint64_t* sk = which ? b.k1 : b.k0;
  const uint32_t* sp = which ? b.p1 : b.p0;
  SearchLaunch ss = sl;
  ss.sorted = true;
  if (out_mode == HM_OUT_PSA_ORDER) {
    HM_TRY(launch_search(tv, ss, sk, sp, dvals, dfound, n, OUT_DIRECT, s, dn));
    HM_TRY(hipMemcpyAsync(dpos, sp, n * 4, hipMemcpyDeviceToDevice, s));
  } else {
    HM_TRY(launch_search(tv, ss, sk, sp, dvals, dfound, n, OUT_SCATTER, s, dn));
  }
  return HM_OK;
}

}  // namespace hm

namespace {

// Level sizes for n_leaves leaves; returns height.
int plan_levels(uint64_t n_leaves, int F, uint32_t* level_start, uint32_t* n_inner) {
  std::vector<uint64_t> sizes;
  uint64_t m = n_leaves;
  do {
    m = (m + F - 1) / F;
    sizes.push_back(m);
  } while (m > 1);
  std::reverse(sizes.begin(), sizes.end());
  if ((int)sizes.size() > MAX_LEVELS) return -1;
  uint64_t acc = 0;
  for (size_t l = 0; l < sizes.size(); ++l) {
    level_start[l] = (uint32_t)acc;
    acc += sizes[l];
  }
  for (size_t l = sizes.size(); l <= MAX_LEVELS; ++l) level_start[l] = (uint32_t)acc;
  *n_inner = (uint32_t)acc;
  return (int)sizes.size();
}

// Grow-only scratch of re-layouts and inner rebuilds (a hipFree synchronizes the device).
int ensure_rl(hm_tree* t, size_t bytes, hipStream_t s) {
  if (t->rl_bytes >= bytes) return HM_OK;
  const size_t want = std::max(bytes, t->rl_bytes + t->rl_bytes / 2);
  if (t->p.mem_budget && t->device_bytes() - t->rl_bytes + want > t->p.mem_budget)
    return set_err(HM_EBUDGET, "re-layout scratch of %llu bytes exceeds the device-memory budget",
                   (unsigned long long)want);
  if (t->rl_buf) {
    HM_TRY(hipStreamSynchronize(s));
    HM_TRY(hipFree(t->rl_buf));
    t->rl_buf = nullptr;
    t->rl_bytes = 0;
  }
  HM_TRY(hipMalloc(&t->rl_buf, want));
  t->rl_bytes = want;
  return HM_OK;
}

size_t inner_scratch_bytes(uint64_t n_leaves, int F, uint32_t n_inner) {
  return align256(std::max<uint64_t>(n_leaves, 1) * 8) + align256(std::max<uint64_t>((n_leaves + F - 1) / F, 1) * 8) +
         align256((size_t)n_inner * 4) + align256(scan_tmp_bytes(n_inner));
}

// Build every inner level and the prefix array of `leaves` into (inner, prefix), which must hold
// the n_inner nodes plan_levels gives; scratch (inner_scratch_bytes) comes from rl_buf at offset
// `scratch_off`.  Enqueued on s, no host synchronization.
int build_inner_into(hm_tree* t, const Leaf* leaves, uint64_t n_leaves, uint64_t* inner, uint32_t* prefix,
                     const Levels& lv, size_t scratch_off, hipStream_t s, bool mins_ready) {
  char* cur = (char*)t->rl_buf + scratch_off;
  uint64_t* minA = (uint64_t*)cur;
  cur += align256(std::max<uint64_t>(n_leaves, 1) * 8);
  uint64_t* minB = (uint64_t*)cur;
  cur += align256(std::max<uint64_t>((n_leaves + t->F - 1) / t->F, 1) * 8);
  uint32_t* nchild = (uint32_t*)cur;
  cur += align256((size_t)lv.n_inner * 4);
  void* scan_tmp = cur;
  if (!mins_ready) HM_TRY(leaf_min_keys(leaves, n_leaves, minA, s));
  uint64_t m_child = n_leaves;
  uint64_t *cmin = minA, *nmin = minB;
  for (int l = lv.height - 1; l >= 0; --l) {
    const uint64_t m = lv.level_start[l + 1] - lv.level_start[l];
    HM_TRY(build_level(cmin, m_child, t->F, inner + (size_t)lv.level_start[l] * t->NS, nmin,
                       nchild + lv.level_start[l], m, s));
    std::swap(cmin, nmin);
    m_child = m;
  }
  HM_TRY(exclusive_scan_u32(nchild, prefix, lv.n_inner, 1u, scan_tmp, s));
  HM_TRY(fold_prefix(inner, prefix, lv.n_inner, t->F, s));
  return HM_OK;
}

void set_levels(hm_tree* t, const Levels& lv) {
  t->height = lv.height;
  t->n_inner = lv.n_inner;
  for (int i = 0; i <= MAX_LEVELS; ++i) t->level_start[i] = lv.level_start[i];
  t->lds_levels = choose_lds_levels(t, t->p.lds_levels);
  t->layout_version++;  // captured pipelines hold the old pointers and level table
}

// Rebuild every inner level and the prefix array of the active leaves in place (build / load).
int rebuild_inner(hm_tree* t, hipStream_t s) {
  Levels lv;
  lv.height = plan_levels(t->n_leaves, t->F, lv.level_start, &lv.n_inner);
  if (lv.height < 0) return set_err(HM_EINVAL, "tree would exceed %d levels", MAX_LEVELS);
  if (lv.n_inner > t->inner_cap) {
    const uint64_t extra = (uint64_t)(lv.n_inner - t->inner_cap) * (t->NS * 8 + 4);
    if (t->p.mem_budget && t->device_bytes() + extra > t->p.mem_budget)
      return set_err(HM_EBUDGET, "inner levels need %llu more bytes than the budget allows",
                     (unsigned long long)extra);
    HM_TRY(hipStreamSynchronize(s));
    if (t->d_inner) HM_TRY(hipFree(t->d_inner));
    if (t->d_prefix) HM_TRY(hipFree(t->d_prefix));
    t->d_inner = nullptr;
    t->d_prefix = nullptr;
    const uint32_t cap = lv.n_inner + lv.n_inner / 8 + 8;
    HM_TRY(hipMalloc(&t->d_inner, (size_t)cap * t->NS * 8));
    HM_TRY(hipMalloc(&t->d_prefix, (size_t)cap * 4));
    t->inner_cap = cap;
  }
  int rc = ensure_rl(t, inner_scratch_bytes(t->n_leaves, t->F, lv.n_inner), s);
  if (rc) return rc;
  rc = build_inner_into(t, t->d_leaves, t->n_leaves, t->d_inner, t->d_prefix, lv, 0, s, false);
  if (rc) return rc;
  HM_TRY(hipStreamSynchronize(s));
  set_levels(t, lv);
  return HM_OK;
}

void free_tree(hm_tree* t) {
  if (!t) return;
  if (t->own) (void)hipStreamSynchronize(t->own);
  for (hipStream_t h : t->hs)
    if (h) (void)hipStreamSynchronize(h);
  for (auto& kv : t->ws) {
    if (kv.second.buf) (void)hipFree(kv.second.buf);
    if (kv.second.last_use) (void)hipEventDestroy(kv.second.last_use);
  }
  if (t->d_inner) (void)hipFree(t->d_inner);
  if (t->d_prefix) (void)hipFree(t->d_prefix);
  if (t->d_leaves) (void)hipFree(t->d_leaves);
  if (t->pend_k) (void)hipFree(t->pend_k);
  if (t->pend_v) (void)hipFree(t->pend_v);
  if (t->upd_buf) (void)hipFree(t->upd_buf);
  if (t->upd_host) (void)hipHostFree(t->upd_host);
  if (t->pin_stage) (void)hipHostFree(t->pin_stage);
  if (t->stage_ev) (void)hipEventDestroy(t->stage_ev);
  if (t->sb_leaves) (void)hipFree(t->sb_leaves);
  if (t->sb_inner) (void)hipFree(t->sb_inner);
  if (t->sb_prefix) (void)hipFree(t->sb_prefix);
  if (t->rl_buf) (void)hipFree(t->rl_buf);
  if (t->bl_child) (void)hipFree(t->bl_child);
  for (int k = 0; k < 2; ++k) {
    if (t->hs_stage[k]) (void)hipHostFree(t->hs_stage[k]);
    if (t->hs_ev[k]) (void)hipEventDestroy(t->hs_ev[k]);
    if (t->hs[k]) (void)hipStreamDestroy(t->hs[k]);
  }
  if (t->publish_ev) (void)hipEventDestroy(t->publish_ev);
  if (t->na_ev) (void)hipEventDestroy(t->na_ev);
  if (t->own) (void)hipStreamDestroy(t->own);
  delete t;
}

int check_params(const hm_params* p) {
  if (!p) return set_err(HM_EINVAL, "params is NULL");
  if (p->fanout != 32 && p->fanout != 64 && p->fanout != 128)
    return set_err(HM_EINVAL, "fanout %u not in {32, 64, 128}", p->fanout);
  if (p->leaf_fill < 1 || p->leaf_fill > 8) return set_err(HM_EINVAL, "leaf_fill %u not in 1..8", p->leaf_fill);
  if (p->lds_levels < -1 || p->lds_levels > 3) return set_err(HM_EINVAL, "lds_levels %d not in -1..3", p->lds_levels);
  if (p->group_size > 64 || (p->group_size & (p->group_size - 1)) || p->group_size > p->fanout)
    return set_err(HM_EINVAL, "group_size %u not 0 or a power of two in 1..min(64, fanout)", p->group_size);
  if (p->psa_bits > 32) return set_err(HM_EINVAL, "psa_bits %u > 32", p->psa_bits);
  return HM_OK;
}

constexpr uint64_t kMaxBatch = (1ull << 31) - 1;

}  // namespace

// ---------------------------------------------------------------------------------------------
extern "C" {

int hm_params_init(hm_params* p) {
  if (!p) return set_err(HM_EINVAL, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->fanout = 64;
  p->leaf_fill = 8;
  p->lds_levels = -1;
  p->group_size = 0;
  p->psa_bits = 16;
  p->sort_threshold = 1ull << 20;  // below this the sort passes cost more launches than they save
  p->mem_budget = 0;
  p->flags = 0;
  return HM_OK;
}

// An empty tree object (no layout yet) on the current device.
static int create_tree(const hm_params* p, hm_tree** out) {
  hm_tree* t = new hm_tree();
  t->p = *p;
  t->F = (int)p->fanout;
  t->fill = (int)p->leaf_fill;
  t->NS = node_slots(t->F);
  if (p->group_size) t->cfg_direct.group = t->cfg_psa.group = (int)p->group_size;
  int rc = HM_OK;
  if (hipGetDevice(&t->device) != hipSuccess)
    rc = set_err(HM_EDEVICE, "no HIP device");
  else if (hipStreamCreateWithFlags(&t->own, hipStreamNonBlocking) != hipSuccess)
    rc = set_err(HM_EDEVICE, "hipStreamCreate failed");
  else if (hipEventCreateWithFlags(&t->publish_ev, hipEventDisableTiming) != hipSuccess ||
           hipEventCreateWithFlags(&t->na_ev, hipEventDisableTiming) != hipSuccess)
    rc = set_err(HM_EDEVICE, "hipEventCreate failed");
  if (rc) {
    free_tree(t);
    return rc;
  }
  *out = t;
  return HM_OK;
}

int hm_tree_build(const uint64_t* keys, const uint64_t* vals, uint64_t n, const hm_params* p, hm_tree** out) {
  if (!out) return set_err(HM_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  if (n > 0 && (!keys || !vals)) return set_err(HM_EINVAL, "keys/vals is NULL with n = %llu", (unsigned long long)n);
  if (n > (1ull << 34)) return set_err(HM_EINVAL, "n = %llu too large", (unsigned long long)n);

  hm_tree* t = nullptr;
  rc = create_tree(p, &t);
  if (rc) return rc;
  auto fail = [&](int code) {
    free_tree(t);
    return code;
  };
  hipStream_t s = t->own;

  t->n_keys = n;
  t->n_leaves = std::max<uint64_t>(1, (n + t->fill - 1) / t->fill);
  t->leaf_cap = t->n_leaves;
  {
    uint32_t ls[MAX_LEVELS + 1], ni = 0;
    if (plan_levels(t->n_leaves, t->F, ls, &ni) < 0) return fail(set_err(HM_EINVAL, "too many levels"));
    const uint64_t need = (uint64_t)ni * (t->NS * 8 + 4) + t->leaf_cap * sizeof(Leaf);
    if (p->mem_budget && need > p->mem_budget)
      return fail(set_err(HM_EBUDGET, "tree needs %llu device bytes, budget is %llu", (unsigned long long)need,
                          (unsigned long long)p->mem_budget));
  }

  // stage input, sort (key, value) pairs, validate
  uint64_t *k_in = nullptr, *v_in = nullptr, *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
  void* tmp = nullptr;
  uint32_t* d_flags = nullptr;
  uint64_t* d_orand = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    for (void* q : {(void*)k_in, (void*)v_in, (void*)k0, (void*)v0, (void*)k1, (void*)v1, tmp, (void*)d_flags,
                    (void*)d_orand})
      if (q) (void)hipFree(q);
  };
#define BUILD_TRY(expr)            \
  do {                             \
    int r_ = [&]() -> int {        \
      HM_TRY(expr);                \
      return HM_OK;                \
    }();                           \
    if (r_) {                      \
      cleanup();                   \
      return fail(r_);             \
    }                              \
  } while (0)

  const size_t nb = std::max<uint64_t>(n, 1) * 8;
  const uint64_t* dk = keys;
  const uint64_t* dv = vals;
  BUILD_TRY(hipMalloc(&d_flags, 256));
  BUILD_TRY(hipMalloc(&d_orand, 256));
  BUILD_TRY(hipMemsetAsync(d_flags, 0, 256, s));
  if (n > 0 && !is_device_ptr(keys)) {
    BUILD_TRY(hipMalloc(&k_in, nb));
    BUILD_TRY(hipMemcpyAsync(k_in, keys, n * 8, hipMemcpyHostToDevice, s));
    dk = k_in;
  }
  if (n > 0 && !is_device_ptr(vals)) {
    BUILD_TRY(hipMalloc(&v_in, nb));
    BUILD_TRY(hipMemcpyAsync(v_in, vals, n * 8, hipMemcpyHostToDevice, s));
    dv = v_in;
  }
  if (n > 1 && !(p->flags & HM_FLAG_KEYS_SORTED)) {
    uint64_t orand_init[2] = {0ull, ~0ull}, orand[2];
    BUILD_TRY(hipMemcpyAsync(d_orand, orand_init, 16, hipMemcpyHostToDevice, s));
    BUILD_TRY(key_bit_summary(dk, n, d_orand, s));
    BUILD_TRY(hipMemcpyAsync(orand, d_orand, 16, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipStreamSynchronize(s));
    const uint64_t varying = orand[0] ^ orand[1];
    if (varying) {
      const int lo = (__builtin_ctzll(varying) / 8) * 8;
      const int hi = 64 - __builtin_clzll(varying);
      BUILD_TRY(hipMalloc(&k0, nb));
      BUILD_TRY(hipMalloc(&v0, nb));
      BUILD_TRY(hipMalloc(&k1, nb));
      BUILD_TRY(hipMalloc(&v1, nb));
      BUILD_TRY(hipMalloc(&tmp, radix_tmp_bytes(n)));
      int which = 0;
      BUILD_TRY(radix_sort_u64pay(dk, dv, n, lo, hi, k0, v0, k1, v1, tmp, s, &which));
      dk = which ? k1 : k0;
      dv = which ? v1 : v0;
    }
  }
  uint32_t flags = 0;
  uint64_t edge[2] = {~0ull, 0ull};
  BUILD_TRY(validate_sorted(dk, n, d_flags, s));
  BUILD_TRY(hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, s));
  if (n) {
    BUILD_TRY(hipMemcpyAsync(&edge[0], dk, 8, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(&edge[1], dk + n - 1, 8, hipMemcpyDeviceToHost, s));
  }
  BUILD_TRY(hipStreamSynchronize(s));
  if (flags & 1u) {
    cleanup();
    return fail(set_err(HM_EINVAL, "key 0xFFFFFFFFFFFFFFFF is reserved for padding"));
  }
  if (flags & 2u) {
    cleanup();
    return fail(set_err(HM_EINVAL, (p->flags & HM_FLAG_KEYS_SORTED) ? "keys flagged sorted are not strictly ascending"
                                                                     : "duplicate keys in build input"));
  }
  t->widen_key_range(edge[0], edge[1]);
  BUILD_TRY(hipMalloc(&t->d_leaves, t->leaf_cap * sizeof(Leaf)));
  BUILD_TRY(fill_leaves(dk, dv, n, t->fill, t->d_leaves, t->n_leaves, s));
  {
    int r = rebuild_inner(t, s);
    if (r) {
      cleanup();
      return fail(r);
    }
  }
  cleanup();
#undef BUILD_TRY
  *out = t;
  return HM_OK;
}

int hm_tree_destroy(hm_tree* t) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  free_tree(t);
  return HM_OK;
}

}  // extern "C"

namespace hm {
// Host-resident batches above one chunk: the reference streams fixed buckets through two CUDA streams,
// copy in -> SortPairs -> search -> copy out per bucket (H/regular_bpt/prefix-sort.cu:261-309).  Here
// chunk c runs on stream hs[c % 2] (copy in, the whole sort -> search -> un-sort pipeline, copy out), so
// the PCIe copies of one chunk overlap the other chunk's kernels and copies in the other direction.
// Pinned caller memory (hipHostMalloc, torch pin_memory) is copied by DMA in place; pageable memory is
// staged through the tree's pinned buffers by the calling thread (several threads for large copies)
// while the device works on the other chunk.
constexpr uint64_t kStreamChunk = 1ull << 22;

bool is_pinned_host(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

void par_copy(void* dst, const void* src, size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, hw), bytes >> 22);  // >= 4 MiB per thread
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes + nt - 1) / nt;
  std::vector<std::thread> th;
  for (size_t i = 1; i < nt; ++i) {
    const size_t o = i * per;
    if (o >= bytes) break;
    th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, std::min(per, bytes - o)); });
  }
  std::memcpy(dst, src, std::min(per, bytes));
  for (auto& x : th) x.join();
}

int search_host_stream_body(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                            const hm_search_opts* o, hipStream_t s) {
  const uint64_t C = kStreamChunk;
  SearchLaunch sl;
  int bits = 0, mode = 0;
  int rc = resolve_search(t, C, o, &sl, &bits, &mode);
  if (rc) return rc;
  const bool pinned = is_pinned_host(q) && is_pinned_host(out_vals) && is_pinned_host(out_found);
  for (int k = 0; k < 2; ++k) {
    if (!t->hs[k]) HM_TRY(hipStreamCreateWithFlags(&t->hs[k], hipStreamNonBlocking));
    if (!t->hs_ev[k]) HM_TRY(hipEventCreateWithFlags(&t->hs_ev[k], hipEventDisableTiming));
  }
  if (!pinned && t->hs_stage_n < C) {
    for (int k = 0; k < 2; ++k) {
      if (t->hs_stage[k]) HM_TRY(hipHostFree(t->hs_stage[k]));  // every earlier use was synchronized
      t->hs_stage[k] = nullptr;
    }
    t->hs_stage_n = 0;
    for (int k = 0; k < 2; ++k)
      if ((rc = host_alloc_mapped(&t->hs_stage[k], C * 17))) return rc;
    t->hs_stage_n = C;
  }
  // device buffers per stream: queries, values, found flags, then the pipeline's scratch
  const size_t kq = align256(C * 8), kf = align256(C);
  Workspace* w[2] = {nullptr, nullptr};
  for (int k = 0; k < 2; ++k)
    if ((rc = ensure_ws(t, t->hs[k], 2 * kq + kf + pipe_bytes(C, bits, mode), &w[k]))) return rc;
  // after the caller's earlier work on s and the latest publish
  Workspace* ws_s = nullptr;
  if ((rc = ensure_ws(t, s, 0, &ws_s))) return rc;
  HM_TRY(hipEventRecord(ws_s->last_use, s));
  for (int k = 0; k < 2; ++k) {
    HM_TRY(hipStreamWaitEvent(t->hs[k], ws_s->last_use, 0));
    if (t->publish_valid) HM_TRY(hipStreamWaitEvent(t->hs[k], t->publish_ev, 0));
  }
  const TreeView tv = t->view();
  const uint64_t chunks = (n + C - 1) / C;
  auto drain = [&](uint64_t c) -> int {  // pageable: results of chunk c from staging to the caller
    const int k = (int)(c & 1);
    const uint64_t off = c * C, m = std::min(C, n - off);
    HM_TRY(hipEventSynchronize(t->hs_ev[k]));
    const char* st = (const char*)t->hs_stage[k];
    par_copy(out_vals + off, st + C * 8, m * 8);
    par_copy(out_found + off, st + C * 16, m);
    return HM_OK;
  };
  for (uint64_t c = 0; c < chunks; ++c) {
    const int k = (int)(c & 1);
    const hipStream_t hk = t->hs[k];
    const uint64_t off = c * C, m = std::min(C, n - off);
    char* cur = (char*)w[k]->buf;
    uint64_t* dq = (uint64_t*)cur;
    uint64_t* dv = (uint64_t*)(cur + kq);
    uint8_t* df = (uint8_t*)(cur + 2 * kq);
    cur += 2 * kq + kf;
    PipeBufs pb = carve_pipe(cur, C, bits, mode);
    const void* src_q = q + off;
    void* dst_v = out_vals + off;
    void* dst_f = out_found + off;
    if (!pinned) {
      if (c >= 2 && (rc = drain(c - 2))) return rc;  // also frees staging slot k
      char* st = (char*)t->hs_stage[k];
      par_copy(st, q + off, m * 8);
      src_q = st;
      dst_v = st + C * 8;
      dst_f = st + C * 16;
    }
    HM_TRY(hipMemcpyAsync(dq, src_q, m * 8, hipMemcpyHostToDevice, hk));
    if ((rc = enqueue_pipeline(tv, sl, bits, mode, dq, m, nullptr, pb, dv, df, nullptr, hk))) return rc;
    HM_TRY(hipMemcpyAsync(dst_v, dv, m * 8, hipMemcpyDeviceToHost, hk));
    HM_TRY(hipMemcpyAsync(dst_f, df, m, hipMemcpyDeviceToHost, hk));
    HM_TRY(hipEventRecord(t->hs_ev[k], hk));
    HM_TRY(hipEventRecord(w[k]->last_use, hk));
  }
  if (!pinned)
    for (uint64_t c = chunks > 2 ? chunks - 2 : 0; c < chunks; ++c)
      if ((rc = drain(c))) return rc;
  for (int k = 0; k < 2; ++k) HM_TRY(hipStreamSynchronize(t->hs[k]));
  return HM_OK;
}

int search_host_stream(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                       const hm_search_opts* o, hipStream_t s) {
  const int rc = search_host_stream_body(t, q, n, out_vals, out_found, o, s);
  if (rc)  // a failed call leaves no copy in flight into the staging slots or the caller's buffers
    for (hipStream_t h : t->hs)
      if (h) (void)hipStreamSynchronize(h);
  return rc;
}

int search_impl(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                const hm_search_opts* o, hipStream_t s, bool mapped) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (n == 0) return HM_OK;
  if (!q || !out_vals || !out_found) return set_err(HM_EINVAL, "NULL buffer");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "batch of %llu queries exceeds 2^31-1", (unsigned long long)n);
  std::lock_guard<std::mutex> g(t->mu);

  SearchLaunch sl;
  int bits = 0, out_mode = 0;
  int rc = resolve_search(t, n, o, &sl, &bits, &out_mode);
  if (rc) return rc;
  const bool psa_order = out_mode == HM_OUT_PSA_ORDER;
  const bool hq = !mapped && !is_device_ptr(q), hv = !mapped && !is_device_ptr(out_vals),
             hf = !mapped && !is_device_ptr(out_found);
  if (hq && hv && hf && !psa_order && n > kStreamChunk) return search_host_stream(t, q, n, out_vals, out_found, o, s);
  const bool hp = psa_order && !mapped && !is_device_ptr(o->out_pos);
  const size_t kq = align256(n * 8), kp = align256(n * 4);
  size_t need = pipe_bytes(n, bits, out_mode);
  if (hq) need += kq;
  if (hv) need += kq;
  if (hf) need += align256(n);
  if (hp) need += kp;
  Workspace* w = nullptr;
  rc = ensure_ws(t, s, need, &w);
  if (rc) return rc;
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(s, t->publish_ev, 0));
  char* cur = (char*)w->buf;
  const uint64_t* dq = q;
  uint64_t* dvals = out_vals;
  uint8_t* dfound = out_found;
  uint32_t* dpos = psa_order ? o->out_pos : nullptr;
  if (hq) {
    uint64_t* b = (uint64_t*)cur;
    cur += kq;
    HM_TRY(hipMemcpyAsync(b, q, n * 8, hipMemcpyHostToDevice, s));
    dq = b;
  }
  if (hv) dvals = (uint64_t*)cur, cur += kq;
  if (hf) dfound = (uint8_t*)cur, cur += align256(n);
  if (hp) dpos = (uint32_t*)cur, cur += kp;
  PipeBufs pb = carve_pipe(cur, n, bits, out_mode);
  rc = enqueue_pipeline(t->view(), sl, bits, out_mode, dq, n, nullptr, pb, dvals, dfound, dpos, s);
  if (rc) return rc;
  if (hv) HM_TRY(hipMemcpyAsync(out_vals, dvals, n * 8, hipMemcpyDeviceToHost, s));
  if (hf) HM_TRY(hipMemcpyAsync(out_found, dfound, n, hipMemcpyDeviceToHost, s));
  if (hp) HM_TRY(hipMemcpyAsync(o->out_pos, dpos, n * 4, hipMemcpyDeviceToHost, s));
  HM_TRY(hipEventRecord(w->last_use, s));
  if (hq || hv || hf || hp) HM_TRY(hipStreamSynchronize(s));
  return HM_OK;
}
}  // namespace hm

extern "C" {

int hm_search_ex(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                 const hm_search_opts* o, void* stream) {
  return search_impl(t, q, n, out_vals, out_found, o, (hipStream_t)stream, false);
}

int hm_search(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found, void* stream) {
  return hm_search_ex(t, q, n, out_vals, out_found, nullptr, stream);
}

int hm_baseline_search(hm_tree* t, int kind, uint32_t group_size, const uint64_t* q, uint64_t n, uint64_t* out_vals,
                       uint8_t* out_found, void* stream) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (kind != HM_BASELINE_POINTER_CHILD && kind != HM_BASELINE_POINTER_LINEAR && kind != HM_BASELINE_PREFIX_PLAIN)
    return set_err(HM_EINVAL, "unknown baseline kind %d", kind);
  if (group_size != 1 && group_size != 2 && group_size != 4 && group_size != 8)
    return set_err(HM_EINVAL, "group_size %u not in {1, 2, 4, 8}", group_size);
  if (n == 0) return HM_OK;
  if (!q || !out_vals || !out_found) return set_err(HM_EINVAL, "NULL buffer");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "batch of %llu queries exceeds 2^31-1", (unsigned long long)n);
  if (!is_device_ptr(q) || !is_device_ptr(out_vals) || !is_device_ptr(out_found))
    return set_err(HM_EINVAL, "baseline search takes device buffers only");
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(t->mu);
  const size_t need = (size_t)t->n_inner * t->F;
  if (need > t->bl_cap) {
    if (t->p.mem_budget && t->device_bytes() + (need - t->bl_cap) * 4 > t->p.mem_budget)
      return set_err(HM_EBUDGET, "baseline child array exceeds the device-memory budget");
    if (t->bl_child) {
      HM_TRY(hipDeviceSynchronize());  // searches on any stream may still read it
      HM_TRY(hipFree(t->bl_child));
      t->bl_child = nullptr;
      t->bl_cap = 0;
    }
    HM_TRY(hipMalloc(&t->bl_child, need * 4));
    t->bl_cap = need;
    t->bl_version = ~0ull;
  }
  Workspace* w = nullptr;
  int rc = ensure_ws(t, s, 0, &w);  // registers the stream, so a publish waits for these reads
  if (rc) return rc;
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(s, t->publish_ev, 0));
  if (t->bl_version != t->layout_version || t->bl_epoch != t->epoch) {
    HM_TRY(build_child_array(t->d_prefix, t->n_inner, t->F, t->bl_child, s));
    t->bl_version = t->layout_version;
    t->bl_epoch = t->epoch;
  }
  const int k = kind == HM_BASELINE_POINTER_CHILD ? 0 : kind == HM_BASELINE_POINTER_LINEAR ? 1 : 2;
  HM_TRY(launch_search_baseline(t->view(), t->bl_child, q, out_vals, out_found, (uint32_t)n, (int)group_size, k, s));
  HM_TRY(hipEventRecord(w->last_use, s));
  return HM_OK;
}

}  // extern "C"

namespace hm {
size_t range_ws_bytes(size_t n) { return align256((n + 1) * 8) + align256(scan_tmp_bytes(n + 1)); }

static int range_common(hm_tree* t, const uint64_t* start, const uint64_t* hi, const uint32_t* len, uint32_t cap_const,
                        uint64_t n, uint64_t* out_offsets, uint64_t* out_keys, uint64_t* out_vals,
                        uint32_t* out_counts, uint64_t capacity, hipStream_t s, bool mapped) {
  std::lock_guard<std::mutex> g(t->mu);
  auto host = [&](const void* p) { return p && !mapped && !is_device_ptr(p); };
  const bool hs = host(start), hh = host(hi), hl = host(len);
  const bool ho = host(out_offsets), hk = host(out_keys), hv = host(out_vals), hc = host(out_counts);
  const size_t k8 = align256(n * 8), k8p = align256((n + 1) * 8), k4 = align256(n * 4);
  const size_t cap_bytes = align256(capacity * 8);
  size_t need = (hs ? k8 : 0) + (hh ? k8 : 0) + (hl ? k4 : 0) + (out_offsets ? k8p : 0) + (hk ? cap_bytes : 0) +
                (hv ? cap_bytes : 0) + (hc ? k4 : 0) + align256(scan_tmp_bytes(n + 1));
  Workspace* w = nullptr;
  int rc = ensure_ws(t, s, need, &w);
  if (rc) return rc;
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(s, t->publish_ev, 0));
  char* cur = (char*)w->buf;
  auto take = [&](size_t b) {
    char* r = cur;
    cur += b;
    return r;
  };
  const uint64_t* ds = start;
  const uint64_t* dh = hi;
  const uint32_t* dl = len;
  if (hs) {
    uint64_t* b = (uint64_t*)take(k8);
    HM_TRY(hipMemcpyAsync(b, start, n * 8, hipMemcpyHostToDevice, s));
    ds = b;
  }
  if (hh) {
    uint64_t* b = (uint64_t*)take(k8);
    HM_TRY(hipMemcpyAsync(b, hi, n * 8, hipMemcpyHostToDevice, s));
    dh = b;
  }
  if (hl) {
    uint32_t* b = (uint32_t*)take(k4);
    HM_TRY(hipMemcpyAsync(b, len, n * 4, hipMemcpyHostToDevice, s));
    dl = b;
  }
  uint64_t* doff = nullptr;
  if (out_offsets) {
    doff = ho ? (uint64_t*)take(k8p) : out_offsets;
    void* scan_tmp = take(align256(scan_tmp_bytes(n + 1)));
    HM_TRY(exclusive_scan_u32_to_u64(dl, doff, n, scan_tmp, s));
    HM_TRY(range_offsets_tail(dl, doff, n, s));
  }
  uint64_t* dk = hk ? (uint64_t*)take(cap_bytes) : out_keys;
  uint64_t* dv = hv ? (uint64_t*)take(cap_bytes) : out_vals;
  uint32_t* dc = hc ? (uint32_t*)take(k4) : out_counts;
  HM_TRY(launch_range(t->view(), ds, dh, dl, cap_const, n, doff, cap_const, dk, dv, dc, capacity, s));
  if (ho) HM_TRY(hipMemcpyAsync(out_offsets, doff, (n + 1) * 8, hipMemcpyDeviceToHost, s));
  if (hk) HM_TRY(hipMemcpyAsync(out_keys, dk, capacity * 8, hipMemcpyDeviceToHost, s));
  if (hv) HM_TRY(hipMemcpyAsync(out_vals, dv, capacity * 8, hipMemcpyDeviceToHost, s));
  if (hc) HM_TRY(hipMemcpyAsync(out_counts, dc, n * 4, hipMemcpyDeviceToHost, s));
  HM_TRY(hipEventRecord(w->last_use, s));
  if (hs || hh || hl || ho || hk || hv || hc) HM_TRY(hipStreamSynchronize(s));
  return HM_OK;
}

int range_impl(hm_tree* t, const uint64_t* start, const uint32_t* len, uint64_t n, uint64_t* out_offsets,
               uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, hipStream_t s,
               bool mapped) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (n == 0) {
    if (out_offsets) {
      uint64_t z = 0;
      if (!mapped && is_device_ptr(out_offsets))
        HM_TRY(hipMemcpy(out_offsets, &z, 8, hipMemcpyHostToDevice));
      else
        out_offsets[0] = 0;
    }
    return HM_OK;
  }
  if (!start || !len || !out_offsets || !out_keys || !out_vals || !out_counts)
    return set_err(HM_EINVAL, "NULL buffer");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "too many ranges");
  return range_common(t, start, nullptr, len, 0, n, out_offsets, out_keys, out_vals, out_counts, capacity, s, mapped);
}
}  // namespace hm

extern "C" {

int hm_range(hm_tree* t, const uint64_t* start, const uint32_t* len, uint64_t n, uint64_t* out_offsets,
             uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts, uint64_t capacity, void* stream) {
  return range_impl(t, start, len, n, out_offsets, out_keys, out_vals, out_counts, capacity, (hipStream_t)stream,
                    false);
}

int hm_range_between(hm_tree* t, const uint64_t* lo, const uint64_t* hi, uint64_t n, uint32_t cap, uint64_t* out_keys,
                     uint64_t* out_vals, uint32_t* out_counts, void* stream) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (n == 0) return HM_OK;
  if (!lo || !hi || !out_keys || !out_vals || !out_counts) return set_err(HM_EINVAL, "NULL buffer");
  if (cap == 0) return set_err(HM_EINVAL, "cap must be > 0");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "too many ranges");
  return range_common(t, lo, hi, nullptr, cap, n, nullptr, out_keys, out_vals, out_counts, n * (uint64_t)cap,
                      (hipStream_t)stream, false);
}

// Device scratch of one publish over n staged updates (carved from t->upd_buf).
struct UpdBufs {
  uint64_t *uk, *uv;          // sorted unique batch
  uint64_t *k0, *v0, *k1, *v1;  // radix ping-pong (large batches)
  void* rtmp;
  uint32_t *keep, *keep_scan;
  void* stmp;
  uint32_t *leaf, *slot;
  uint32_t* nlist;  // nodes to re-spread: (node, first batch entry) pairs
  uint8_t* found;
  uint32_t* summary;  // [0] overflow, [1] new keys, [2] unique count, [3] validation flags, [4] nodes listed
};

static size_t upd_need(size_t n) {
  const size_t k8 = align256(n * 8), k4 = align256((n + 1) * 4);
  return 6 * k8 + align256(radix_tmp_bytes(n)) + 2 * k4 + align256(scan_tmp_bytes(n + 1)) + 4 * k4 + align256(n) +
         256;
}

static UpdBufs carve_upd(void* base, size_t n) {
  char* cur = (char*)base;
  const size_t k8 = align256(n * 8), k4 = align256((n + 1) * 4);
  UpdBufs b;
  b.summary = (uint32_t*)cur, cur += 256;
  b.uk = (uint64_t*)cur, cur += k8;
  b.uv = (uint64_t*)cur, cur += k8;
  b.k0 = (uint64_t*)cur, cur += k8;
  b.v0 = (uint64_t*)cur, cur += k8;
  b.k1 = (uint64_t*)cur, cur += k8;
  b.v1 = (uint64_t*)cur, cur += k8;
  b.rtmp = cur, cur += align256(radix_tmp_bytes(n));
  b.keep = (uint32_t*)cur, cur += k4;
  b.keep_scan = (uint32_t*)cur, cur += k4;
  b.stmp = cur, cur += align256(scan_tmp_bytes(n + 1));
  b.leaf = (uint32_t*)cur, cur += k4;
  b.slot = (uint32_t*)cur, cur += k4;
  b.nlist = (uint32_t*)cur, cur += 2 * k4;
  b.found = (uint8_t*)cur;
  return b;
}

// Grow-only: a publish in steady state allocates nothing (hipFree would synchronize the device).
static int ensure_upd(hm_tree* t, size_t n, hipStream_t s) {
  if (!t->upd_host) {
    int rc = host_alloc_mapped((void**)&t->upd_host, 256);
    if (rc) return rc;
  }
  const size_t need = upd_need(std::max<size_t>(n, 1));
  if (t->upd_bytes >= need) return HM_OK;
  const size_t want = std::max(need, std::min<size_t>(2 * t->upd_bytes, upd_need(kMaxBatch)));
  if (t->p.mem_budget && t->device_bytes() - t->upd_bytes + want > t->p.mem_budget)
    return set_err(HM_EBUDGET, "update scratch for %llu keys exceeds the device-memory budget", (unsigned long long)n);
  if (t->upd_buf) {
    HM_TRY(hipStreamSynchronize(s));
    HM_TRY(hipFree(t->upd_buf));
    t->upd_buf = nullptr;
    t->upd_bytes = 0;
  }
  HM_TRY(hipMalloc(&t->upd_buf, want));
  t->upd_bytes = want;
  return HM_OK;
}

}  // extern "C"

namespace hm {
// Pinned bounce buffer for host update batches (keys and values of up to cap staged updates).
static int ensure_pin(hm_tree* t, size_t cap) {
  if (t->pin_cap >= cap) return HM_OK;
  if (t->pin_stage) {
    if (t->stage_valid) HM_TRY(hipEventSynchronize(t->stage_ev));  // its last copies are done
    HM_TRY(hipHostFree(t->pin_stage));
    t->pin_stage = nullptr;
    t->pin_cap = 0;
  }
  int rc = host_alloc_mapped((void**)&t->pin_stage, cap * 16);
  if (rc) return rc;
  t->pin_cap = cap;
  return HM_OK;
}

// Stage an upsert batch; the caller holds t->pub_mu.
int stage_updates(hm_tree* t, const uint64_t* keys, const uint64_t* vals, uint64_t n, hipStream_t s) {
  if (n == 0) return HM_OK;
  if (!keys || !vals) return set_err(HM_EINVAL, "NULL buffer");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "update batch too large");
  const bool dk = is_device_ptr(keys), dv = is_device_ptr(vals);
  if (!dk) {  // host keys: check for the reserved key before anything touches the device
    for (uint64_t i = 0; i < n; ++i)
      if (keys[i] == HM_KEY_MAX) return set_err(HM_EINVAL, "key 0xFFFFFFFFFFFFFFFF is reserved for padding");
  }
  if (t->pend_n + n > kMaxBatch) return set_err(HM_EINVAL, "too many staged updates; publish first");
  if (t->pend_n + n > t->pend_cap) {
    size_t cap = std::max<size_t>(t->pend_n + n, 2 * t->pend_cap);
    if (cap < 65536) cap = 65536;
    if (t->p.mem_budget && t->device_bytes() + (cap - t->pend_cap) * 16 > t->p.mem_budget)
      return set_err(HM_EBUDGET, "staging %llu updates exceeds the device-memory budget", (unsigned long long)n);
    uint64_t *nk = nullptr, *nv = nullptr;
    HM_TRY(hipMalloc(&nk, cap * 8));
    HM_TRY(hipMalloc(&nv, cap * 8));
    if (t->pend_n) {
      if (t->stage_valid) HM_TRY(hipStreamWaitEvent(s, t->stage_ev, 0));  // staged on other streams
      HM_TRY(copy_words(nk, t->pend_k, t->pend_n, s));
      HM_TRY(copy_words(nv, t->pend_v, t->pend_n, s));
    }
    HM_TRY(hipStreamSynchronize(s));
    if (t->pend_k) HM_TRY(hipFree(t->pend_k));
    if (t->pend_v) HM_TRY(hipFree(t->pend_v));
    t->pend_k = nk;
    t->pend_v = nv;
    t->pend_cap = cap;
  }
  if (!t->stage_ev) HM_TRY(hipEventCreateWithFlags(&t->stage_ev, hipEventDisableTiming));
  if (!dk || !dv) {
    // host sources go through mapped pinned memory that a copy kernel reads: no DMA-engine copy
    // (those serialize across every stream of the process) and no pageable copy that blocks.
    // Slots [pend_n, pend_n + n) are free: a publish finishes every staging copy before it resets
    // pend_n.
    int rc = ensure_pin(t, t->pend_cap);
    if (rc) return rc;
    if (!dk) {
      std::memcpy(t->pin_stage + t->pend_n, keys, n * 8);
      keys = t->pin_stage + t->pend_n;
    }
    if (!dv) {
      std::memcpy(t->pin_stage + t->pin_cap + t->pend_n, vals, n * 8);
      vals = t->pin_stage + t->pin_cap + t->pend_n;
    }
  }
  // chain the staging streams, so stage_ev covers every batch staged so far
  if (t->stage_valid) HM_TRY(hipStreamWaitEvent(s, t->stage_ev, 0));
  HM_TRY(copy_words(t->pend_k + t->pend_n, keys, n, s));
  HM_TRY(copy_words(t->pend_v + t->pend_n, vals, n, s));
  HM_TRY(hipEventRecord(t->stage_ev, s));
  t->stage_valid = true;
  if (dk) {  // device keys: check on the device (validate_sorted sets bit 0 for KEY_MAX; bit 1 is ignored)
    int rc = ensure_upd(t, t->pend_n + n, s);
    if (rc) return rc;
    UpdBufs b = carve_upd(t->upd_buf, t->pend_n + n);
    HM_TRY(hipMemsetAsync(b.summary + 3, 0, 4, s));
    HM_TRY(validate_sorted(t->pend_k + t->pend_n, n, b.summary + 3, s));
    HM_TRY(copy_u32s(t->upd_host + 3, b.summary + 3, 1, s));
    HM_TRY(hipStreamSynchronize(s));
    if (t->upd_host[3] & 1u) return set_err(HM_EINVAL, "key 0xFFFFFFFFFFFFFFFF is reserved for padding");
  }
  t->pend_n += n;
  return HM_OK;
}
}  // namespace hm

extern "C" {

int hm_update(hm_tree* t, const uint64_t* keys, const uint64_t* vals, uint64_t n, void* stream) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  std::lock_guard<std::mutex> g(t->pub_mu);
  return stage_updates(t, keys, vals, n, (hipStream_t)stream);
}

// Make s wait for every search / range issued so far (before leaves change in place).
static int wait_for_readers(hm_tree* t, hipStream_t s) {
  for (auto& kv : t->ws)
    if (kv.second.last_use) HM_TRY(hipStreamWaitEvent(s, kv.second.last_use, 0));
  if (t->na_valid) HM_TRY(hipStreamWaitEvent(s, t->na_ev, 0));  // the last allocation-free search
  return HM_OK;
}

}  // extern "C"

namespace hm {

// Prepare the staged batch on s without changing what searches read: sort and de-duplicate it,
// locate its keys and plan the epoch (one host round trip).  If some leaf overflows, build the
// whole new layout (leaves with the batch merged, inner levels, prefix array) into the standby
// buffers.  Searches on other streams keep running on the active layout meanwhile.  The caller
// holds t->pub_mu until publish_commit.
int publish_prepare(hm_tree* t, hipStream_t s, PublishPlan* plan) {
  *plan = PublishPlan{};
  const size_t n = t->pend_n;
  plan->n = n;
  if (!n) return HM_OK;
  int rc = ensure_upd(t, n, s);
  if (rc) return rc;
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(s, t->publish_ev, 0));  // the previous epoch's writes
  if (t->stage_valid) HM_TRY(hipStreamWaitEvent(s, t->stage_ev, 0));       // staging on other streams
  const UpdBufs b = carve_upd(t->upd_buf, n);
  uint32_t* dm = b.summary + 2;
  HM_TRY(hipMemsetAsync(b.summary, 0, 32, s));
  // sort by key (stable: a later write of a key follows the earlier one), keep the last write
  if (n <= kSmallUpdate) {
    HM_TRY(small_sort_dedupe(t->pend_k, t->pend_v, (uint32_t)n, b.uk, b.uv, dm, s));
  } else {
    const uint64_t* sk;
    const uint64_t* sv;
    if (n <= kMergeUpdate) {  // the index ping-pong uses keep / keep_scan before they are needed
      HM_TRY(merge_sort_pairs(t->pend_k, t->pend_v, (uint32_t)n, b.k0, b.k1, b.keep, b.keep_scan, b.v0, s, &sk));
      sv = b.v0;
    } else {
      int which = 0;
      HM_TRY(radix_sort_u64pay(t->pend_k, t->pend_v, n, 0, 64, b.k0, b.v0, b.k1, b.v1, b.rtmp, s, &which));
      sk = which ? b.k1 : b.k0;
      sv = which ? b.v1 : b.v0;
    }
    HM_TRY(mark_last_of_runs(sk, n, b.keep, s));
    HM_TRY(exclusive_scan_u32(b.keep, b.keep_scan, n, 0u, b.stmp, s));
    HM_TRY(compact_pairs(sk, sv, b.keep, b.keep_scan, n, b.uk, b.uv, s));
    HM_TRY(unique_count(b.keep, b.keep_scan, n, dm, s));
  }
  const TreeView tv = t->view();
  HM_TRY(launch_locate(tv, b.uk, n, b.leaf, b.slot, b.found, s, dm));
  HM_TRY(plan_runs(t->d_leaves, (uint32_t)t->n_leaves, t->F, b.leaf, b.found, n, dm, b.summary, b.nlist, s));
  HM_TRY(copy_u32s(t->upd_host, b.summary, 5, s));
  uint64_t* edge = reinterpret_cast<uint64_t*>(t->upd_host + 8);  // bytes 32..47 of the pinned summary
  HM_TRY(edge_keys(b.uk, dm, edge, s));
  HM_TRY(hipStreamSynchronize(s));
  plan->key_lo = edge[0];
  plan->key_hi = edge[1];
  plan->n_new = t->upd_host[1];
  plan->m = t->upd_host[2];
  // an overflowing leaf whose last-level node has room: that node is re-spread in place at commit
  // (bounded cost); only when some node is full do all leaves get re-laid out
  plan->relayout = (t->upd_host[0] & 2u) != 0;
  if (!plan->relayout) {
    plan->respread = (t->upd_host[0] & 1u) ? t->upd_host[4] : 0u;
    if (plan->respread) HM_TRY(mark_nodes(b.nlist, plan->respread, t->F, b.leaf, b.found, dm, s));
    return HM_OK;
  }

  // ---- re-layout into the standby buffers ----
  ProfScope prof("relayout", s);
  const uint64_t nl = t->n_leaves;
  const size_t a4 = align256((nl + 1) * 4);
  {
    // size the scratch for the counts and the inner rebuild at once (growing it later means a
    // device-wide free and a copy): a touched leaf becomes at most 8 more leaves (fill 1), so
    // nl + 8 * n_new bounds the new leaf count
    const uint64_t bound = nl + 8 * (uint64_t)plan->n_new;
    uint32_t ls[MAX_LEVELS + 1], ni = 0;
    const int h = plan_levels(bound, t->F, ls, &ni);
    const size_t counts = align256(4 * a4 + align256(scan_tmp_bytes(nl + 1)));
    rc = ensure_rl(t, counts + (h >= 0 ? inner_scratch_bytes(bound, t->F, ni) : 0), s);
    if (rc == HM_EBUDGET) rc = ensure_rl(t, counts, s);  // the bound is loose: try the exact size later
  }
  if (rc) return rc;
  char* cur = (char*)t->rl_buf;
  uint32_t* ins = (uint32_t*)cur;
  uint32_t* first = (uint32_t*)(cur + a4);
  uint32_t* cnt = (uint32_t*)(cur + 2 * a4);
  uint32_t* nstart = (uint32_t*)(cur + 3 * a4);
  void* ntmp = cur + 4 * a4;
  HM_TRY(hipMemsetAsync(ins, 0, nl * 4, s));
  HM_TRY(hipMemsetAsync(first, 0xFF, nl * 4, s));
  HM_TRY(count_inserts(b.leaf, b.found, plan->m, ins, s));
  HM_TRY(leaf_first_keys(b.leaf, plan->m, first, s));
  HM_TRY(new_leaf_counts(t->d_leaves, ins, nl, t->fill, cnt, s));
  HM_TRY(exclusive_scan_u32(cnt, nstart, nl, 0u, ntmp, s));
  HM_TRY(copy_u32s(t->upd_host + 4, nstart + nl - 1, 1, s));
  HM_TRY(copy_u32s(t->upd_host + 5, cnt + nl - 1, 1, s));
  HM_TRY(hipStreamSynchronize(s));
  const uint64_t new_leaves = (uint64_t)t->upd_host[4] + t->upd_host[5];
  if (new_leaves >= (1ull << 32)) return set_err(HM_EINVAL, "leaf count would exceed 2^32");
  Levels lv;
  lv.height = plan_levels(new_leaves, t->F, lv.level_start, &lv.n_inner);
  if (lv.height < 0) return set_err(HM_EINVAL, "tree would exceed %d levels", MAX_LEVELS);
  // standby capacity (grow-only, with slack so the next re-layouts fit)
  const uint64_t leaf_cap = t->sb_leaf_cap >= new_leaves ? t->sb_leaf_cap : new_leaves + new_leaves / 16 + 64;
  const uint32_t inner_cap = t->sb_inner_cap >= lv.n_inner ? t->sb_inner_cap : lv.n_inner + lv.n_inner / 8 + 8;
  const uint64_t grow = (leaf_cap - t->sb_leaf_cap) * sizeof(Leaf) + (uint64_t)(inner_cap - t->sb_inner_cap) * (t->NS * 8 + 4);
  const size_t ib_off = align256(4 * a4 + align256(scan_tmp_bytes(nl + 1)));
  const size_t ib_need = ib_off + inner_scratch_bytes(new_leaves, t->F, lv.n_inner);
  if (t->p.mem_budget &&
      t->device_bytes() + grow + (ib_need > t->rl_bytes ? ib_need - t->rl_bytes : 0) > t->p.mem_budget)
    return set_err(HM_EBUDGET, "re-layout to %llu leaves exceeds the device-memory budget",
                   (unsigned long long)new_leaves);
  if (ib_need > t->rl_bytes) {  // not expected (sized above); keep the counts: grow and copy them over
    void* nb = nullptr;
    HM_TRY(hipMalloc(&nb, ib_need));
    HM_TRY(copy_words((uint64_t*)nb, (const uint64_t*)t->rl_buf, ib_off / 8, s));
    HM_TRY(hipStreamSynchronize(s));
    HM_TRY(hipFree(t->rl_buf));
    t->rl_buf = nb;
    t->rl_bytes = ib_need;
    cur = (char*)t->rl_buf;
    ins = (uint32_t*)cur;
    first = (uint32_t*)(cur + a4);
    nstart = (uint32_t*)(cur + 3 * a4);
  }
  if (leaf_cap != t->sb_leaf_cap) {
    HM_TRY(hipStreamSynchronize(s));
    if (t->sb_leaves) HM_TRY(hipFree(t->sb_leaves));
    t->sb_leaves = nullptr;
    t->sb_leaf_cap = 0;
    HM_TRY(hipMalloc(&t->sb_leaves, leaf_cap * sizeof(Leaf)));
    t->sb_leaf_cap = leaf_cap;
  }
  if (inner_cap != t->sb_inner_cap) {
    HM_TRY(hipStreamSynchronize(s));
    if (t->sb_inner) HM_TRY(hipFree(t->sb_inner));
    if (t->sb_prefix) HM_TRY(hipFree(t->sb_prefix));
    t->sb_inner = nullptr;
    t->sb_prefix = nullptr;
    t->sb_inner_cap = 0;
    HM_TRY(hipMalloc(&t->sb_inner, (size_t)inner_cap * t->NS * 8));
    HM_TRY(hipMalloc(&t->sb_prefix, (size_t)inner_cap * 4));
    t->sb_inner_cap = inner_cap;
  }
  // the standby arrays were the active ones before the last swap: searches issued before it may
  // still read them
  {
    std::lock_guard<std::mutex> g(t->mu);  // the workspace map changes under mu
    rc = wait_for_readers(t, s);
  }
  if (rc) return rc;
  // the leaf minima go straight to the inner rebuild's scratch (its first array)
  uint64_t* new_mins = (uint64_t*)((char*)t->rl_buf + ib_off);
  HM_TRY(relayout_touched(t->d_leaves, ins, b.leaf, b.slot, b.uk, b.uv, b.found, plan->m, nstart, t->fill,
                          t->sb_leaves, new_mins, s));
  // the streaming copy goes in paced chunks of 256 MiB (~0.1 ms): a search whose stream shares a
  // hardware queue with this one then waits behind one chunk, not the whole copy
  constexpr uint64_t kChunk = 1ull << 21;
  for (uint64_t lo = 0; lo < nl; lo += kChunk) {
    const uint64_t hi = std::min<uint64_t>(nl, lo + kChunk);
    HM_TRY(relayout_copy(t->d_leaves, lo, hi, first, nstart, t->sb_leaves, new_mins, s));
    if (hi < nl) HM_TRY(hipStreamSynchronize(s));
  }
  rc = build_inner_into(t, t->sb_leaves, new_leaves, t->sb_inner, t->sb_prefix, lv, ib_off, s, true);
  if (rc) return rc;
  HM_TRY(hipStreamSynchronize(s));
  plan->new_leaves = new_leaves;
  plan->lv = lv;
  return HM_OK;
}

// Size the update staging and publish scratch for batches of up to n keys now (grow-only).
int reserve_updates(hm_tree* t, size_t n, hipStream_t s) {
  std::lock_guard<std::mutex> pub(t->pub_mu);
  int rc = ensure_upd(t, n, s);
  if (rc) return rc;
  if (t->pend_cap < n) {
    if (t->p.mem_budget && t->device_bytes() + (n - t->pend_cap) * 16 > t->p.mem_budget)
      return set_err(HM_EBUDGET, "update staging exceeds the device-memory budget");
    uint64_t *nk = nullptr, *nv = nullptr;
    HM_TRY(hipMalloc(&nk, n * 8));
    HM_TRY(hipMalloc(&nv, n * 8));
    if (t->pend_n) {
      HM_TRY(hipMemcpyAsync(nk, t->pend_k, t->pend_n * 8, hipMemcpyDeviceToDevice, s));
      HM_TRY(hipMemcpyAsync(nv, t->pend_v, t->pend_n * 8, hipMemcpyDeviceToDevice, s));
    }
    HM_TRY(hipStreamSynchronize(s));
    if (t->pend_k) HM_TRY(hipFree(t->pend_k));
    if (t->pend_v) HM_TRY(hipFree(t->pend_v));
    t->pend_k = nk;
    t->pend_v = nv;
    t->pend_cap = n;
  }
  return ensure_pin(t, t->pend_cap);
}

// Allocate the standby layout and re-layout scratch ahead of time (a serving loop does this at
// start, so the first overflow does not pay for multi-GB allocations).  Best effort under the
// memory budget: returns HM_EBUDGET without allocating if it would not fit.
int reserve_relayout(hm_tree* t, hipStream_t s) {
  std::lock_guard<std::mutex> pub(t->pub_mu);
  const uint64_t nl = t->n_leaves;
  const uint64_t leaf_cap = nl + nl / 16 + 64;
  uint32_t ls[MAX_LEVELS + 1], ni = 0;
  if (plan_levels(leaf_cap, t->F, ls, &ni) < 0) return set_err(HM_EINVAL, "too many levels");
  const uint32_t inner_cap = ni + ni / 8 + 8;
  const size_t a4 = align256((nl + 1) * 4);
  const size_t rl = align256(4 * a4 + align256(scan_tmp_bytes(nl + 1))) + inner_scratch_bytes(leaf_cap, t->F, ni);
  uint64_t grow = 0;
  if (t->sb_leaf_cap < leaf_cap) grow += (leaf_cap - t->sb_leaf_cap) * sizeof(Leaf);
  if (t->sb_inner_cap < inner_cap) grow += (uint64_t)(inner_cap - t->sb_inner_cap) * (t->NS * 8 + 4);
  if (t->rl_bytes < rl) grow += rl - t->rl_bytes;
  if (t->p.mem_budget && t->device_bytes() + grow > t->p.mem_budget)
    return set_err(HM_EBUDGET, "no room for a standby layout under the memory budget");
  HM_TRY(hipStreamSynchronize(s));
  if (t->sb_leaf_cap < leaf_cap) {
    if (t->sb_leaves) HM_TRY(hipFree(t->sb_leaves));
    t->sb_leaves = nullptr;
    t->sb_leaf_cap = 0;
    HM_TRY(hipMalloc(&t->sb_leaves, leaf_cap * sizeof(Leaf)));
    t->sb_leaf_cap = leaf_cap;
  }
  if (t->sb_inner_cap < inner_cap) {
    if (t->sb_inner) HM_TRY(hipFree(t->sb_inner));
    if (t->sb_prefix) HM_TRY(hipFree(t->sb_prefix));
    t->sb_inner = nullptr;
    t->sb_prefix = nullptr;
    t->sb_inner_cap = 0;
    HM_TRY(hipMalloc(&t->sb_inner, (size_t)inner_cap * t->NS * 8));
    HM_TRY(hipMalloc(&t->sb_prefix, (size_t)inner_cap * 4));
    t->sb_inner_cap = inner_cap;
  }
  return ensure_rl(t, rl, s);
}

// Drop the staged batch of an epoch that failed.  The caller holds t->pub_mu; the staging copies
// still in flight complete first, so the slots can be reused.
void discard_staged(hm_tree* t) {
  if (t->stage_valid) (void)hipEventSynchronize(t->stage_ev);
  t->pend_n = 0;
}

// Make the prepared epoch visible: merge in place (after every earlier search) or swap in the
// standby layout; later searches wait for it on the device.  The caller holds t->pub_mu.
int publish_commit(hm_tree* t, hipStream_t s, const PublishPlan& plan, int* rebuilt) {
  std::lock_guard<std::mutex> g(t->mu);
  if (rebuilt) *rebuilt = plan.relayout ? 1 : 0;
  if (plan.n && !plan.relayout) {
    const UpdBufs b = carve_upd(t->upd_buf, plan.n);
    int rc = wait_for_readers(t, s);  // leaves change in place
    if (rc) return rc;
    HM_TRY(apply_runs(t->d_leaves, b.leaf, b.slot, b.found, b.uk, b.uv, plan.n, b.summary + 2, s));
    if (plan.respread) {
      uint64_t* last = t->d_inner + (size_t)t->level_start[t->height - 1] * t->NS;
      HM_TRY(node_respread(t->d_leaves, last, (uint32_t)t->n_leaves, t->F, b.nlist, plan.respread, b.leaf, b.slot,
                           b.found, b.uk, b.uv, b.summary + 2, s));
      t->respreads += plan.respread;
    }
  } else if (plan.relayout) {
    std::swap(t->d_leaves, t->sb_leaves);
    std::swap(t->leaf_cap, t->sb_leaf_cap);
    std::swap(t->d_inner, t->sb_inner);
    std::swap(t->d_prefix, t->sb_prefix);
    std::swap(t->inner_cap, t->sb_inner_cap);
    t->n_leaves = plan.new_leaves;
    set_levels(t, plan.lv);
    t->relayouts++;
  }
  t->n_keys += plan.n_new;
  t->widen_key_range(plan.key_lo, plan.key_hi);
  t->pend_n = 0;
  HM_TRY(hipEventRecord(t->publish_ev, s));
  t->publish_valid = true;
  t->epoch++;
  return HM_OK;
}

}  // namespace hm

extern "C" {

int hm_epoch_publish(hm_tree* t, void* stream, int* out_rebuilt) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (out_rebuilt) *out_rebuilt = 0;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(t->pub_mu);
  PublishPlan plan;
  int rc = publish_prepare(t, s, &plan);
  if (!rc) rc = publish_commit(t, s, plan, out_rebuilt);
  if (rc) discard_staged(t);  // the failed epoch is dropped, not carried into the next one
  return rc;
}

// ---- binary snapshot ------------------------------------------------------------------------
// File = header, leaves (n_leaves x 128 B), inner nodes (n_inner x node_slots u64), prefix
// (n_inner u32), all little-endian.  Loading uploads the leaves, rebuilds the inner levels and
// prefix array on the device and requires them to hash to the stored sections: the inner layout
// is a function of the leaves, so a file that disagrees is rejected instead of searched.
namespace {
constexpr char kSnapMagic[8] = {'H', 'M', 'S', 'N', 'A', 'P', '0', '1'};
constexpr int kSnapLevels = 17;
struct SnapHeader {
  char magic[8];
  uint32_t version, fanout, leaf_fill, node_slots;
  uint32_t height, n_inner;
  uint64_t n_leaves, n_keys, epoch;
  uint32_t level_start[kSnapLevels];
  uint32_t reserved;
  uint64_t fnv_leaves, fnv_inner, fnv_prefix;
};
static_assert(sizeof(SnapHeader) == 152, "snapshot header layout");

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 0xCBF29CE484222325ull) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
  return h;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};
}  // namespace

static int tree_save_impl(const hm_tree* t, const char* path);

int hm_tree_save(const hm_tree* t, const char* path) {
  if (!t || !path) return set_err(HM_EINVAL, "NULL argument");
  try {
    return tree_save_impl(t, path);
  } catch (const std::bad_alloc&) {
    return set_err(HM_ENOMEM, "%s: host allocation failed", path);
  }
}

static int tree_save_impl(const hm_tree* t, const char* path) {
  if (t->pend_n) return set_err(HM_EINVAL, "%llu staged updates: publish before saving", (unsigned long long)t->pend_n);
  std::vector<uint64_t> leaves(t->n_leaves * 16), inner((size_t)t->n_inner * t->NS);
  std::vector<uint32_t> prefix(t->n_inner);
  int rc = hm_tree_export(t, inner.data(), prefix.data(), leaves.data());
  if (rc) return rc;
  SnapHeader h{};
  std::memcpy(h.magic, kSnapMagic, 8);
  h.version = 1;
  h.fanout = (uint32_t)t->F;
  h.leaf_fill = (uint32_t)t->fill;
  h.node_slots = (uint32_t)t->NS;
  h.height = (uint32_t)t->height;
  h.n_inner = t->n_inner;
  h.n_leaves = t->n_leaves;
  h.n_keys = t->n_keys;
  h.epoch = t->epoch;
  for (int i = 0; i < kSnapLevels; ++i) h.level_start[i] = t->level_start[i];
  h.fnv_leaves = fnv1a(leaves.data(), leaves.size() * 8);
  h.fnv_inner = fnv1a(inner.data(), inner.size() * 8);
  h.fnv_prefix = fnv1a(prefix.data(), prefix.size() * 4);
  File f;
  f.f = fopen(path, "wb");
  if (!f.f) return set_err(HM_EINVAL, "cannot create %s", path);
  if (fwrite(&h, sizeof h, 1, f.f) != 1 || fwrite(leaves.data(), 8, leaves.size(), f.f) != leaves.size() ||
      fwrite(inner.data(), 8, inner.size(), f.f) != inner.size() ||
      fwrite(prefix.data(), 4, prefix.size(), f.f) != prefix.size() || fflush(f.f) != 0)
    return set_err(HM_EINVAL, "write to %s failed", path);
  return HM_OK;
}

static int tree_load_impl(const char* path, const hm_params* p, hm_tree** out) {
  hm_params q;
  hm_params_init(&q);
  if (p) q = *p;
  File f;
  f.f = fopen(path, "rb");
  if (!f.f) return set_err(HM_EINVAL, "cannot open %s", path);
  SnapHeader h{};
  if (fread(&h, sizeof h, 1, f.f) != 1 || std::memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != 1)
    return set_err(HM_EINVAL, "%s is not a harmonia snapshot", path);
  q.fanout = h.fanout;  // the layout decides these
  q.leaf_fill = h.leaf_fill;
  int rc = check_params(&q);
  if (rc) return rc;
  if (h.node_slots != (uint32_t)node_slots((int)h.fanout) || h.n_leaves == 0 || h.n_leaves >= (1ull << 32) ||
      h.n_keys > h.n_leaves * LEAF_KEYS)
    return set_err(HM_EINVAL, "%s: inconsistent snapshot header", path);
  // no hash covers the header: before anything is sized from it, the file must be exactly as long
  // as the header says, and the level table must be the one the leaf count defines
  {
    uint32_t ls[MAX_LEVELS + 1], ni = 0;
    const int height = plan_levels(h.n_leaves, (int)h.fanout, ls, &ni);
    if (height < 0 || (uint32_t)height != h.height || ni != h.n_inner)
      return set_err(HM_EINVAL, "%s: header level table does not match %llu leaves", path,
                     (unsigned long long)h.n_leaves);
    for (int i = 0; i <= MAX_LEVELS; ++i)
      if (ls[i] != h.level_start[i]) return set_err(HM_EINVAL, "%s: header level table is corrupt", path);
    const uint64_t want = sizeof(SnapHeader) + h.n_leaves * sizeof(Leaf) +
                          (uint64_t)h.n_inner * ((uint64_t)h.node_slots * 8 + 4);
    if (fseeko(f.f, 0, SEEK_END) != 0) return set_err(HM_EINVAL, "%s: cannot seek", path);
    const off_t size = ftello(f.f);
    if (size < 0 || (uint64_t)size != want)
      return set_err(HM_EINVAL, "%s: %lld bytes, the header describes %llu", path, (long long)size,
                     (unsigned long long)want);
    if (fseeko(f.f, (off_t)sizeof(SnapHeader), SEEK_SET) != 0) return set_err(HM_EINVAL, "%s: cannot seek", path);
  }
  std::vector<uint64_t> leaves(h.n_leaves * 16);
  if (fread(leaves.data(), 8, leaves.size(), f.f) != leaves.size() ||
      fnv1a(leaves.data(), leaves.size() * 8) != h.fnv_leaves)
    return set_err(HM_EINVAL, "%s: leaf section truncated or corrupt", path);
  // the leaf chain must hold the keys in strictly ascending order, padding at the end of each leaf
  uint64_t n_keys = 0, prev = 0, first = 0;
  bool any = false;
  for (uint64_t l = 0; l < h.n_leaves; ++l) {
    const uint64_t* k = &leaves[l * 16];
    int c = 0;
    while (c < LEAF_KEYS && k[c] != HM_KEY_MAX) ++c;
    for (int i = c; i < LEAF_KEYS; ++i)
      if (k[i] != HM_KEY_MAX) return set_err(HM_EINVAL, "%s: leaf %llu has a key after padding", path, (unsigned long long)l);
    if (c == 0 && h.n_leaves > 1) return set_err(HM_EINVAL, "%s: empty leaf %llu", path, (unsigned long long)l);
    for (int i = 0; i < c; ++i) {
      if (any && k[i] <= prev) return set_err(HM_EINVAL, "%s: keys out of order in leaf %llu", path, (unsigned long long)l);
      if (!any) first = k[i];
      prev = k[i];
      any = true;
    }
    n_keys += (uint64_t)c;
  }
  if (n_keys != h.n_keys) return set_err(HM_EINVAL, "%s: key count mismatch", path);
  const uint64_t need = h.n_leaves * sizeof(Leaf) + (uint64_t)h.n_inner * ((uint64_t)h.node_slots * 8 + 4);
  if (q.mem_budget && need > q.mem_budget)
    return set_err(HM_EBUDGET, "tree needs %llu device bytes, budget is %llu", (unsigned long long)need,
                   (unsigned long long)q.mem_budget);
  // host buffers first: nothing below may throw once the tree object exists
  std::vector<uint64_t> inner((size_t)h.n_inner * h.node_slots), finner(inner.size());
  std::vector<uint32_t> prefix(h.n_inner), fprefix(prefix.size());
  hm_tree* t = nullptr;
  rc = create_tree(&q, &t);
  if (rc) return rc;
  t->n_keys = n_keys;
  t->n_leaves = h.n_leaves;
  t->leaf_cap = h.n_leaves;
  if (any) t->widen_key_range(first, prev);
  auto fail = [&](int code) {
    free_tree(t);
    return code;
  };
  if (hipMalloc(&t->d_leaves, t->leaf_cap * sizeof(Leaf)) != hipSuccess)
    return fail(set_err(HM_ENOMEM, "leaf allocation failed"));
  if (hipMemcpy(t->d_leaves, leaves.data(), leaves.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(HM_EDEVICE, "leaf upload failed"));
  rc = rebuild_inner(t, t->own);
  if (rc) return fail(rc);
  // the stored inner levels and prefix must be the ones the leaves define
  if (t->n_inner != h.n_inner || (uint32_t)t->height != h.height ||
      fread(finner.data(), 8, finner.size(), f.f) != finner.size() ||
      fread(fprefix.data(), 4, fprefix.size(), f.f) != fprefix.size())
    return fail(set_err(HM_EINVAL, "%s: inner section truncated or inconsistent", path));
  rc = hm_tree_export(t, inner.data(), prefix.data(), nullptr);
  if (rc) return fail(rc);
  if (fnv1a(finner.data(), finner.size() * 8) != h.fnv_inner || fnv1a(fprefix.data(), fprefix.size() * 4) != h.fnv_prefix ||
      std::memcmp(inner.data(), finner.data(), inner.size() * 8) != 0 ||
      std::memcmp(prefix.data(), fprefix.data(), prefix.size() * 4) != 0)
    return fail(set_err(HM_EINVAL, "%s: stored inner levels do not match the leaves", path));
  t->epoch = h.epoch;
  *out = t;
  return HM_OK;
}

int hm_tree_load(const char* path, const hm_params* p, hm_tree** out) {
  if (!path || !out) return set_err(HM_EINVAL, "NULL argument");
  *out = nullptr;
  try {
    return tree_load_impl(path, p, out);
  } catch (const std::bad_alloc&) {
    return set_err(HM_ENOMEM, "%s: host allocation failed", path);
  } catch (const std::exception& e) {
    return set_err(HM_EINVAL, "%s: %s", path, e.what());
  }
}

int hm_tree_stats(const hm_tree* t, hm_stats* st) {
  if (!t || !st) return set_err(HM_EINVAL, "NULL argument");
  std::memset(st, 0, sizeof(*st));
  st->n_keys = t->n_keys;
  st->n_leaves = t->n_leaves;
  st->n_inner = t->n_inner;
  st->height = (uint32_t)t->height;
  st->fanout = (uint32_t)t->F;
  st->node_slots = (uint32_t)t->NS;
  st->lds_levels = (uint32_t)t->lds_levels;
  for (int i = 0; i <= MAX_LEVELS; ++i) st->level_start[i] = t->level_start[i];
  st->bytes_inner = (uint64_t)t->n_inner * t->NS * 8;
  st->bytes_prefix = (uint64_t)t->n_inner * 4;
  st->bytes_leaves = t->n_leaves * sizeof(Leaf);
  st->bytes_device = t->device_bytes();
  st->epoch = t->epoch;
  st->relayouts = t->relayouts;
  st->respreads = t->respreads;
  st->key_min = t->key_lo;
  st->key_max = t->key_hi;
  st->psa_window_top = (uint32_t)t->psa_top;
  return HM_OK;
}

int hm_tree_export(const hm_tree* t, uint64_t* inner, uint32_t* prefix, uint64_t* leaves) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  hm_tree* mt = const_cast<hm_tree*>(t);
  std::lock_guard<std::mutex> g(mt->mu);
  HM_TRY(hipStreamSynchronize(t->own));
  HM_TRY(hipDeviceSynchronize());
  if (inner) {
    HM_TRY(hipMemcpy(inner, t->d_inner, (size_t)t->n_inner * t->NS * 8, hipMemcpyDeviceToHost));
    const int is = idx_slots(t->F);  // the device keeps prefix[node] in the last index slot
    for (uint32_t j = 0; j < t->n_inner; ++j) inner[(size_t)j * t->NS + is - 1] = KEY_MAX;
  }
  if (prefix) HM_TRY(hipMemcpy(prefix, t->d_prefix, (size_t)t->n_inner * 4, hipMemcpyDeviceToHost));
  if (leaves) HM_TRY(hipMemcpy(leaves, t->d_leaves, t->n_leaves * sizeof(Leaf), hipMemcpyDeviceToHost));
  return HM_OK;
}

int hm_psa_sort(const uint64_t* keys, uint64_t n, int begin_bit, int end_bit, uint64_t* keys_out, uint32_t* pos_out,
                void* stream) {
  if (n == 0) return HM_OK;
  if (!keys || !keys_out || !pos_out) return set_err(HM_EINVAL, "NULL buffer");
  if (begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) return set_err(HM_EINVAL, "bad bit range");
  if (n > kMaxBatch) return set_err(HM_EINVAL, "too many keys");
  hipStream_t s = (hipStream_t)stream;
  const size_t kq = align256(n * 8), kp = align256(n * 4);
  struct Scratch {  // freed on every return path, after the stream's work on it
    char* p = nullptr;
    hipStream_t s = nullptr;
    ~Scratch() {
      if (p) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(p);
      }
    }
  } scratch;
  scratch.s = s;
  HM_TRY(hipMalloc(&scratch.p, 3 * kq + 2 * kp + align256(radix_tmp_bytes(n))));
  char* cur = scratch.p;
  const uint64_t* dk = keys;
  if (!is_device_ptr(keys)) {
    HM_TRY(hipMemcpyAsync(cur, keys, n * 8, hipMemcpyHostToDevice, s));
    dk = (const uint64_t*)cur;
  }
  cur += kq;
  uint64_t* k0 = (uint64_t*)cur;
  cur += kq;
  uint64_t* k1 = (uint64_t*)cur;
  cur += kq;
  uint32_t* p0 = (uint32_t*)cur;
  cur += kp;
  uint32_t* p1 = (uint32_t*)cur;
  cur += kp;
  int which = 0;
  HM_TRY(radix_sort_pos(dk, n, begin_bit, end_bit, k0, p0, k1, p1, cur, s, &which));
  const bool dko = is_device_ptr(keys_out), dpo = is_device_ptr(pos_out);
  HM_TRY(hipMemcpyAsync(keys_out, which ? k1 : k0, n * 8, dko ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HM_TRY(hipMemcpyAsync(pos_out, which ? p1 : p0, n * 4, dpo ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HM_TRY(hipStreamSynchronize(s));
  return HM_OK;
}

int hm_profile_enable(int on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_on = on != 0;
  return HM_OK;
}

int hm_profile_reset(void) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto& r : g_recs) {
    (void)hipEventSynchronize(r.b);
    g_ev_pool.push_back(r.a);
    g_ev_pool.push_back(r.b);
  }
  g_recs.clear();
  return HM_OK;
}

int hm_profile_get(const char* stage, double* total_ms, uint64_t* launches) {
  if (!stage || !total_ms || !launches) return set_err(HM_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(g_prof_mu);
  double tot = 0;
  uint64_t cnt = 0;
  for (auto& r : g_recs) {
    if (g_stage_names[r.stage] != stage) continue;
    HM_TRY(hipEventSynchronize(r.b));
    float ms = 0;
    HM_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return HM_OK;
}

int hm_device_synchronize(void) {
  HM_TRY(hipDeviceSynchronize());
  return HM_OK;
}

int hm_last_error(char* buf, uint64_t len) {
  if (!buf || !len) return HM_EINVAL;
  std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return HM_OK;
}

}  // extern "C"
