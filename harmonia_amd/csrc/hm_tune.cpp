// Autotuner for the search launch shape.
//
// The reference fixes its knobs at compile time (fanout DEFAULT_ORDER and Narrow_Rate in CMake,
// H/CMakeLists.txt:55-56; PSA bits 48..64 in every SortPairs call) and the paper picks the thread
// group size with a static profiling model (Harmonia.pdf p.140-141).  Here the candidates are timed
// on the device with a sample batch: lanes per query x queries in flight per group x LDS-resident
// levels x blocks per CU for the unsorted path; PSA bits x output mode x the same shape for the
// sorted path; then the sort threshold (smallest batch size at which PSA wins) and, in
// hm_tree_build_tuned, the fanout.  Winners become the tree's defaults.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hm_tree.h"

using namespace hm;

namespace {

struct Bufs {
  uint64_t* q = nullptr;
  uint64_t* vals = nullptr;
  uint8_t* found = nullptr;
  void* scratch = nullptr;
  ~Bufs() {
    for (void* p : {(void*)q, (void*)vals, (void*)found, scratch})
      if (p) (void)hipFree(p);
  }
};

// Average milliseconds of `reps` pipeline runs (after one warm-up) on stream s.
int time_pipeline(hm_tree* t, const SearchLaunch& sl, int bits, int mode, const Bufs& b, size_t n, hipStream_t s,
                  int reps, float* ms) {
  char* cur = (char*)b.scratch;
  PipeBufs pb = carve_pipe(cur, n, bits, mode);
  const TreeView tv = t->view();
  int rc = enqueue_pipeline(tv, sl, bits, mode, b.q, n, nullptr, pb, b.vals, b.found, nullptr, s);
  if (rc) return rc;
  struct Events {  // destroyed on every return path
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HM_TRY(hipEventCreate(&ev.e0));
  HM_TRY(hipEventCreate(&ev.e1));
  HM_TRY(hipEventRecord(ev.e0, s));
  for (int r = 0; r < reps && !rc; ++r) rc = enqueue_pipeline(tv, sl, bits, mode, b.q, n, nullptr, pb, b.vals,
                                                              b.found, nullptr, s);
  HM_TRY(hipEventRecord(ev.e1, s));
  HM_TRY(hipEventSynchronize(ev.e1));
  if (rc) return rc;
  float t_ms = 0;
  HM_TRY(hipEventElapsedTime(&t_ms, ev.e0, ev.e1));
  *ms = t_ms / reps;
  return HM_OK;
}

struct Best {
  SearchLaunch sl;
  int bits = 0, mode = 0;
  float ms = 1e30f;
};

int tune_tree(hm_tree* t, const uint64_t* sample, uint64_t n, uint32_t flags, hm_tune_result* res) {
  hipStream_t s = t->own;
  Bufs b;
  HM_TRY(hipMalloc(&b.q, n * 8));
  HM_TRY(hipMalloc(&b.vals, n * 8));
  HM_TRY(hipMalloc(&b.found, n));
  HM_TRY(hipMalloc(&b.scratch, pipe_bytes(n, 32, HM_OUT_UNPERMUTE)));
  HM_TRY(hipMemcpyAsync(b.q, sample, n * 8, is_device_ptr(sample) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        s));
  const int reps = (flags & HM_TUNE_QUICK) ? 2 : 4;
  const int maxl = std::min(3, t->height);
  // 1..8 lanes rank through the index granule; 16, 32, 64 lanes rank the whole node in one step (the
  // reference's narrowed groups, wider than its 32-lane warps allow); at most F lanes
  const int groups[] = {1, 2, 4, 8, 16, 32, 64};
  const int qpgs[] = {1, 2, 4};
  const int bpcs[] = {4, 5, 8, 10};  // persistent grids: whole rounds of resident blocks matter
  uint32_t trials = 0;
  Best bd, bp;
  int rc = 0;
  // Wide groups cost G/8 times the bytes of a level and are rarely competitive: before their grid, one
  // probe (16 lanes, 4 queries, 8 blocks per CU) must come within 1.5x of the best narrow shape.
  auto wide_probe = [&](int bits, int mode, const Best& best, bool* ok) {
    *ok = false;
    if (t->F < 16) return HM_OK;
    SearchLaunch sl = best.sl;
    sl.group = 16;
    sl.queries_per_group = 4;
    sl.blocks_per_cu = 8;
    float ms = 0;
    const int r = time_pipeline(t, sl, bits, mode, b, n, s, reps, &ms);
    if (r) return r;
    ++trials;
    *ok = ms < 1.5f * best.ms;
    return HM_OK;
  };
  bool wide = false;
  for (int g : groups)
    for (int q : qpgs)
      for (int bpc : bpcs)
        for (int lds = 0; lds <= maxl; ++lds) {
          if (g > t->F) continue;
          if (g == 16 && q == qpgs[0] && bpc == bpcs[0] && lds == 0 && (rc = wide_probe(0, 0, bd, &wide))) return rc;
          if (g >= 16 && !wide) continue;
          SearchLaunch sl;
          sl.group = g;
          sl.queries_per_group = q;
          sl.blocks_per_cu = bpc;
          sl.lds_levels = choose_lds_levels(t, lds);
          if (sl.lds_levels != lds) continue;  // does not fit
          float ms = 0;
          if ((rc = time_pipeline(t, sl, 0, 0, b, n, s, reps, &ms))) return rc;
          ++trials;
          if (ms < bd.ms) bd = Best{sl, 0, 0, ms};
        }
  // PSA path: bits x output mode x shape (LDS levels fixed to the direct winner's)
  const int bitss[] = {12, 16, 20, 24};
  const int modes[] = {HM_OUT_UNPERMUTE, HM_OUT_SCATTER};
  for (int bits : bitss)
    for (int mode : modes)
      for (int g : groups)
        for (int q : qpgs)
          for (int bpc : bpcs) {
            if (g > t->F) continue;
            if (g == 16 && q == qpgs[0] && bpc == bpcs[0] && (rc = wide_probe(bits, mode, bp, &wide))) return rc;
            if (g >= 16 && !wide) continue;
            SearchLaunch sl = bd.sl;
            sl.group = g;
            sl.queries_per_group = q;
            sl.blocks_per_cu = bpc;
            float ms = 0;
            if ((rc = time_pipeline(t, sl, bits, mode, b, n, s, reps, &ms))) return rc;
            ++trials;
            if (ms < bp.ms) bp = Best{sl, bits, mode, ms};
          }
  // sort threshold: the smallest batch where the PSA winner beats the direct winner, searched over
  // 1024 * 2^(k/2) (a factor of ~1.41 per step, so a crossover between two powers of two is not
  // rounded up to the next one) and then the sample size n itself; if no smaller batch wins but PSA
  // won the grid at n, the threshold is n
  uint64_t threshold = ~0ull;
  std::vector<uint64_t> sizes;
  for (uint64_t m = 1024; m < n; m = (m * 181 + 127) / 128) sizes.push_back(m);  // 181/128 ~ sqrt(2)
  sizes.push_back(n);
  for (uint64_t m : sizes) {
    float md = 0, mp = 0;
    if ((rc = time_pipeline(t, bd.sl, 0, 0, b, m, s, reps, &md))) return rc;
    if ((rc = time_pipeline(t, bp.sl, bp.bits, bp.mode, b, m, s, reps, &mp))) return rc;
    trials += 2;
    if (mp < md) {
      threshold = m;
      break;
    }
  }
  if (threshold == ~0ull && bp.ms < bd.ms) threshold = n;
  // apply
  t->cfg_direct = SearchCfg{bd.sl.group, bd.sl.queries_per_group, bd.sl.blocks_per_cu, HM_OUT_UNPERMUTE};
  t->cfg_psa = SearchCfg{bp.sl.group, bp.sl.queries_per_group, bp.sl.blocks_per_cu, bp.mode};
  t->lds_levels = bd.sl.lds_levels;
  t->p.psa_bits = (uint32_t)bp.bits;
  t->p.sort_threshold = threshold;
  if (res) {
    std::memset(res, 0, sizeof(*res));
    res->fanout = (uint32_t)t->F;
    res->direct_group = (uint32_t)bd.sl.group;
    res->direct_queries_per_group = (uint32_t)bd.sl.queries_per_group;
    res->direct_blocks_per_cu = (uint32_t)bd.sl.blocks_per_cu;
    res->lds_levels = bd.sl.lds_levels;
    res->psa_group = (uint32_t)bp.sl.group;
    res->psa_queries_per_group = (uint32_t)bp.sl.queries_per_group;
    res->psa_blocks_per_cu = (uint32_t)bp.sl.blocks_per_cu;
    res->psa_bits = (uint32_t)bp.bits;
    res->output_mode = (uint32_t)bp.mode;
    res->sort_threshold = threshold;
    res->direct_ms = bd.ms;
    res->psa_ms = bp.ms;
    res->best_ms = std::min(bd.ms, bp.ms);
    res->trials = trials;
  }
  HM_TRY(hipStreamSynchronize(s));
  return HM_OK;
}

}  // namespace

extern "C" {

int hm_autotune(hm_tree* t, const uint64_t* sample, uint64_t n, uint32_t flags, hm_tune_result* out) {
  if (!t || !sample) return set_err(HM_EINVAL, "NULL argument");
  if (n < 1024 || n > (1ull << 31) - 1) return set_err(HM_EINVAL, "sample size must be in 1024..2^31-1");
  std::lock_guard<std::mutex> g(t->mu);
  HM_TRY(hipStreamSynchronize(t->own));
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(t->own, t->publish_ev, 0));
  return tune_tree(t, sample, n, flags, out);
}

int hm_tree_build_tuned(const uint64_t* keys, const uint64_t* vals, uint64_t n, const hm_params* p,
                        const uint64_t* sample, uint64_t m, uint32_t flags, hm_tree** out, hm_tune_result* res) {
  if (!out || !p) return set_err(HM_EINVAL, "NULL argument");
  *out = nullptr;
  const uint32_t fanouts[] = {32, 64, 128};
  hm_tree* best = nullptr;
  hm_tune_result best_r{};
  for (uint32_t F : fanouts) {
    hm_params q = *p;
    q.fanout = F;
    hm_tree* t = nullptr;
    int rc = hm_tree_build(keys, vals, n, &q, &t);
    if (rc) {
      if (best) hm_tree_destroy(best);
      return rc;
    }
    hm_tune_result r{};
    rc = hm_autotune(t, sample, m, flags, &r);
    if (rc) {
      hm_tree_destroy(t);
      if (best) hm_tree_destroy(best);
      return rc;
    }
    if (!best || r.best_ms < best_r.best_ms) {
      if (best) hm_tree_destroy(best);
      best = t;
      best_r = r;
    } else {
      hm_tree_destroy(t);
    }
  }
  *out = best;
  if (res) *res = best_r;
  return HM_OK;
}

}  // extern "C"
