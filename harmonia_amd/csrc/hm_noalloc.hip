// Allocation-free batched search: hm_search's sort -> search -> un-sort pipeline with its scratch in
// a fixed device array instead of the stream workspace (which grows with hipMalloc).
//
// The array lives in this library's code object, one copy per GPU, held until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only a stand-in for the
// symbol): the first call on a device resolves the device address with hipGetSymbolAddress, with
// that device current, and every later call on that device uses the recorded address.
//
// Every kernel of the pipeline writes its part of the scratch before reading it (histogram counts,
// digit offsets, staging indices, sorted keys, results), so nothing has to start at zero and the
// contents left by the previous call do not matter.  Calls that overlap on one GPU share the array
// and race: serialize them (one stream, or a wait on the previous call's completion).
#include <mutex>

#include "hm_tree.h"

namespace hm {
namespace {

// 2^20 queries with PSA over up to 32 bits and results in input order need 43 MiB (sorted keys
// x2, four replay records, sorted results and their un-sort buffer).
constexpr uint64_t kNoAllocMaxQueries = 1ull << 20;
constexpr size_t kScratchBytes = 44ull << 20;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (resolved once per device; dev is current).
int scratch_on(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_err(HM_EDEVICE, "device %d out of range", dev);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HM_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(HM_EDEVICE, "hipGetSymbolAddress returned NULL for the search scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HM_OK;
}

}  // namespace
}  // namespace hm

using namespace hm;

extern "C" {

uint64_t hm_search_noalloc_max(void) { return kNoAllocMaxQueries; }

int hm_search_noalloc(hm_tree* t, const uint64_t* q, uint64_t n, uint64_t* out_vals, uint8_t* out_found,
                      void* stream) {
  if (!t) return set_err(HM_EINVAL, "tree is NULL");
  if (n == 0) return HM_OK;
  if (!q || !out_vals || !out_found) return set_err(HM_EINVAL, "NULL buffer");
  if (n > kNoAllocMaxQueries)
    return set_err(HM_EINVAL, "hm_search_noalloc: %llu queries exceed its fixed scratch (at most %llu per call)",
                   (unsigned long long)n, (unsigned long long)kNoAllocMaxQueries);
  if (!is_device_ptr(q) || !is_device_ptr(out_vals) || !is_device_ptr(out_found))
    return set_err(HM_EINVAL, "hm_search_noalloc takes device buffers only");
  int dev = -1;
  HM_TRY(hipGetDevice(&dev));
  if (dev != t->device)
    return set_err(HM_EINVAL, "hm_search_noalloc: current device %d is not the tree's device %d", dev, t->device);
  void* base = nullptr;
  int rc = scratch_on(dev, &base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(t->mu);
  SearchLaunch sl;
  int bits = 0, out_mode = 0;
  rc = resolve_search(t, n, nullptr, &sl, &bits, &out_mode);
  if (rc) return rc;
  const size_t need = pipe_bytes(n, bits, out_mode);
  if (need > kScratchBytes)
    return set_err(HM_EINVAL, "hm_search_noalloc: %llu bytes of scratch needed, %llu available",
                   (unsigned long long)need, (unsigned long long)kScratchBytes);
  if (t->publish_valid) HM_TRY(hipStreamWaitEvent(s, t->publish_ev, 0));
  char* cur = (char*)base;
  const PipeBufs pb = carve_pipe(cur, n, bits, out_mode);
  {
    ProfMute mute;  // stage profiling would create events
    rc = enqueue_pipeline(t->view(), sl, bits, out_mode, q, n, nullptr, pb, out_vals, out_found, nullptr, s);
  }
  if (rc) return rc;
  // a publish waits for this call's reads (wait_for_readers); the event was created with the tree
  HM_TRY(hipEventRecord(t->na_ev, s));
  t->na_valid = true;
  return HM_OK;
}

}  // extern "C"
