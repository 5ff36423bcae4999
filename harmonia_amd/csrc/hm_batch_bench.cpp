// Uneven-batch benchmark driver (plain C++ against the public C ABI), called in-process by bench.py.
//
// Runs the same sequence of device-resident query batches two ways and times each with the host
// clock around the whole sequence:
//   plain   -- one hm_search call per batch (its own launches), results into one device buffer;
//   batched -- the graph batcher: hm_batcher_add_many packs batches until the buffer is full, then
//              hm_batcher_flush replays the captured sort -> search -> un-sort pipeline once.
// Both loops run here in C++, so neither pays an interpreter's per-batch overhead.  A first, untimed
// pass of each warms the kernels, workspaces and captured graphs up.  The results are checked by the
// caller on a separate pass through the same batcher configuration.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/harmonia.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int plain_pass(hm_tree* t, const uint64_t* const* qs, const uint64_t* ns, uint64_t count, uint64_t* dv, uint8_t* df,
               void* stream) {
  for (uint64_t i = 0; i < count; ++i) {
    const int rc = hm_search(t, qs[i], ns[i], dv, df, stream);
    if (rc) return rc;
  }
  return hm_device_synchronize();
}

int batched_pass(hm_batcher* b, const uint64_t* const* qs, const uint64_t* ns, uint64_t count, uint64_t* flushes) {
  uint64_t i = 0, f = 0;
  while (i < count) {
    uint64_t added = 0;
    int rc = hm_batcher_add_many(b, qs + i, ns + i, count - i, HM_BATCH_DEVICE, nullptr, &added);
    if (rc) return rc;
    if (added == 0) return HM_EINVAL;  // a single batch larger than the capacity
    i += added;
    if ((rc = hm_batcher_flush(b))) return rc;
    ++f;
  }
  *flushes = f;
  return hm_batcher_sync(b);
}

}  // namespace

// qs[i] / ns[i]: device pointer and size of batch i; dv / df: device buffers of at least max(ns)
// entries for the plain path.  Returns HM_OK and the two wall-clock times.
extern "C" int hm_batch_bench_run(hm_tree* t, const uint64_t* const* qs, const uint64_t* ns, uint64_t count,
                                  uint64_t capacity, void* stream, uint64_t* dv, uint8_t* df, double* plain_s,
                                  double* batched_s, uint64_t* flushes) {
  int rc = plain_pass(t, qs, ns, count < 64 ? count : 64, dv, df, stream);  // warm-up
  if (rc) return rc;
  double t0 = now_s();
  if ((rc = plain_pass(t, qs, ns, count, dv, df, stream))) return rc;
  *plain_s = now_s() - t0;
  hm_batcher* b = nullptr;
  if ((rc = hm_batcher_create(t, capacity, nullptr, stream, &b))) return rc;
  uint64_t f = 0;
  rc = batched_pass(b, qs, ns, count, &f);  // warm-up: captures the graphs
  if (!rc) {
    t0 = now_s();
    rc = batched_pass(b, qs, ns, count, &f);
    *batched_s = now_s() - t0;
    *flushes = f;
  }
  hm_batcher_destroy(b);
  return rc;
}
