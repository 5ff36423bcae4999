// Batched upsert on the device.
//
// The reference updates on the CPU: 56 threads upsert under per-leaf mutexes, splits are deferred
// to a serial pass, and the inner levels plus the prefix array are rebuilt breadth-first on the
// host (insert_parallel_Newupdate_batch / rebuild_without_leaf, H/src/bpt-ppi-update-batch.cpp:86-200,
// H/src/bpt-ppi-update.cpp:320-463).  Here a batch is sorted and de-duplicated (last write wins),
// each key is located in the tree, existing keys get their value overwritten in place, and new keys
// are merged into their leaf's free slots.  When some leaf would overflow, every leaf is re-laid
// out (overflowing leaves split evenly into ceil(total / fill) leaves) at offsets given by a device
// scan, and the caller rebuilds the inner levels and re-scans the prefix array on the device.
#include "hm_internal.h"

namespace hm {
namespace {

int grid_for(size_t work, int threads) {
  size_t g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

__device__ __forceinline__ int leaf_count(const Leaf& l) {
  int c = 0;
#pragma unroll
  for (int i = 0; i < LEAF_KEYS; ++i) c += l.key[i] != KEY_MAX ? 1 : 0;
  return c;
}

__global__ void mark_last_kernel(const uint64_t* __restrict__ k, size_t n, uint32_t* __restrict__ keep) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    keep[i] = (i + 1 == n || k[i] != k[i + 1]) ? 1u : 0u;
}

__global__ void compact_kernel(const uint64_t* __restrict__ k, const uint64_t* __restrict__ v,
                               const uint32_t* __restrict__ keep, const uint32_t* __restrict__ scan, size_t n,
                               uint64_t* __restrict__ ko, uint64_t* __restrict__ vo) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (keep[i]) {
      ko[scan[i]] = k[i];
      vo[scan[i]] = v[i];
    }
}

__global__ void apply_values_kernel(Leaf* __restrict__ leaves, const uint32_t* __restrict__ leaf,
                                    const uint32_t* __restrict__ slot, const uint8_t* __restrict__ found,
                                    const uint64_t* __restrict__ vals, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (found[i]) leaves[leaf[i]].val[slot[i]] = vals[i];
}

__global__ void count_inserts_kernel(const uint32_t* __restrict__ leaf, const uint8_t* __restrict__ found, size_t n,
                                     uint32_t* __restrict__ ins) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (!found[i]) atomicAdd(&ins[leaf[i]], 1u);
}

// The batch is sorted by key, so leaf ids are nondecreasing: the first not-found index of a leaf is
// the smallest i with !found[i] whose leaf has no earlier not-found key.
__global__ void insert_starts_kernel(const uint32_t* __restrict__ leaf, const uint8_t* __restrict__ found, size_t n,
                                     uint32_t* __restrict__ start) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (!found[i]) atomicMin(&start[leaf[i]], (uint32_t)i);
}

__global__ void overflow_kernel(const Leaf* __restrict__ leaves, const uint32_t* __restrict__ ins, size_t n_leaves,
                                uint32_t* __restrict__ flags) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_leaves; i += stride) {
    const uint32_t a = ins[i];
    if (a && (uint32_t)leaf_count(leaves[i]) + a > (uint32_t)LEAF_KEYS) atomicOr(flags, 1u);
  }
}

// Stream the merge of a leaf's keys with its inserts (the batch keys from `start` on that are not
// found and map to this leaf; they are contiguous among the not-found keys because the batch is
// sorted) in key order, calling emit(i, key, value) for i = 0, 1, ...; returns the merged count.
template <typename Emit>
__device__ int merge_leaf(const Leaf& l, uint32_t n_ins, uint32_t start, const uint64_t* __restrict__ bk,
                          const uint64_t* __restrict__ bv, const uint8_t* __restrict__ found, Emit emit) {
  const int cur = leaf_count(l);
  int a = 0, c = 0;
  uint32_t b = start, taken = 0;
  while (taken < n_ins && found[b]) ++b;  // skip existing keys interleaved with the inserts
  while (a < cur || taken < n_ins) {
    const bool take_leaf = taken >= n_ins || (a < cur && l.key[a] < bk[b]);
    if (take_leaf) {
      emit(c, l.key[a], l.val[a]);
      ++a;
    } else {
      emit(c, bk[b], bv[b]);
      ++taken;
      ++b;
      while (taken < n_ins && found[b]) ++b;
    }
    ++c;
  }
  return c;
}

__global__ void merge_in_place_kernel(Leaf* __restrict__ leaves, const uint32_t* __restrict__ ins,
                                      const uint32_t* __restrict__ start, const uint64_t* __restrict__ bk,
                                      const uint64_t* __restrict__ bv, const uint8_t* __restrict__ found,
                                      size_t n_leaves) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_leaves; i += stride) {
    const uint32_t a = ins[i];
    if (!a) continue;
    const Leaf l = leaves[i];
    Leaf o;
#pragma unroll
    for (int s = 0; s < LEAF_KEYS; ++s) {
      o.key[s] = KEY_MAX;
      o.val[s] = 0ull;
    }
    // the overflow check guarantees at most LEAF_KEYS merged keys here
    merge_leaf(l, a, start[i], bk, bv, found, [&](int idx, uint64_t k, uint64_t v) {
      if (idx < LEAF_KEYS) {
        o.key[idx] = k;
        o.val[idx] = v;
      }
    });
    leaves[i] = o;
  }
}

__global__ void new_counts_kernel(const Leaf* __restrict__ leaves, const uint32_t* __restrict__ ins, size_t n_leaves,
                                  int fill, uint32_t* __restrict__ cnt) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_leaves; i += stride) {
    const uint32_t total = (uint32_t)leaf_count(leaves[i]) + ins[i];
    cnt[i] = total <= (uint32_t)LEAF_KEYS ? 1u : (total + fill - 1) / fill;
  }
}

__global__ void relayout_kernel(const Leaf* __restrict__ old_leaves, size_t n_old, const uint32_t* __restrict__ ins,
                                const uint32_t* __restrict__ start, const uint64_t* __restrict__ bk,
                                const uint64_t* __restrict__ bv, const uint8_t* __restrict__ found,
                                const uint32_t* __restrict__ new_start, int fill, Leaf* __restrict__ new_leaves) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_old; i += stride) {
    const uint32_t a = ins[i];
    const Leaf l = old_leaves[i];
    Leaf* dst = new_leaves + new_start[i];
    if (!a) {
      *dst = l;
      continue;
    }
    // split the merged run evenly: part p holds merged keys [p*c/parts, (p+1)*c/parts),
    // at most ceil(c/parts) <= fill <= 8 keys each
    const int c = leaf_count(l) + (int)a;
    const int parts = c <= LEAF_KEYS ? 1 : (c + fill - 1) / fill;
    for (int p = 0; p < parts; ++p)
#pragma unroll
      for (int s = 0; s < LEAF_KEYS; ++s) {
        dst[p].key[s] = KEY_MAX;
        dst[p].val[s] = 0ull;
      }
    int part = 0, lo = 0, hi = c / parts;
    merge_leaf(l, a, start[i], bk, bv, found, [&](int idx, uint64_t k, uint64_t v) {
      while (idx >= hi && part + 1 < parts) {
        ++part;
        lo = hi;
        hi = (int)(((long long)(part + 1) * c) / parts);
      }
      const int slot = idx - lo;
      if (slot < LEAF_KEYS) {
        dst[part].key[slot] = k;
        dst[part].val[slot] = v;
      }
    });
  }
}

}  // namespace

hipError_t mark_last_of_runs(const uint64_t* sorted_keys, size_t n, uint32_t* keep, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(mark_last_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, sorted_keys, n, keep);
  return hipGetLastError();
}

hipError_t compact_pairs(const uint64_t* k, const uint64_t* v, const uint32_t* keep, const uint32_t* keep_scan,
                         size_t n, uint64_t* ko, uint64_t* vo, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(compact_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, k, v, keep, keep_scan, n, ko, vo);
  return hipGetLastError();
}

hipError_t apply_value_updates(Leaf* leaves, const uint32_t* leaf, const uint32_t* slot, const uint8_t* found,
                               const uint64_t* vals, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(apply_values_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaves, leaf, slot, found, vals, n);
  return hipGetLastError();
}

hipError_t count_inserts(const uint32_t* leaf, const uint8_t* found, size_t n, uint32_t* ins_per_leaf,
                         hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(count_inserts_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf, found, n, ins_per_leaf);
  return hipGetLastError();
}

hipError_t insert_starts(const uint32_t* leaf, const uint8_t* found, size_t n, uint32_t* ins_start, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(insert_starts_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf, found, n, ins_start);
  return hipGetLastError();
}

hipError_t overflow_check(const Leaf* leaves, const uint32_t* ins_per_leaf, size_t n_leaves, uint32_t* d_flags,
                          hipStream_t s) {
  ProfScope prof("update", s);
  hipLaunchKernelGGL(overflow_kernel, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, leaves, ins_per_leaf, n_leaves,
                     d_flags);
  return hipGetLastError();
}

hipError_t merge_in_place(Leaf* leaves, const uint32_t* ins_per_leaf, const uint32_t* ins_start, const uint64_t* bkeys,
                          const uint64_t* bvals, const uint8_t* found, size_t n_leaves, hipStream_t s) {
  ProfScope prof("update", s);
  hipLaunchKernelGGL(merge_in_place_kernel, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, leaves, ins_per_leaf,
                     ins_start, bkeys, bvals, found, n_leaves);
  return hipGetLastError();
}

hipError_t new_leaf_counts(const Leaf* leaves, const uint32_t* ins_per_leaf, size_t n_leaves, int fill, uint32_t* cnt,
                           hipStream_t s) {
  ProfScope prof("update", s);
  hipLaunchKernelGGL(new_counts_kernel, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, leaves, ins_per_leaf,
                     n_leaves, fill, cnt);
  return hipGetLastError();
}

hipError_t relayout_leaves(const Leaf* old_leaves, size_t n_old, const uint32_t* ins_per_leaf,
                           const uint32_t* ins_start, const uint64_t* bkeys, const uint64_t* bvals,
                           const uint8_t* found, const uint32_t* new_start, int fill, Leaf* new_leaves,
                           hipStream_t s) {
  ProfScope prof("update", s);
  hipLaunchKernelGGL(relayout_kernel, dim3(grid_for(n_old, 256)), dim3(256), 0, s, old_leaves, n_old, ins_per_leaf,
                     ins_start, bkeys, bvals, found, new_start, fill, new_leaves);
  return hipGetLastError();
}

}  // namespace hm
