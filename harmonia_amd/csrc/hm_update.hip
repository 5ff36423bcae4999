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
//
// The common serving case (a few thousand keys, no leaf overflows) stays O(batch): one block sorts
// and de-duplicates the batch in LDS, the keys are located, each leaf's keys form one run of the
// sorted batch, and one thread per run plans (overflow? how many new keys?) and then applies it.
#include <utility>

#include "hm_block.h"
#include "hm_internal.h"

namespace hm {
namespace {

int grid_for(size_t work, int threads) {
  size_t g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// Register-array access with a run-time index, as unrolled selects: a dynamic index into a local
// array would put the whole leaf in scratch memory.
__device__ __forceinline__ uint64_t pick(const uint64_t (&a)[LEAF_KEYS], int i) {
  // masks rather than selects: the compiler folds a select chain back into one indexed load
  uint64_t r = 0;
#pragma unroll
  for (int s = 0; s < LEAF_KEYS; ++s) r |= a[s] & (0ull - (uint64_t)(i == s));
  return r;
}

__device__ __forceinline__ void put(uint64_t (&a)[LEAF_KEYS], int i, uint64_t v) {
#pragma unroll
  for (int s = 0; s < LEAF_KEYS; ++s) a[s] = i == s ? v : a[s];
}

// A leaf held in registers (a Leaf struct copy would live in scratch memory).
struct LeafRegs {
  uint64_t key[LEAF_KEYS];
  uint64_t val[LEAF_KEYS];
};

__device__ __forceinline__ LeafRegs load_leaf(const Leaf* p) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
  LeafRegs r;
#pragma unroll
  for (int i = 0; i < LEAF_KEYS / 2; ++i) {
    const ulonglong2 a = q[i];
    r.key[2 * i] = a.x;
    r.key[2 * i + 1] = a.y;
  }
#pragma unroll
  for (int i = 0; i < LEAF_KEYS / 2; ++i) {
    const ulonglong2 b = q[LEAF_KEYS / 2 + i];
    r.val[2 * i] = b.x;
    r.val[2 * i + 1] = b.y;
  }
  return r;
}

__device__ __forceinline__ void store_leaf(Leaf* p, const LeafRegs& r) {
  ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
#pragma unroll
  for (int i = 0; i < LEAF_KEYS / 2; ++i) q[i] = make_ulonglong2(r.key[2 * i], r.key[2 * i + 1]);
#pragma unroll
  for (int i = 0; i < LEAF_KEYS / 2; ++i) q[LEAF_KEYS / 2 + i] = make_ulonglong2(r.val[2 * i], r.val[2 * i + 1]);
}

__device__ __forceinline__ int leaf_count(const LeafRegs& l) {
  int c = 0;
#pragma unroll
  for (int i = 0; i < LEAF_KEYS; ++i) c += l.key[i] != KEY_MAX ? 1 : 0;
  return c;
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

__global__ void count_inserts_kernel(const uint32_t* __restrict__ leaf, const uint8_t* __restrict__ found, size_t n,
                                     uint32_t* __restrict__ ins) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (!found[i]) atomicAdd(&ins[leaf[i]], 1u);
}

// Stream the merge of a leaf's keys with its inserts (the batch keys from `start` on that are not
// found and map to this leaf; they are contiguous among the not-found keys because the batch is
// sorted) in key order, calling emit(i, key, value) for i = 0, 1, ...; returns the merged count.
template <typename Emit>
__device__ __forceinline__ int merge_leaf(const LeafRegs& l, uint32_t n_ins, uint32_t start, const uint64_t* __restrict__ bk,
                          const uint64_t* __restrict__ bv, const uint8_t* __restrict__ found, Emit emit) {
  const int cur = leaf_count(l);
  int a = 0, c = 0;
  uint32_t b = start, taken = 0;
  while (taken < n_ins && found[b]) ++b;  // skip existing keys interleaved with the inserts
  while (a < cur || taken < n_ins) {
    const uint64_t ka = pick(l.key, a);
    const bool take_leaf = taken >= n_ins || (a < cur && ka < bk[b]);
    if (take_leaf) {
      emit(c, ka, pick(l.val, a));
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

__global__ void new_counts_kernel(const Leaf* __restrict__ leaves, const uint32_t* __restrict__ ins, size_t n_leaves,
                                  int fill, uint32_t* __restrict__ cnt) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_leaves; i += stride) {
    const uint32_t a = ins[i];
    if (!a) {  // most leaves: no new key, one leaf; the leaf itself is not read
      cnt[i] = 1u;
      continue;
    }
    const uint32_t total = (uint32_t)leaf_count(leaves[i]) + a;
    cnt[i] = total <= (uint32_t)LEAF_KEYS ? 1u : (total + fill - 1) / fill;
  }
}

__global__ void leaf_first_kernel(const uint32_t* __restrict__ leaf, size_t n, uint32_t* __restrict__ first) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (i == 0 || leaf[i] != leaf[i - 1]) first[leaf[i]] = (uint32_t)i;  // run heads: one writer per leaf
}

// Re-layout, part 1: every leaf in [lo, hi) the batch does not touch moves unchanged to
// new_start[i].  Runs of untouched leaves shift by a constant, so this is a coalesced streaming copy
// in 16-byte pieces; the piece holding key[0] also records the leaf's minimum key for the inner
// rebuild (which then does not read a cache line per leaf again).
__global__ __launch_bounds__(256) void relayout_copy_kernel(const Leaf* __restrict__ old_leaves, size_t lo, size_t hi,
                                                            const uint32_t* __restrict__ first,
                                                            const uint32_t* __restrict__ new_start,
                                                            Leaf* __restrict__ new_leaves,
                                                            uint64_t* __restrict__ new_mins) {
  constexpr int kPieces = sizeof(Leaf) / 16;
  const ulonglong2* src = reinterpret_cast<const ulonglong2*>(old_leaves);
  ulonglong2* dst = reinterpret_cast<ulonglong2*>(new_leaves);
  const size_t end = hi * kPieces;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t u = lo * kPieces + (size_t)blockIdx.x * blockDim.x + threadIdx.x; u < end; u += stride) {
    const size_t i = u / kPieces;
    if (first[i] == 0xFFFFFFFFu) {
      const ulonglong2 v = src[u];
      dst[(size_t)new_start[i] * kPieces + u % kPieces] = v;
      if (u % kPieces == 0) new_mins[new_start[i]] = v.x;
    }
  }
}

// Re-layout, part 2: one thread per leaf the batch touches (a run head of the sorted batch): existing
// keys get their new value in the copy, new keys are merged in, and a leaf that overflows is split
// evenly into ceil(total / fill) leaves at new_start[L].
__global__ void relayout_touched_kernel(const Leaf* __restrict__ old_leaves, const uint32_t* __restrict__ ins,
                                        const uint32_t* __restrict__ bleaf, const uint32_t* __restrict__ bslot,
                                        const uint64_t* __restrict__ bk, const uint64_t* __restrict__ bv,
                                        const uint8_t* __restrict__ found, size_t m,
                                        const uint32_t* __restrict__ new_start, int fill,
                                        Leaf* __restrict__ new_leaves, uint64_t* __restrict__ new_mins) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t h = (size_t)blockIdx.x * blockDim.x + threadIdx.x; h < m; h += stride) {
    if (h > 0 && bleaf[h] == bleaf[h - 1]) continue;  // not a run head
    const uint32_t L = bleaf[h];
    const uint32_t a = ins[L];
    LeafRegs l = load_leaf(old_leaves + L);
    for (size_t j = h; j < m && bleaf[j] == L; ++j)
      if (found[j]) put(l.val, (int)bslot[j], bv[j]);
    Leaf* dst = new_leaves + new_start[L];
    uint64_t* dmin = new_mins + new_start[L];
    if (!a) {
      store_leaf(dst, l);
      *dmin = l.key[0];
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
    merge_leaf(l, a, (uint32_t)h, bk, bv, found, [&](int idx, uint64_t k, uint64_t v) {
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
      if (slot == 0) dmin[part] = k;  // every part gets at least one key
    });
  }
}

constexpr int kSortThreads = 1024;

// Bitonic network over N (a power of two, <= kSmallUpdate) (key, index) pairs in LDS, ascending by
// key, then index.  Pair q (elements i, i + j) is handled by thread q % kSortThreads, so every lane
// has work in every step.  Each step moves every pair through LDS (about 98 KB at N = 4096), and that
// bounds it: wave-level syncs for the steps whose pairs stay inside one wave measured the same as block
// barriers.  Ends with a block barrier.
__device__ __forceinline__ void bitonic_pairs(uint64_t* sk, uint32_t* si, uint32_t N) {
  for (uint32_t k = 2; k <= N; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      // pair q: i = q with a zero bit inserted at j's position, partner i + j (every lane busy)
      for (uint32_t q = threadIdx.x; q < N / 2; q += kSortThreads) {
        const uint32_t i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
        const uint32_t p = i + j;
        const uint64_t a = sk[i], b = sk[p];
        const uint32_t ia = si[i], ib = si[p];
        const bool gt = a > b || (a == b && ia > ib);
        if (gt == ((i & k) == 0)) {
          sk[i] = b;
          sk[p] = a;
          si[i] = ib;
          si[p] = ia;
        }
      }
      __syncthreads();
    }
  }
}

// LSD radix sort, in LDS, of kSmallUpdate (key, index) pairs held in (sk[0], si[0]), over the 8 key
// bytes; stable, so equal keys keep their index order.  8 passes that each move the pairs through LDS
// once (a bitonic network over the same pairs takes 78 steps).  Wave w owns positions
// [256 w, 256 w + 256), element r of lane l being 256 w + 64 r + l; ranks within a wave come from the
// digit ballots (match_digit), in that order.  Returns the buffer (0 or 1) holding the result; ends
// with a block barrier.
constexpr int kChunkWaves = kSortThreads / 64;
constexpr int kChunkItems = kSmallUpdate / kSortThreads;
static_assert(kChunkItems * 64 * kChunkWaves == (int)kSmallUpdate, "each wave owns 64 x kChunkItems positions");
struct RadixLds {
  uint64_t sk[2][kSmallUpdate];
  uint32_t si[2][kSmallUpdate];
  uint32_t wcnt[kChunkWaves][256];
  uint32_t boff[256];
  uint32_t scan_sh[kChunkWaves + 1];
};
__device__ __forceinline__ int radix_sort_lds(RadixLds& L) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint32_t base = (uint32_t)w * 64 * kChunkItems;
  int cur = 0;
  for (int shift = 0; shift < 64; shift += 8) {  // uniform across the block
    for (int i = threadIdx.x; i < kChunkWaves * 256; i += kSortThreads) (&L.wcnt[0][0])[i] = 0;
    __syncthreads();  // the pairs (first pass) and the cleared counters
    uint64_t k[kChunkItems];
    uint32_t ix[kChunkItems], pkd[kChunkItems];
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r) {
      k[r] = L.sk[cur][base + r * 64 + lane];
      ix[r] = L.si[cur][base + r * 64 + lane];
    }
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r) {
      const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
      const uint64_t peers = match_digit(d, ~0ull);
      const uint32_t prior = L.wcnt[w][d];
      if ((peers & lt) == 0) L.wcnt[w][d] = prior + (uint32_t)__popcll(peers);
      pkd[r] = ((prior + (uint32_t)__popcll(peers & lt)) << 8) | d;
    }
    __syncthreads();
    {
      // per digit: each wave's start inside the digit, then the digit's start
      const int d = threadIdx.x;
      uint32_t run = 0;
      if (d < 256) {
#pragma unroll
        for (int ww = 0; ww < kChunkWaves; ++ww) {
          const uint32_t t = L.wcnt[ww][d];
          L.wcnt[ww][d] = run;
          run += t;
        }
      }
      uint32_t total;
      const uint32_t e = block_exclusive_scan<uint32_t, kSortThreads>(run, L.scan_sh, total);
      if (d < 256) L.boff[d] = e;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r) {
      const uint32_t d = pkd[r] & 255u;
      const uint32_t np = L.boff[d] + L.wcnt[w][d] + (pkd[r] >> 8);
      L.sk[cur ^ 1][np] = k[r];
      L.si[cur ^ 1][np] = ix[r];
    }
    cur ^= 1;
    __syncthreads();  // every thread has read wcnt / boff before the next pass clears them
  }
  return cur;
}

// First stage of the merge sort of mid-size batches: chunk c of kSmallUpdate staged keys sorted by
// key with their staging indices.  Padding (KEY_MAX, ~0) has digit 255 in every pass and sorts after
// every staged key.
__global__ __launch_bounds__(kSortThreads) void chunk_sort_kernel(const uint64_t* __restrict__ pk, uint32_t n,
                                                                  uint64_t* __restrict__ ko,
                                                                  uint32_t* __restrict__ io) {
  __shared__ RadixLds L;
  const uint32_t c0 = blockIdx.x * kSmallUpdate;
  const uint32_t cnt = n - c0 < kSmallUpdate ? n - c0 : kSmallUpdate;
  for (uint32_t i = threadIdx.x; i < kSmallUpdate; i += kSortThreads) {
    L.sk[0][i] = i < cnt ? pk[c0 + i] : KEY_MAX;
    L.si[0][i] = i < cnt ? c0 + i : 0xFFFFFFFFu;
  }
  const int cur = radix_sort_lds(L);
  for (uint32_t i = threadIdx.x; i < cnt; i += kSortThreads) {
    ko[c0 + i] = L.sk[cur][i];
    io[c0 + i] = L.si[cur][i];
  }
}

// One merge pass over sorted runs of length L (pairs [2gL, 2gL + L) and [2gL + L, 2gL + 2L)): each
// element goes to its own rank in the run plus the number of elements of the other run that order
// before it.  (key, index) pairs are all distinct, so the ranks never collide and the merge keeps
// the staging order of equal keys.
__global__ __launch_bounds__(256) void merge_pass_kernel(const uint64_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ iin,
                                                         uint64_t* __restrict__ kout, uint32_t* __restrict__ iout,
                                                         uint32_t n, uint32_t L) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= n) return;
  const uint32_t base = e / (2 * L) * (2 * L);
  const uint32_t mid = base + L < n ? base + L : n;
  const uint32_t end = mid + L < n ? mid + L : n;
  const uint64_t k = kin[e];
  const uint32_t ix = iin[e];
  const bool in_a = e < mid;
  uint32_t lo = in_a ? mid : base, hi = in_a ? end : mid;  // search the other run
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    const uint64_t km = kin[m];
    if (km < k || (km == k && iin[m] < ix))
      lo = m + 1;
    else
      hi = m;
  }
  const uint32_t pos = in_a ? base + (e - base) + (lo - mid) : base + (e - mid) + (lo - base);
  kout[pos] = k;
  iout[pos] = ix;
}

__global__ void gather_by_index_kernel(const uint64_t* __restrict__ v, const uint32_t* __restrict__ ix, uint32_t n,
                                       uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v[ix[i]];
}

// Sort of (key, staging index) in LDS, stable in the index, so the last write of a key ends its run
// of equal keys; then the last of each run is kept.  Padding (KEY_MAX, ~0) sorts after every real key
// (KEY_MAX is never staged).
__global__ __launch_bounds__(kSortThreads) void small_sort_dedupe_kernel(const uint64_t* __restrict__ pk,
                                                                         const uint64_t* __restrict__ pv, uint32_t n,
                                                                         uint64_t* __restrict__ uk,
                                                                         uint64_t* __restrict__ uv,
                                                                         uint32_t* __restrict__ dm) {
  __shared__ RadixLds L;
  __shared__ uint32_t scan_sh[kSortThreads / 64 + 1];
  // up to 1024 keys: the bitonic network over the next power of two; above: the radix sort (8 passes
  // over all kSmallUpdate slots, padding included), which is cheaper there
  const bool radix = n > 1024;
  uint32_t N = 1;
  while (N < n) N <<= 1;
  if (radix) N = kSmallUpdate;
  for (uint32_t i = threadIdx.x; i < N; i += kSortThreads) {
    L.sk[0][i] = i < n ? pk[i] : KEY_MAX;
    L.si[0][i] = i < n ? i : 0xFFFFFFFFu;
  }
  int cur = 0;
  if (radix) {  // block-uniform
    cur = radix_sort_lds(L);
  } else {
    __syncthreads();
    bitonic_pairs(L.sk[0], L.si[0], N);
  }
  const uint64_t* sk = L.sk[cur];
  const uint32_t* si = L.si[cur];
  constexpr int kPer = kSmallUpdate / kSortThreads;  // consecutive elements per thread
  const uint32_t b0 = threadIdx.x * kPer;
  bool keep[kPer];
  uint32_t cnt = 0;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const uint32_t i = b0 + r;
    keep[r] = i < n && (i + 1 == n || sk[i] != sk[i + 1]);
    cnt += keep[r] ? 1u : 0u;
  }
  uint32_t total = 0;
  uint32_t off = block_exclusive_scan<uint32_t, kSortThreads>(cnt, scan_sh, total);
#pragma unroll
  for (int r = 0; r < kPer; ++r)
    if (keep[r]) {
      uk[off] = sk[b0 + r];
      uv[off] = pv[si[b0 + r]];
      ++off;
    }
  if (threadIdx.x == 0) *dm = total;
}

__global__ void unique_count_kernel(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ scan, size_t n,
                                    uint32_t* __restrict__ dm) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dm = scan[n - 1] + keep[n - 1];
}

__device__ __forceinline__ bool run_head(const uint32_t* __restrict__ leaf, uint32_t i) {
  return i == 0 || leaf[i] != leaf[i - 1];
}

// Plan an epoch, one thread per last-level node the batch touches (its first batch entry: the batch
// is sorted, so a node's entries are contiguous).  Last-level node j holds leaves
// [j F, min((j + 1) F, n_leaves)): the bulk build and re-layouts fill every node but the last, and
// an in-node re-spread keeps each node's leaf count.  summary[1] += new keys; if a leaf of the node
// would overflow, summary[0] |= 1 and the node either goes on the re-spread list (nlist: node, first
// batch entry; count in summary[4]) when its leaves can hold all its keys, or summary[0] |= 2 (the
// epoch needs a full re-layout).
__global__ void plan_runs_kernel(const Leaf* __restrict__ leaves, uint32_t n_leaves, int F,
                                 const uint32_t* __restrict__ leaf, const uint8_t* __restrict__ found, size_t n_max,
                                 const uint32_t* __restrict__ dm, uint32_t* __restrict__ summary,
                                 uint32_t* __restrict__ nlist) {
  __shared__ uint32_t block_new;
  if (threadIdx.x == 0) block_new = 0;
  __syncthreads();
  const uint32_t m = dm ? *dm : (uint32_t)n_max;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t uF = (uint32_t)F;
  uint32_t my_new = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    my_new += found[i] ? 0u : 1u;
    const uint32_t node = leaf[i] / uF;
    if (i > 0 && leaf[i - 1] / uF == node) continue;  // not the node's first entry
    bool over = false;
    uint32_t node_ins = 0;
    for (uint32_t j = i; j < m && leaf[j] / uF == node;) {
      const uint32_t L = leaf[j];
      const uint32_t room = (uint32_t)(LEAF_KEYS - leaf_count(leaves[L]));
      uint32_t ins = 0;
      for (; j < m && leaf[j] == L; ++j) ins += found[j] ? 0u : 1u;
      over |= ins > room;
      node_ins += ins;
    }
    if (!over) continue;
    atomicOr(&summary[0], 1u);
    const uint32_t x0 = node * uF, x1 = x0 + uF < n_leaves ? x0 + uF : n_leaves;
    uint32_t total = node_ins;
    for (uint32_t x = x0; x < x1; ++x) total += (uint32_t)leaf_count(leaves[x]);
    if (total <= (x1 - x0) * (uint32_t)LEAF_KEYS) {
      const uint32_t k = atomicAdd(&summary[4], 1u);
      nlist[2 * k] = node;
      nlist[2 * k + 1] = i;
    } else {
      atomicOr(&summary[0], 2u);
    }
  }
  // new-key count: wave sum, then one LDS atomic per wave and one global atomic per block (a
  // global atomic per wave on one address serializes)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) my_new += __shfl_xor(my_new, o);
  if ((threadIdx.x & 63) == 0 && my_new) atomicAdd(&block_new, my_new);
  __syncthreads();
  if (threadIdx.x == 0 && block_new) atomicAdd(&summary[1], block_new);
}

// Mark the batch entries of the listed nodes (found |= 2) so the in-place merge leaves them to the
// node re-spread.  One block per listed node.
__global__ void mark_nodes_kernel(const uint32_t* __restrict__ nlist, uint32_t n_nodes, int F,
                                  const uint32_t* __restrict__ leaf, uint8_t* __restrict__ found,
                                  const uint32_t* __restrict__ dm) {
  const uint32_t m = *dm;
  for (uint32_t k = blockIdx.x; k < n_nodes; k += gridDim.x) {
    const uint32_t node = nlist[2 * k], i0 = nlist[2 * k + 1];
    for (uint32_t j = i0 + threadIdx.x;; j += blockDim.x) {
      if (j >= m || leaf[j] / (uint32_t)F != node) break;
      found[j] |= 2u;
    }
  }
}

// Re-spread one last-level node in place (one 64-lane block per listed node): its leaves' keys with
// the batch's values applied and its new keys merged in, spread evenly over the same leaves (leaf s
// gets merged keys [s K / c, (s + 1) K / c), at least one and at most 8 each), then the node's
// separators rewritten as the leaves' minima (key[s] = min of leaf s + 1) and its index granule
// (idx[g] = key[8 g + 7]; the last index slot, the node's prefix entry, is kept).  The node's
// smallest key does not change (new keys are >= the separator that leads to the node, which is that
// key), so the levels above stay valid.  Bounded cost: at most F leaves and one node per overflow.
template <int F>
__global__ __launch_bounds__(64) void node_respread_kernel(Leaf* __restrict__ leaves, uint64_t* __restrict__ nodes,
                                                           uint32_t n_leaves, const uint32_t* __restrict__ nlist,
                                                           uint32_t n_nodes, const uint32_t* __restrict__ leaf,
                                                           const uint32_t* __restrict__ slot,
                                                           const uint8_t* __restrict__ found,
                                                           const uint64_t* __restrict__ uk,
                                                           const uint64_t* __restrict__ uv,
                                                           const uint32_t* __restrict__ dm) {
  constexpr int CAP = LEAF_KEYS * F;
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  __shared__ uint64_t ek[CAP], ev[CAP];  // the node's keys in order (values updated), then merged
  __shared__ uint64_t nk[CAP], nv[CAP];  // its new keys in order
  __shared__ uint64_t mk[CAP], mv[CAP];
  __shared__ uint32_t off[F + 1];
  const int lane = threadIdx.x;
  const uint32_t m = *dm;
  for (uint32_t k = blockIdx.x; k < n_nodes; k += gridDim.x) {
    const uint32_t node = nlist[2 * k], i0 = nlist[2 * k + 1];
    const uint32_t x0 = node * (uint32_t)F;
    const uint32_t c = (x0 + F < n_leaves ? x0 + F : n_leaves) - x0;
    // leaf offsets in the node's key list: counts, then a wave scan (c <= F <= 128: two per lane)
    uint32_t c0 = 0, c1 = 0;
    if ((uint32_t)lane < c) c0 = (uint32_t)leaf_count(leaves[x0 + lane]);
    if ((uint32_t)lane + 64 < c) c1 = (uint32_t)leaf_count(leaves[x0 + 64 + lane]);
    const uint32_t s0 = wave_inclusive_scan(c0);
    const uint32_t t0 = (uint32_t)__shfl((int)s0, 63);
    const uint32_t s1 = wave_inclusive_scan(c1) + t0;
    if ((uint32_t)lane < c) off[lane] = s0 - c0;
    if ((uint32_t)lane + 64 < c) off[64 + lane] = s1 - c1;
    const uint32_t ke = (uint32_t)__shfl((int)s1, 63);  // existing keys
    if (lane == 0) off[c] = ke;
    __syncthreads();
    for (uint32_t u = lane; u < c * LEAF_KEYS; u += 64) {
      const uint32_t s = u / LEAF_KEYS, q = u % LEAF_KEYS;
      if (q < off[s + 1] - off[s]) {
        ek[off[s] + q] = leaves[x0 + s].key[q];
        ev[off[s] + q] = leaves[x0 + s].val[q];
      }
    }
    __syncthreads();
    // the batch entries of this node: values of existing keys, and the new keys compacted in order
    uint32_t kn = 0;
    for (uint32_t j0 = i0;; j0 += 64) {  // uniform: the node's entries are contiguous from i0
      const uint32_t j = j0 + lane;
      const bool in = j < m && leaf[j] / (uint32_t)F == node;
      const uint8_t f = in ? (found[j] & 1u) : 0u;
      if (in && f) ev[off[leaf[j] - x0] + slot[j]] = uv[j];
      const uint32_t isnew = in && !f ? 1u : 0u;
      const uint32_t inc = wave_inclusive_scan(isnew);
      if (isnew) {
        nk[kn + inc - 1] = uk[j];
        nv[kn + inc - 1] = uv[j];
      }
      kn += (uint32_t)__shfl((int)inc, 63);
      if (!__any(in) || j0 + 64 >= m) break;
    }
    __syncthreads();
    // merge: an element's place = its index + how many of the other list are smaller (keys distinct)
    const uint32_t K = ke + kn;
    for (uint32_t a = lane; a < ke; a += 64) {
      uint32_t lo = 0, hi = kn;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (nk[mid] < ek[a]) lo = mid + 1; else hi = mid;
      }
      mk[a + lo] = ek[a];
      mv[a + lo] = ev[a];
    }
    for (uint32_t b = lane; b < kn; b += 64) {
      uint32_t lo = 0, hi = ke;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ek[mid] < nk[b]) lo = mid + 1; else hi = mid;
      }
      mk[b + lo] = nk[b];
      mv[b + lo] = nv[b];
    }
    __syncthreads();
    // leaves: leaf s = merged [s K / c, (s + 1) K / c), padded
    for (uint32_t u = lane; u < c * 2 * LEAF_KEYS; u += 64) {
      const uint32_t s = u / (2 * LEAF_KEYS), q = u % (2 * LEAF_KEYS);
      const uint32_t lo = s * K / c, hi = (s + 1) * K / c;
      const uint32_t p = lo + (q % LEAF_KEYS);
      Leaf* L = leaves + x0 + s;
      if (q < LEAF_KEYS)
        L->key[q] = p < hi ? mk[p] : KEY_MAX;
      else
        L->val[q - LEAF_KEYS] = p < hi ? mv[p] : 0ull;
    }
    // the node: separators, then the index slots below the last
    uint64_t* nd = nodes + (size_t)node * NS;
    for (uint32_t s = lane; s < (uint32_t)F; s += 64)
      nd[IS + s] = s + 1 < c ? mk[(s + 1) * K / c] : KEY_MAX;
    for (uint32_t g = lane; g + 1 < (uint32_t)IS; g += 64) {
      const uint32_t ks = 8 * g + 7;
      nd[g] = (g < (uint32_t)(F / 8) && ks + 1 < c) ? mk[(ks + 1) * K / c] : KEY_MAX;
    }
    __syncthreads();  // the LDS lists are reused by the next node
  }
}

__global__ void apply_runs_kernel(Leaf* __restrict__ leaves, const uint32_t* __restrict__ leaf,
                                  const uint32_t* __restrict__ slot, const uint8_t* __restrict__ found,
                                  const uint64_t* __restrict__ uk, const uint64_t* __restrict__ uv, size_t n_max,
                                  const uint32_t* __restrict__ dm) {
  const uint32_t m = dm ? *dm : (uint32_t)n_max;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    if (!run_head(leaf, i) || (found[i] & 2u)) continue;  // marked: the node re-spread handles it
    const uint32_t L = leaf[i];
    uint32_t ins = 0;
    for (uint32_t j = i; j < m && leaf[j] == L; ++j) {
      if (found[j])
        leaves[L].val[slot[j]] = uv[j];
      else
        ++ins;
    }
    if (!ins) continue;
    const LeafRegs l = load_leaf(leaves + L);  // this thread's value writes above are visible to its own load
    LeafRegs o;
#pragma unroll
    for (int s = 0; s < LEAF_KEYS; ++s) {
      o.key[s] = KEY_MAX;
      o.val[s] = 0ull;
    }
    // the plan guarantees at most LEAF_KEYS merged keys
    merge_leaf(l, ins, i, uk, uv, found, [&](int idx, uint64_t k, uint64_t v) {
      put(o.key, idx, k);  // idx >= LEAF_KEYS matches no slot
      put(o.val, idx, v);
    });
    store_leaf(leaves + L, o);
  }
}

__global__ void edge_keys_kernel(const uint64_t* __restrict__ uk, const uint32_t* __restrict__ dm,
                                 uint64_t* __restrict__ edge) {
  if (threadIdx.x == 0) {
    const uint32_t m = *dm;
    edge[0] = m ? uk[0] : KEY_MAX;
    edge[1] = m ? uk[m - 1] : 0ull;
  }
}

}  // namespace

hipError_t edge_keys(const uint64_t* uk, const uint32_t* dm, uint64_t* edge, hipStream_t s) {
  hipLaunchKernelGGL(edge_keys_kernel, dim3(1), dim3(64), 0, s, uk, dm, edge);
  return hipGetLastError();
}

hipError_t small_sort_dedupe(const uint64_t* pk, const uint64_t* pv, uint32_t n, uint64_t* uk, uint64_t* uv,
                             uint32_t* dm, hipStream_t s) {
  if (n == 0 || n > kSmallUpdate) return hipErrorInvalidValue;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(small_sort_dedupe_kernel, dim3(1), dim3(kSortThreads), 0, s, pk, pv, n, uk, uv, dm);
  return hipGetLastError();
}

hipError_t merge_sort_pairs(const uint64_t* pk, const uint64_t* pv, uint32_t n, uint64_t* k0, uint64_t* k1,
                            uint32_t* i0, uint32_t* i1, uint64_t* vout, hipStream_t s, const uint64_t** sorted_keys) {
  if (n == 0) return hipErrorInvalidValue;
  ProfScope prof("update", s);
  const uint32_t chunks = (n + kSmallUpdate - 1) / kSmallUpdate;
  hipLaunchKernelGGL(chunk_sort_kernel, dim3(chunks), dim3(kSortThreads), 0, s, pk, n, k0, i0);
  uint64_t *ka = k0, *kb = k1;
  uint32_t *ia = i0, *ib = i1;
  const unsigned grid = (unsigned)((n + 255) / 256);
  for (uint64_t L = kSmallUpdate; L < n; L *= 2) {
    hipLaunchKernelGGL(merge_pass_kernel, dim3(grid), dim3(256), 0, s, ka, ia, kb, ib, n, (uint32_t)L);
    std::swap(ka, kb);
    std::swap(ia, ib);
  }
  hipLaunchKernelGGL(gather_by_index_kernel, dim3(grid), dim3(256), 0, s, pv, ia, n, vout);
  *sorted_keys = ka;
  return hipGetLastError();
}

hipError_t unique_count(const uint32_t* keep, const uint32_t* keep_scan, size_t n, uint32_t* dm, hipStream_t s) {
  if (!n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(unique_count_kernel, dim3(1), dim3(64), 0, s, keep, keep_scan, n, dm);
  return hipGetLastError();
}

hipError_t plan_runs(const Leaf* leaves, uint32_t n_leaves, int F, const uint32_t* leaf, const uint8_t* found, size_t n,
                     const uint32_t* dm, uint32_t* summary, uint32_t* nlist, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(plan_runs_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaves, n_leaves, F, leaf, found, n,
                     dm, summary, nlist);
  return hipGetLastError();
}

hipError_t mark_nodes(const uint32_t* nlist, uint32_t n_nodes, int F, const uint32_t* leaf, uint8_t* found,
                      const uint32_t* dm, hipStream_t s) {
  if (!n_nodes) return hipSuccess;
  ProfScope prof("update", s);
  const unsigned g = n_nodes < 4096 ? n_nodes : 4096;
  hipLaunchKernelGGL(mark_nodes_kernel, dim3(g), dim3(64), 0, s, nlist, n_nodes, F, leaf, found, dm);
  return hipGetLastError();
}

hipError_t node_respread(Leaf* leaves, uint64_t* last_level, uint32_t n_leaves, int F, const uint32_t* nlist,
                         uint32_t n_nodes, const uint32_t* leaf, const uint32_t* slot, const uint8_t* found,
                         const uint64_t* uk, const uint64_t* uv, const uint32_t* dm, hipStream_t s) {
  if (!n_nodes) return hipSuccess;
  ProfScope prof("update", s);
  const unsigned g = n_nodes < 4096 ? n_nodes : 4096;
  switch (F) {
    case 32:
      hipLaunchKernelGGL(node_respread_kernel<32>, dim3(g), dim3(64), 0, s, leaves, last_level, n_leaves, nlist,
                         n_nodes, leaf, slot, found, uk, uv, dm);
      break;
    case 64:
      hipLaunchKernelGGL(node_respread_kernel<64>, dim3(g), dim3(64), 0, s, leaves, last_level, n_leaves, nlist,
                         n_nodes, leaf, slot, found, uk, uv, dm);
      break;
    case 128:
      hipLaunchKernelGGL(node_respread_kernel<128>, dim3(g), dim3(64), 0, s, leaves, last_level, n_leaves, nlist,
                         n_nodes, leaf, slot, found, uk, uv, dm);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t apply_runs(Leaf* leaves, const uint32_t* leaf, const uint32_t* slot, const uint8_t* found,
                      const uint64_t* uk, const uint64_t* uv, size_t n, const uint32_t* dm, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(apply_runs_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaves, leaf, slot, found, uk, uv, n,
                     dm);
  return hipGetLastError();
}

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

hipError_t count_inserts(const uint32_t* leaf, const uint8_t* found, size_t n, uint32_t* ins_per_leaf,
                         hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(count_inserts_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf, found, n, ins_per_leaf);
  return hipGetLastError();
}

hipError_t new_leaf_counts(const Leaf* leaves, const uint32_t* ins_per_leaf, size_t n_leaves, int fill, uint32_t* cnt,
                           hipStream_t s) {
  ProfScope prof("update", s);
  hipLaunchKernelGGL(new_counts_kernel, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, leaves, ins_per_leaf,
                     n_leaves, fill, cnt);
  return hipGetLastError();
}

hipError_t leaf_first_keys(const uint32_t* leaf, size_t n, uint32_t* first, hipStream_t s) {
  if (!n) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(leaf_first_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf, n, first);
  return hipGetLastError();
}

hipError_t relayout_copy(const Leaf* old_leaves, size_t lo, size_t hi, const uint32_t* first, const uint32_t* new_start,
                         Leaf* new_leaves, uint64_t* new_mins, hipStream_t s) {
  if (hi <= lo) return hipSuccess;
  ProfScope prof("update", s);
  // a few blocks per CU: the copy streams at full rate and searches on other streams still get CUs
  const size_t pieces = (hi - lo) * (sizeof(Leaf) / 16);
  size_t nb = (pieces + 255) / 256;
  const size_t cap = (size_t)launch_cfg().num_cus * 4;
  if (nb > cap) nb = cap;
  hipLaunchKernelGGL(relayout_copy_kernel, dim3((unsigned)nb), dim3(256), 0, s, old_leaves, lo, hi, first, new_start,
                     new_leaves, new_mins);
  return hipGetLastError();
}

hipError_t relayout_touched(const Leaf* old_leaves, const uint32_t* ins_per_leaf, const uint32_t* bleaf,
                            const uint32_t* bslot, const uint64_t* bkeys, const uint64_t* bvals, const uint8_t* found,
                            size_t m, const uint32_t* new_start, int fill, Leaf* new_leaves, uint64_t* new_mins,
                            hipStream_t s) {
  if (!m) return hipSuccess;
  ProfScope prof("update", s);
  hipLaunchKernelGGL(relayout_touched_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, old_leaves, ins_per_leaf,
                     bleaf, bslot, bkeys, bvals, found, m, new_start, fill, new_leaves, new_mins);
  return hipGetLastError();
}

}  // namespace hm
