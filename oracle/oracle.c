/*
 * oracle.c -- CPU oracle for the Harmonia B+tree path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load this
 * library, and only as the checker or as the timed CPU reference path; the product
 * (libharmonia_amd) never calls it.
 *
 * It restates, in plain C, the reference algorithms the GPU path must match:
 *   - hmo_build_bfs:    breadth-first relayout with the prefix recurrence prefix[0] = 1,
 *                       prefix[k+1] = prefix[k] + children_k (H/regular_bpt/traversal.cpp:36-78,
 *                       H/src/traversal.cpp:180-195), over the bulk-built layout (see below);
 *   - hmo_search_cfb:   CFB::search (H/regular_bpt/cfb.cpp:4-41): per inner node a linear scan
 *                       `while (i < F-1 && key >= k[i]) i++` (upper_bound), next = prefix[node] + i,
 *                       then an equality scan of the leaf (regular find, H/regular_bpt/regular_bpt.cpp:383-393);
 *                       the OpenMP form restates search_cpu_multithread (H/regular_bpt/main.cpp:92-116);
 *   - hmo_search_harmonia: Harmonia's BPlusTree::search (H/src/bpt.cpp:673-680) over this layout:
 *                       per inner node the index-then-key step of getChildIdx_avx2_version2
 *                       (H/src/avx.cpp:2-54: 4 index slots compared at once, the next 4 only if all
 *                       4 are <= key; then the same 4 + 4 over the chosen 8-key group), next =
 *                       prefix[node] + slot (search_last_level_inner_node_PPItree,
 *                       H/src/bpt-ppi.cpp:118-140), then findKey's backward equality scan of the
 *                       leaf (H/src/bpt.h:252-274); the OpenMP form restates search_cpu_test
 *                       (H/src/main.cpp:772-812).  The reference's keys are signed with padding
 *                       LLONG_MAX; here keys are unsigned with padding 2^64-1, so the AVX2 signed
 *                       compares run on keys with the top bit flipped (same order);
 *   - hmo_search_sorted: the closed form: lower_bound in the sorted key array + equality;
 *   - hmo_psa_sort:     cub::DeviceRadixSort::SortPairs(keys, pos, n, begin_bit, end_bit)
 *                       (H/regular_bpt/prefix-sort.cu:288-293) = stable sort of (key, pos) by the
 *                       key bits [begin_bit, end_bit);
 *   - hmo_range_walk:   the leaf-chain scan of H/src/range.cu:168-223 / BPlusTree::range_search
 *                       (H/src/bpt.cpp:695-769) generalised to a per-range length;
 *   - hmo_range_sorted: the closed form slice [lower_bound(start), +len) of the sorted keys;
 *   - hmo_upsert_sorted: closed form of a batched upsert (last write of a key wins).
 *
 * Layout shared with the device build (values are bit-compared in the tests):
 *   leaves: n_leaves = max(1, ceil(n / fill)); leaf i = keys [i*fill, min((i+1)*fill, n)),
 *           16 u64 per leaf: key[8] (KEY_MAX padding) then val[8] (0 padding);
 *   levels: m_H = n_leaves, m_l = ceil(m_{l+1} / F) until 1; node j of level l has children
 *           j*F .. min((j+1)*F, m_{l+1}) - 1 of level l+1;
 *   node:   IS index slots (IS = max(8, F/8); idx[j] = key[8j+7] for j < F/8, else KEY_MAX)
 *           then F key slots: key[s] = min key of child s+1 for s < nchild-1, else KEY_MAX.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KEY_MAX UINT64_MAX
#define LEAF_KEYS 8
#define MAX_LEVELS 16

static int idx_slots(int F) { return F / 8 < 8 ? 8 : F / 8; }
static int node_slots(int F) { return idx_slots(F) + F; }

/* Vigna's splitmix64: state += golden gamma, then the two xor-shift-multiply rounds. */
uint64_t hmo_splitmix64_next(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* value of a key: the first output of a splitmix64 stream whose state is the key */
uint64_t hmo_splitmix64_mix(uint64_t key) {
  uint64_t s = key;
  return hmo_splitmix64_next(&s);
}

/* Level plan for n keys.  Returns height (inner levels) or -1. */
int hmo_plan(uint64_t n, int F, int fill, uint32_t* level_start /* MAX_LEVELS+1 */, uint32_t* n_inner,
             uint64_t* n_leaves) {
  if ((F != 32 && F != 64 && F != 128) || fill < 1 || fill > 8) return -1;
  uint64_t nl = n == 0 ? 1 : (n + (uint64_t)fill - 1) / (uint64_t)fill;
  uint64_t sizes[MAX_LEVELS + 1];
  int h = 0;
  uint64_t m = nl;
  do {
    m = (m + (uint64_t)F - 1) / (uint64_t)F;
    if (h >= MAX_LEVELS) return -1;
    sizes[h++] = m;
  } while (m > 1);
  /* sizes are bottom-up; level 0 = root */
  uint64_t acc = 0;
  for (int l = 0; l < h; ++l) {
    level_start[l] = (uint32_t)acc;
    acc += sizes[h - 1 - l];
  }
  for (int l = h; l <= MAX_LEVELS; ++l) level_start[l] = (uint32_t)acc;
  *n_inner = (uint32_t)acc;
  *n_leaves = nl;
  return h;
}

/* Size of the subtree unit: number of leaves under one node of level l. */
static uint64_t leaves_under(int l, int h, int F) {
  uint64_t u = 1;
  for (int i = l; i < h; ++i) u *= (uint64_t)F;
  return u;
}

/* Breadth-first emission of the layout from strictly ascending keys.  A queue of (level, j) is
 * processed in order; each inner node writes its separators, index granule and prefix entry
 * through the running recurrence, as traverse_transform does. */
int hmo_build_bfs(const uint64_t* keys, const uint64_t* vals, uint64_t n, int F, int fill, uint64_t* inner,
                  uint32_t* prefix, uint64_t* leaves, uint32_t* level_start, uint32_t* out_height,
                  uint32_t* out_n_inner, uint64_t* out_n_leaves) {
  uint32_t n_inner = 0;
  uint64_t nl = 0;
  const int h = hmo_plan(n, F, fill, level_start, &n_inner, &nl);
  if (h < 0) return -1;
  const int IS = idx_slots(F), NS = node_slots(F);
  /* leaves */
  for (uint64_t i = 0; i < nl; ++i) {
    for (int s = 0; s < LEAF_KEYS; ++s) {
      const uint64_t src = i * (uint64_t)fill + (uint64_t)s;
      const int have = s < fill && src < n;
      leaves[i * 16 + s] = have ? keys[src] : KEY_MAX;
      leaves[i * 16 + 8 + s] = have ? vals[src] : 0;
    }
  }
  /* BFS over inner nodes: the queue is implicit because every level is a contiguous run of
   * (level, j) pairs; node k of the BFS order is (l, k - level_start[l]). */
  prefix[0] = 1;
  for (uint32_t k = 0; k < n_inner; ++k) {
    int l = 0;
    while (l + 1 < h && k >= level_start[l + 1]) ++l;
    const uint64_t j = k - level_start[l];
    const uint64_t m_child = (l + 1 < h) ? (uint64_t)(level_start[l + 2 <= h ? l + 2 : h] - level_start[l + 1]) : nl;
    const uint64_t first_child = j * (uint64_t)F;
    const uint64_t nchild = (m_child - first_child) < (uint64_t)F ? (m_child - first_child) : (uint64_t)F;
    const uint64_t unit = leaves_under(l + 1, h, F); /* leaves under one child */
    uint64_t* node = inner + (uint64_t)k * NS;
    uint64_t* kslot = node + IS;
    for (int s = 0; s < F; ++s) {
      uint64_t v = KEY_MAX;
      if ((uint64_t)s + 1 < nchild) {
        const uint64_t leaf = (first_child + (uint64_t)s + 1) * unit; /* leftmost leaf of child s+1 */
        v = leaves[leaf * 16 + 0];
      }
      kslot[s] = v;
    }
    for (int s = 0; s < IS; ++s) node[s] = (s < F / 8) ? kslot[8 * s + 7] : KEY_MAX;
    if (k + 1 < n_inner) prefix[k + 1] = prefix[k] + (uint32_t)nchild;
  }
  *out_height = (uint32_t)h;
  *out_n_inner = n_inner;
  *out_n_leaves = nl;
  return 0;
}

typedef struct {
  const uint64_t* inner;
  const uint32_t* prefix;
  const uint64_t* leaves;
  uint32_t n_inner;
  uint32_t height;
  int F;
} hmo_view;

/* CFB::search over the layout (one query). */
static int cfb_find(const hmo_view* t, uint64_t key, uint64_t* val) {
  if (key == KEY_MAX) return 0; /* padding value: never stored */
  const int IS = idx_slots(t->F), NS = node_slots(t->F);
  uint64_t idx = 0;
  for (uint32_t h = 0; h < t->height; ++h) {
    const uint64_t* k = t->inner + idx * NS + IS;
    int i = 0;
    while (i < t->F - 1 && key >= k[i]) i++;
    idx = t->prefix[idx] + (uint64_t)i;
  }
  const uint64_t* lf = t->leaves + (idx - t->n_inner) * 16;
  for (int i = 0; i < LEAF_KEYS; ++i)
    if (lf[i] == key) {
      *val = lf[8 + i];
      return 1;
    }
  return 0;
}

void hmo_search_cfb(const uint64_t* inner, const uint32_t* prefix, const uint64_t* leaves, uint32_t n_inner,
                    uint32_t height, int F, const uint64_t* q, uint64_t m, uint64_t* out_vals, uint8_t* out_found,
                    int nthreads) {
  hmo_view t = {inner, prefix, leaves, n_inner, height, F};
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    uint64_t v = 0;
    const int f = cfb_find(&t, q[i], &v);
    out_vals[i] = f ? v : 0;
    out_found[i] = (uint8_t)f;
  }
  (void)nthreads;
}

/* --- Harmonia's index-then-key node search (BPlusTree::search) --- */
#if defined(__x86_64__)
#include <immintrin.h>
/* #entries <= key among 4 consecutive u64 (signed AVX2 compare after flipping the top bit) */
__attribute__((target("avx2"))) static inline int le4_avx2(const uint64_t* p, __m256i vq) {
  const __m256i flip = _mm256_set1_epi64x((long long)0x8000000000000000ull);
  const __m256i v = _mm256_xor_si256(_mm256_loadu_si256((const __m256i*)p), flip);
  const __m256i le = _mm256_or_si256(_mm256_cmpgt_epi64(vq, v), _mm256_cmpeq_epi64(vq, v));
  return __builtin_popcount((unsigned)_mm256_movemask_epi8(le) & 0x01010101u);
}
/* rank inside a group of 4 + 4: the second half is compared only if the whole first half is <= key */
__attribute__((target("avx2"))) static inline int le8_avx2(const uint64_t* p, __m256i vq) {
  int k = le4_avx2(p, vq);
  if (k == 4) k += le4_avx2(p + 4, vq);
  return k;
}
__attribute__((target("avx2"))) static int harmonia_find_avx2(const hmo_view* t, uint64_t key, uint64_t* val) {
  if (key == KEY_MAX) return 0;
  const int IS = idx_slots(t->F), NS = node_slots(t->F);
  const __m256i vq = _mm256_set1_epi64x((long long)(key ^ 0x8000000000000000ull));
  uint64_t idx = 0;
  for (uint32_t h = 0; h < t->height; ++h) {
    const uint64_t* nd = t->inner + idx * NS;
    int j = 0; /* index slots <= key, groups of 8 as 4 + 4 */
    for (int g = 0; g < IS; g += 8) {
      const int c = le8_avx2(nd + g, vq);
      j += c;
      if (c < 8) break;
    }
    const int s = 8 * j + le8_avx2(nd + IS + 8 * j, vq);
    idx = t->prefix[idx] + (uint64_t)s;
  }
  const uint64_t* lf = t->leaves + (idx - t->n_inner) * 16;
  for (int i = LEAF_KEYS - 1; i >= 0; --i) /* findKey scans backwards */
    if (lf[i] == key) {
      *val = lf[8 + i];
      return 1;
    }
  return 0;
}
#endif

/* the same steps without AVX2 */
static int harmonia_find_scalar(const hmo_view* t, uint64_t key, uint64_t* val) {
  if (key == KEY_MAX) return 0;
  const int IS = idx_slots(t->F), NS = node_slots(t->F);
  uint64_t idx = 0;
  for (uint32_t h = 0; h < t->height; ++h) {
    const uint64_t* nd = t->inner + idx * NS;
    int j = 0;
    while (j < IS && nd[j] <= key) ++j;
    const uint64_t* g = nd + IS + 8 * j;
    int c = 0;
    while (c < 8 && g[c] <= key) ++c;
    idx = t->prefix[idx] + (uint64_t)(8 * j + c);
  }
  const uint64_t* lf = t->leaves + (idx - t->n_inner) * 16;
  for (int i = LEAF_KEYS - 1; i >= 0; --i)
    if (lf[i] == key) {
      *val = lf[8 + i];
      return 1;
    }
  return 0;
}

/* use_avx2: 1 = AVX2 path (falls back to scalar on a CPU without AVX2), 0 = scalar path */
void hmo_search_harmonia(const uint64_t* inner, const uint32_t* prefix, const uint64_t* leaves, uint32_t n_inner,
                         uint32_t height, int F, const uint64_t* q, uint64_t m, uint64_t* out_vals,
                         uint8_t* out_found, int nthreads, int use_avx2) {
  hmo_view t = {inner, prefix, leaves, n_inner, height, F};
#if defined(__x86_64__)
  const int avx = use_avx2 && __builtin_cpu_supports("avx2");
#else
  const int avx = 0;
  (void)use_avx2;
#endif
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    uint64_t v = 0;
#if defined(__x86_64__)
    const int f = avx ? harmonia_find_avx2(&t, q[i], &v) : harmonia_find_scalar(&t, q[i], &v);
#else
    const int f = harmonia_find_scalar(&t, q[i], &v);
#endif
    out_vals[i] = f ? v : 0;
    out_found[i] = (uint8_t)f;
  }
  (void)nthreads;
}

/* 1 if this CPU runs the AVX2 path */
int hmo_have_avx2(void) {
#if defined(__x86_64__)
  return __builtin_cpu_supports("avx2") ? 1 : 0;
#else
  return 0;
#endif
}

static uint64_t lower_bound_u64(const uint64_t* a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

void hmo_search_sorted(const uint64_t* keys, const uint64_t* vals, uint64_t n, const uint64_t* q, uint64_t m,
                       uint64_t* out_vals, uint8_t* out_found) {
  for (uint64_t i = 0; i < m; ++i) {
    const uint64_t p = lower_bound_u64(keys, n, q[i]);
    const int f = p < n && keys[p] == q[i] && q[i] != KEY_MAX;
    out_vals[i] = f ? vals[p] : 0;
    out_found[i] = (uint8_t)f;
  }
}

/* --- PSA: stable sort of (key, pos) by key bits [begin_bit, end_bit) --- */
static int g_begin, g_end;
static const uint64_t* g_keys;
static uint64_t digit_of(uint64_t k) {
  const int w = g_end - g_begin;
  const uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1ull);
  return (k >> g_begin) & mask;
}
static int cmp_pos(const void* a, const void* b) {
  const uint32_t pa = *(const uint32_t*)a, pb = *(const uint32_t*)b;
  const uint64_t da = digit_of(g_keys[pa]), db = digit_of(g_keys[pb]);
  if (da != db) return da < db ? -1 : 1;
  return pa < pb ? -1 : (pa > pb ? 1 : 0); /* stability: original position breaks ties */
}
int hmo_psa_sort(const uint64_t* keys, uint64_t n, int begin_bit, int end_bit, uint64_t* keys_out, uint32_t* pos_out) {
  if (begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) return -1;
  for (uint64_t i = 0; i < n; ++i) pos_out[i] = (uint32_t)i;
  if (begin_bit < end_bit) {
    g_begin = begin_bit;
    g_end = end_bit;
    g_keys = keys;
    qsort(pos_out, n, sizeof(uint32_t), cmp_pos);
  }
  for (uint64_t i = 0; i < n; ++i) keys_out[i] = keys[pos_out[i]];
  return 0;
}

/* --- ranges --- */
/* Leaf-chain walk: descend to the leaf of `start` as cfb_find does, then walk leaf i -> i+1
 * skipping padding, collecting keys >= start (and <= hi) until len keys. */
void hmo_range_walk(const uint64_t* inner, const uint32_t* prefix, const uint64_t* leaves, uint32_t n_inner,
                    uint64_t n_leaves, uint32_t height, int F, const uint64_t* start, const uint64_t* hi,
                    const uint32_t* len, uint32_t cap_const, uint64_t n, const uint64_t* offsets,
                    uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts) {
  const int IS = idx_slots(F), NS = node_slots(F);
  for (uint64_t r = 0; r < n; ++r) {
    const uint64_t s = start[r];
    const uint64_t h = hi ? hi[r] : KEY_MAX;
    const uint32_t L = len ? len[r] : cap_const;
    const uint64_t off = offsets ? offsets[r] : r * (uint64_t)cap_const;
    uint32_t got = 0;
    if (s != KEY_MAX && s <= h && L > 0) {
      uint64_t idx = 0;
      for (uint32_t lv = 0; lv < height; ++lv) {
        const uint64_t* k = inner + idx * NS + IS;
        int i = 0;
        while (i < F - 1 && s >= k[i]) i++;
        idx = prefix[idx] + (uint64_t)i;
      }
      uint64_t leaf = idx - n_inner;
      int stop = 0;
      for (; leaf < n_leaves && !stop && got < L; ++leaf) {
        const uint64_t* lf = leaves + leaf * 16;
        for (int i = 0; i < LEAF_KEYS && got < L; ++i) {
          const uint64_t k = lf[i];
          if (k == KEY_MAX || k < s) continue;
          if (k > h) {
            stop = 1;
            break;
          }
          out_keys[off + got] = k;
          out_vals[off + got] = lf[8 + i];
          got++;
        }
      }
    }
    out_counts[r] = got;
  }
}

void hmo_range_sorted(const uint64_t* keys, const uint64_t* vals, uint64_t n, const uint64_t* start,
                      const uint64_t* hi, const uint32_t* len, uint32_t cap_const, uint64_t m,
                      const uint64_t* offsets, uint64_t* out_keys, uint64_t* out_vals, uint32_t* out_counts) {
  for (uint64_t r = 0; r < m; ++r) {
    const uint64_t s = start[r];
    const uint64_t h = hi ? hi[r] : KEY_MAX;
    const uint32_t L = len ? len[r] : cap_const;
    const uint64_t off = offsets ? offsets[r] : r * (uint64_t)cap_const;
    uint32_t got = 0;
    if (s != KEY_MAX && s <= h) {
      for (uint64_t p = lower_bound_u64(keys, n, s); p < n && got < L && keys[p] <= h; ++p) {
        out_keys[off + got] = keys[p];
        out_vals[off + got] = vals[p];
        got++;
      }
    }
    out_counts[r] = got;
  }
}

/* --- batched upsert closed form: (keys, vals) sorted map; batch applied in order --- */
static const uint64_t* g_bk;
static int cmp_batch(const void* a, const void* b) {
  const uint64_t ia = *(const uint64_t*)a, ib = *(const uint64_t*)b;
  if (g_bk[ia] != g_bk[ib]) return g_bk[ia] < g_bk[ib] ? -1 : 1;
  return ia < ib ? -1 : (ia > ib ? 1 : 0);
}
/* out arrays must hold n + m entries; returns the new key count */
uint64_t hmo_upsert_sorted(const uint64_t* keys, const uint64_t* vals, uint64_t n, const uint64_t* bk,
                           const uint64_t* bv, uint64_t m, uint64_t* out_keys, uint64_t* out_vals) {
  uint64_t* order = (uint64_t*)malloc((m ? m : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) order[i] = i;
  g_bk = bk;
  qsort(order, m, sizeof(uint64_t), cmp_batch);
  uint64_t i = 0, j = 0, o = 0;
  while (i < n || j < m) {
    if (j < m) {
      /* last write of this batch key */
      uint64_t jj = j;
      while (jj + 1 < m && bk[order[jj + 1]] == bk[order[j]]) jj++;
      const uint64_t bkey = bk[order[jj]];
      if (i < n && keys[i] < bkey) {
        out_keys[o] = keys[i];
        out_vals[o++] = vals[i++];
        continue;
      }
      if (i < n && keys[i] == bkey) i++; /* replaced */
      out_keys[o] = bkey;
      out_vals[o++] = bv[order[jj]];
      j = jj + 1;
    } else {
      out_keys[o] = keys[i];
      out_vals[o++] = vals[i++];
    }
  }
  free(order);
  return o;
}

/* --- structural invariants of a layout (used on device-built trees too) ---
 * returns 0 if OK, else a code naming the first violated invariant. */
int hmo_check_layout(const uint64_t* inner, const uint32_t* prefix, const uint64_t* leaves, uint32_t n_inner,
                     uint64_t n_leaves, uint32_t height, const uint32_t* level_start, int F) {
  const int IS = idx_slots(F), NS = node_slots(F);
  if (prefix[0] != 1) return 1;
  /* prefix recurrence: prefix[k+1] - prefix[k] = children of k, and children are contiguous */
  uint64_t total_children = 0;
  for (uint32_t k = 0; k < n_inner; ++k) {
    const uint64_t* node = inner + (uint64_t)k * NS;
    const uint64_t* ks = node + IS;
    int nsep = 0;
    while (nsep < F && ks[nsep] != KEY_MAX) nsep++;
    for (int s = nsep; s < F; ++s)
      if (ks[s] != KEY_MAX) return 2; /* padding must be a suffix */
    for (int s = 1; s < nsep; ++s)
      if (ks[s - 1] >= ks[s]) return 3; /* separators strictly ascending */
    if (nsep > F - 1) return 4;
    for (int s = 0; s < IS; ++s)
      if (node[s] != ((s < F / 8) ? ks[8 * s + 7] : KEY_MAX)) return 5; /* index granule */
    const uint64_t nchild = (uint64_t)nsep + 1;
    if (k + 1 < n_inner && prefix[k + 1] != prefix[k] + nchild) return 6;
    total_children += nchild;
  }
  if (total_children != (uint64_t)n_inner - 1 + n_leaves) return 7;
  /* every separator = min key of its right child subtree; leaves sorted, padding suffix */
  for (uint64_t i = 0; i < n_leaves; ++i) {
    const uint64_t* lf = leaves + i * 16;
    int c = 0;
    while (c < LEAF_KEYS && lf[c] != KEY_MAX) c++;
    for (int s = c; s < LEAF_KEYS; ++s)
      if (lf[s] != KEY_MAX) return 8;
    for (int s = 1; s < c; ++s)
      if (lf[s - 1] >= lf[s]) return 9;
    if (i + 1 < n_leaves && c > 0 && leaves[(i + 1) * 16] != KEY_MAX && lf[c - 1] >= leaves[(i + 1) * 16])
      return 10; /* leaf chain in key order */
  }
  for (uint32_t k = 0; k < n_inner; ++k) {
    const uint64_t* ks = inner + (uint64_t)k * NS + IS;
    for (int s = 0; s < F - 1 && ks[s] != KEY_MAX; ++s) {
      /* leftmost leaf under child s+1 */
      uint64_t c = prefix[k] + (uint64_t)s + 1;
      while (c < n_inner) c = prefix[c];
      if (leaves[(c - n_inner) * 16] != ks[s]) return 11;
    }
  }
  (void)height;
  (void)level_start;
  return 0;
}
