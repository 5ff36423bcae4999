// Hand-written stable LSD radix sort (8-bit digits) for u64 keys, with or without a payload.
//
// The PSA step of the reference sorts (key, position) pairs with
// cub::DeviceRadixSort::SortPairs(keys, pos, n, 48, 64) before every search launch
// (H/regular_bpt/prefix-sort.cu:288-293, H/src/full-multi-prefix-sort-8thread.cu:305-309).
// That call is a stable LSD radix sort over key bits [begin_bit, end_bit); for keys in
// [0, 2^63) CUB's signed-key bit twiddle (H/cub/cub/util_type.cuh:1015-1022) does not change the
// order, so the result equals a stable sort of (key, pos) by key >> begin_bit.  The same kernels
// sort the bulk-build input (u64 value payload over every varying bit).
//
// One pass = histogram kernel (per-tile digit counts, digit-major) -> digit scan (one block per
// digit: exclusive scan of the digit's counts along the tiles, and the digit's total) -> scatter
// kernel.  In the scatter, the digit starts are the exclusive scan of the 256 totals (done in each
// block); a wave-level digit match (8 ballots) gives each item its stable rank inside its wave;
// per-wave counters are scanned across the block; the tile is then staged in LDS in digit order so
// the global writes of each digit run are coalesced.  Tiles are mapped to blocks so that each XCD
// processes a contiguous range of tiles: the digit runs of neighbouring tiles share cache lines at
// their ends, and those partial lines then meet in one L2.
//
// Returning results to input order: a search in PSA order produces one result per sorted query.
// Instead of sorting the results by position (three more passes over 24 position bits) or
// writing each result to its position (one random 9-byte write per query, bound by cache-line
// transactions), the keys-only sort records, per pass, the scanned (digit, tile) offsets and each
// element's index in its tile's LDS staging.  Running the passes backwards, a block reads its
// tile's digit runs from the pass output (contiguous runs: coalesced), stages them in LDS, and
// writes the tile back in its input order (coalesced).
#include "hm_block.h"
#include "hm_internal.h"

namespace hm {
namespace {

// Tiles are NT threads x kItems keys: 256 x 16 for the sorts with a payload (their LDS staging of keys
// and payload then allows several blocks per CU), 512 x 16 for the keys-only sort that PSA replays:
// its 256 digit runs per tile are then ~32 keys long instead of ~16, and the runs are what the scatter
// writes and the un-scatter gathers (each a separate stretch of HBM).
constexpr int kThreads = 256;
constexpr int kRThreads = 512;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kRTile = kRThreads * kItems;
constexpr int kWaves = kThreads / 64;
static_assert(kRTile <= 65536, "staging indices are 16-bit");

// Block -> tile: blocks are dispatched round-robin over the 8 XCDs; XCD x gets tiles
// [x * grid / 8, (x + 1) * grid / 8).  grid % 8 == 0; tiles >= nb are idle.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t grid) { return (b & 7u) * (grid >> 3) + (b >> 3); }
unsigned grid_for(size_t nb) { return (unsigned)((nb + 7) / 8 * 8); }
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Per-tile digit counts, one block per tile.  VEC: the keys are 16-byte aligned and full tiles are
// read with 16-byte loads (two keys per lane per load; the counts do not depend on which lane counts
// which key).  (A persistent grid that loads tile t + 1 while counting tile t measured no faster.)
// In full tiles a wave whose 128 keys of one load share a digit (runs of sorted keys, a hot key
// repeated, the second pass over a skewed batch) adds 128 once instead of 128 atomics serialized on one
// LDS bin.
template <typename K, bool VEC, int NT>
__global__ __launch_bounds__(NT) void radix_hist_kernel(const K* __restrict__ keys, size_t n, int shift, uint32_t mask,
                                                        uint32_t* __restrict__ counts, uint32_t nb,
                                                        const uint32_t* __restrict__ dn) {
  constexpr int TILE = NT * kItems, NW = NT / 64;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= nb) return;  // block-uniform
  if (dn) n = *dn;  // device-side count (graph replays): tiles past n count nothing
  // one histogram per wave (LDS atomics; lanes of a wave that hit one bin are serialized by the LDS,
  // which is cheaper than ranking them when only counts are needed)
  __shared__ uint32_t hist[NW][256];
  for (int i = threadIdx.x; i < NW * 256; i += NT) (&hist[0][0])[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t base = (size_t)tile * TILE + (size_t)w * 64 * kItems;
  if (VEC && (size_t)(tile + 1) * TILE <= n) {  // block-uniform
    static_assert(sizeof(K) == 8, "16-byte loads take two keys");
    const ulonglong2* src = reinterpret_cast<const ulonglong2*>(keys + base);
    ulonglong2 kk[kItems / 2];
#pragma unroll
    for (int r = 0; r < kItems / 2; ++r) kk[r] = src[r * 64 + lane];
#pragma unroll
    for (int r = 0; r < kItems / 2; ++r) {  // every lane of the wave is active here
      const uint32_t dx = (uint32_t)(kk[r].x >> shift) & mask, dy = (uint32_t)(kk[r].y >> shift) & mask;
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dx);
      if (__ballot(dx == d0 && dy == d0) == ~0ull) {  // wave-uniform: the wave's 128 keys share a digit
        if (lane == 0) atomicAdd(&hist[w][d0], 128u);
      } else {
        atomicAdd(&hist[w][dx], 1u);
        atomicAdd(&hist[w][dy], 1u);
      }
    }
  } else {
    const size_t last = n ? n - 1 : 0;
    // issue every load of the tile before the first use, so the loads overlap
    K kk[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
      const size_t i = base + (size_t)r * 64 + lane;
      kk[r] = keys[i < last ? i : last];
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
      const size_t i = base + (size_t)r * 64 + lane;
      if (i < n) atomicAdd(&hist[w][(uint32_t)(kk[r] >> shift) & mask], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < 256) {
    uint32_t c = 0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) c += hist[ww][threadIdx.x];
    counts[(size_t)threadIdx.x * nb + tile] = c;
  }
}

// Write (key, value) to (kout, vout); without HAS_VAL keys only.  With SAVE_LP, lp_out[i] = the
// index of element i in its tile's digit-ordered staging (what un-sorting needs).
template <typename K, typename V, bool HAS_VAL, bool SAVE_LP, int NT>
__global__ __launch_bounds__(NT) void radix_scatter_kernel(const K* __restrict__ kin, const V* __restrict__ vin,
                                                           K* __restrict__ kout, V* __restrict__ vout,
                                                           uint16_t* __restrict__ lp_out, size_t n, int shift,
                                                           uint32_t mask, const uint32_t* __restrict__ offsets,
                                                           uint32_t nb, int gen_pos, const uint32_t* __restrict__ dn) {
  constexpr int TILE = NT * kItems, NW = NT / 64;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (dn) n = *dn;
  if (tile >= nb || (size_t)tile * TILE >= n) return;  // whole tile past the end (block-uniform)
  __shared__ uint32_t wcnt[NW][256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t boff[256];
  __shared__ uint32_t scan_sh[NW + 1];
  __shared__ K skeys[TILE];
  __shared__ V svals[HAS_VAL ? TILE : 1];
  for (int i = threadIdx.x; i < NW * 256; i += NT) (&wcnt[0][0])[i] = 0;
  {
    // digit d starts at the digits below it (scan of the totals after the 256 x nb counts) plus its
    // count in the tiles before this one (threads >= 256 take part in the scans with zeros)
    // (both loads issued before the scan: the tile's own offset does not depend on it)
    const bool dig = threadIdx.x < 256;
    const uint32_t tot = dig ? offsets[(size_t)256 * nb + threadIdx.x] : 0u;
    const uint32_t mine = dig ? offsets[(size_t)threadIdx.x * nb + tile] : 0u;
    uint32_t all;
    const uint32_t dbase = block_exclusive_scan<uint32_t, NT>(tot, scan_sh, all);
    if (dig) goff[threadIdx.x] = dbase + mine;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const size_t base = (size_t)tile * TILE + (size_t)w * 64 * kItems;
  K kk[kItems];
  V vv[HAS_VAL ? kItems : 1];
  uint32_t pk[kItems];  // (rank within wave << 8) | digit
  const size_t last = n - 1;  // n > 0 here
  // issue every load of the tile before the first use, so the loads overlap
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const size_t i = base + (size_t)r * 64 + lane;
    const size_t ic = i < last ? i : last;
    kk[r] = kin[ic];
    if constexpr (HAS_VAL) vv[r] = gen_pos ? (V)i : vin[ic];
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const size_t i = base + (size_t)r * 64 + lane;
    const bool valid = i < n;
    const K k = kk[r];
    const uint32_t d = (uint32_t)(k >> shift) & mask;
    const uint64_t peers = match_digit(d, __ballot(valid));
    const uint32_t prior = wcnt[w][d];
    if (valid && (peers & lt) == 0) wcnt[w][d] = prior + (uint32_t)__popcll(peers);
    pk[r] = ((prior + (uint32_t)__popcll(peers & lt)) << 8) | d;
  }
  __syncthreads();
  // per digit: offsets of each wave inside the digit, then the digit's start inside the tile
  {
    const int d = threadIdx.x;
    uint32_t run = 0;
    if (d < 256) {
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) {
        const uint32_t t = wcnt[ww][d];
        wcnt[ww][d] = run;
        run += t;
      }
    }
    uint32_t total;
    const uint32_t e = block_exclusive_scan<uint32_t, NT>(run, scan_sh, total);
    if (d < 256) boff[d] = e;
  }
  __syncthreads();
  // stage the tile in LDS in digit order ...
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const size_t i = base + (size_t)r * 64 + lane;
    if (i < n) {
      const uint32_t d = pk[r] & 255u;
      const uint32_t lp = boff[d] + wcnt[w][d] + (pk[r] >> 8);
      skeys[lp] = kk[r];
      if constexpr (HAS_VAL) svals[lp] = vv[r];
      if constexpr (SAVE_LP) lp_out[i] = (uint16_t)lp;
    }
  }
  __syncthreads();
  // ... and write it out with consecutive lanes on consecutive addresses of each digit run
  const size_t tile0 = (size_t)tile * TILE;
  const uint32_t cnt = (uint32_t)(n - tile0 < (size_t)TILE ? n - tile0 : (size_t)TILE);
  for (uint32_t i = threadIdx.x; i < cnt; i += NT) {
    const K k = skeys[i];
    const uint32_t d = (uint32_t)(k >> shift) & mask;
    const uint32_t p = goff[d] + (i - boff[d]);
    kout[p] = k;
    if constexpr (HAS_VAL) vout[p] = svals[i];
  }
}

// Undo one recorded pass of the keys-only (replay) sort for per-element results: (vin, fin) are in
// the pass's output order, (vout, fout) receive them in its input order.  One block per forward tile
// (512 threads x 16 items); its LDS staging of 8192 results (72 KiB) allows two blocks per CU.
constexpr int kUThreads = kRThreads;
constexpr int kUItems = kRTile / kUThreads;

// VEC: vout is 16-byte and fout 4-byte aligned.  The write phase then gives each lane 4 consecutive
// outputs: one 8-byte load of their 4 staging indices, two 16-byte value stores and one 4-byte flag
// store, in place of 12 narrow loads and stores (the kernel is bound by load/store instructions, not
// bytes; the 1-byte flag accesses cost as much as the 8-byte values).
template <bool VEC>
__global__ __launch_bounds__(kUThreads) void radix_unscatter_kernel(const uint64_t* __restrict__ vin,
                                                                    const uint8_t* __restrict__ fin,
                                                                    uint64_t* __restrict__ vout,
                                                                    uint8_t* __restrict__ fout,
                                                                    const uint16_t* __restrict__ lp, size_t n,
                                                                    const uint32_t* __restrict__ offsets, uint32_t nb,
                                                                    const uint32_t* __restrict__ dn) {
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (dn) n = *dn;
  if (tile >= nb || (size_t)tile * kRTile >= n) return;  // block-uniform
  __shared__ uint32_t goff[256];
  __shared__ uint32_t boff[257];
  __shared__ uint32_t scan_sh[kUThreads / 64 + 1];
  __shared__ uint64_t sv[kRTile];
  __shared__ uint8_t sf[kRTile];
  {
    // offsets hold, per digit, the exclusive scan of its counts along the tiles, then the 256 digit
    // totals: this tile's count of digit d = the next tile's offset (or the total) - this one
    const uint32_t d = threadIdx.x;
    uint32_t c = 0, g = 0, tot = 0;
    if (d < 256) {
      const size_t idx = (size_t)d * nb + tile;
      tot = offsets[(size_t)256 * nb + d];
      g = offsets[idx];
      // loaded unconditionally so the three loads go out together (idx + 1 <= 256 nb: in bounds)
      const uint32_t nxt = offsets[idx + 1];
      c = (tile + 1 < nb ? nxt : tot) - g;
    }
    uint32_t all;
    const uint32_t dbase = block_exclusive_scan<uint32_t, kUThreads>(tot, scan_sh, all);
    if (d < 256) goff[d] = dbase + g;
    uint32_t total;
    const uint32_t e = block_exclusive_scan<uint32_t, kUThreads>(c, scan_sh, total);
    if (d < 256) boff[d] = e;
    if (d == 0) boff[256] = total;
  }
  __syncthreads();
  const size_t tile0 = (size_t)tile * kRTile;
  const uint32_t cnt = boff[256];
  // gather the tile's digit runs in digit order (consecutive i of one digit read consecutive
  // addresses); the digit of staging index i is the last d with boff[d] <= i.  Every load of the
  // tile is issued before the first one is used, so their latencies overlap.
  uint32_t src[kUItems];
  uint16_t lpv[VEC ? 1 : kUItems];
  uint2 lpw[VEC ? kUItems / 4 : 1];  // VEC: staging indices of outputs 4 j0 .. 4 j0 + 3
  if constexpr (VEC) {
#pragma unroll
    for (int r = 0; r < kUItems / 4; ++r) {
      const uint32_t j0 = (threadIdx.x + r * kUThreads) * 4;
      lpw[r] = j0 + 3 < cnt ? *reinterpret_cast<const uint2*>(lp + tile0 + j0) : make_uint2(0u, 0u);
    }
  }
#pragma unroll
  for (int r = 0; r < kUItems; ++r) {
    const uint32_t i = threadIdx.x + r * kUThreads;
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 128; step; step >>= 1)
      if (boff[lo + step] <= i) lo += step;
    src[r] = goff[lo] + (i - boff[lo]);
    if constexpr (!VEC) lpv[r] = i < cnt ? lp[tile0 + i] : (uint16_t)0;
  }
  uint64_t v[kUItems];
  uint8_t f[kUItems];
#pragma unroll
  for (int r = 0; r < kUItems; ++r) {
    const uint32_t i = threadIdx.x + r * kUThreads;
    v[r] = i < cnt ? vin[src[r]] : 0ull;
    f[r] = i < cnt ? fin[src[r]] : (uint8_t)0;
  }
#pragma unroll
  for (int r = 0; r < kUItems; ++r) {
    const uint32_t i = threadIdx.x + r * kUThreads;
    if (i < cnt) {
      sv[i] = v[r];
      sf[i] = f[r];
    }
  }
  __syncthreads();
  if constexpr (VEC) {
#pragma unroll
    for (int r = 0; r < kUItems / 4; ++r) {
      const uint32_t j0 = (threadIdx.x + r * kUThreads) * 4;
      if (j0 + 3 < cnt) {
        const uint32_t l0 = lpw[r].x & 0xFFFFu, l1 = lpw[r].x >> 16, l2 = lpw[r].y & 0xFFFFu, l3 = lpw[r].y >> 16;
        *reinterpret_cast<ulonglong2*>(vout + tile0 + j0) = make_ulonglong2(sv[l0], sv[l1]);
        *reinterpret_cast<ulonglong2*>(vout + tile0 + j0 + 2) = make_ulonglong2(sv[l2], sv[l3]);
        *reinterpret_cast<uint32_t*>(fout + tile0 + j0) =
            (uint32_t)sf[l0] | ((uint32_t)sf[l1] << 8) | ((uint32_t)sf[l2] << 16) | ((uint32_t)sf[l3] << 24);
      } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
          const uint32_t j = j0 + k;
          if (j < cnt) {
            const uint32_t l = lp[tile0 + j];
            vout[tile0 + j] = sv[l];
            fout[tile0 + j] = sf[l];
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kUItems; ++r) {
      const uint32_t j = threadIdx.x + r * kUThreads;
      if (j < cnt) {
        vout[tile0 + j] = sv[lpv[r]];
        fout[tile0 + j] = sf[lpv[r]];
      }
    }
  }
}

template <typename V>
__global__ void copy_with_pos_kernel(const uint64_t* __restrict__ kin, const V* __restrict__ vin,
                                     uint64_t* __restrict__ kout, V* __restrict__ vout, size_t n, int gen_pos) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    kout[i] = kin[i];
    vout[i] = gen_pos ? (V)i : vin[i];
  }
}


// Per digit (one block each): exclusive scan of counts[d * nb .. d * nb + nb) in place (the digit's
// offsets inside its run, tile by tile); the digit's total goes to counts[256 * nb + d].  One launch
// in place of a three-kernel scan over all 256 * nb counters: the block-level scan of the 256
// totals that the digit starts need is done by every scatter / un-scatter block.
constexpr int kDPer = 16;
__global__ __launch_bounds__(kThreads) void digit_scan_kernel(uint32_t* __restrict__ counts, uint32_t nb) {
  __shared__ uint32_t wsum[kWaves];
  uint32_t* row = counts + (size_t)blockIdx.x * nb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kThreads * kDPer) {  // uniform across the block
    // wave w scans its segment of 64 * kDPer counters in order, 64 at a time (coalesced)
    const uint32_t wb = base + (uint32_t)w * 64 * kDPer;
    uint32_t v[kDPer], ex[kDPer];
#pragma unroll
    for (int i = 0; i < kDPer; ++i) {
      const uint32_t idx = wb + i * 64 + lane;
      v[i] = idx < nb ? row[idx] : 0u;
    }
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < kDPer; ++i) {
      const uint32_t inc = wave_inclusive_scan(v[i]);
      ex[i] = run + inc - v[i];
      run += (uint32_t)__shfl((int)inc, 63);
    }
    if (lane == 0) wsum[w] = run;
    __syncthreads();
    uint32_t before = carry, tot = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t x = wsum[ww];
      before += ww < w ? x : 0u;
      tot += x;
    }
#pragma unroll
    for (int i = 0; i < kDPer; ++i) {
      const uint32_t idx = wb + i * 64 + lane;
      if (idx < nb) row[idx] = before + ex[i];
    }
    carry += tot;
    __syncthreads();  // wsum is rewritten by the next round
  }
  if (threadIdx.x == 0) counts[(size_t)256 * nb + blockIdx.x] = carry;
}

// One digit pass from (sk, sv) into (dk, dv).  counts (256 * (nb + 1) u32) receive the per-digit
// scanned (digit, tile) offsets and the digit totals.
template <typename V, bool HAS_VAL, bool SAVE_LP>
hipError_t radix_pass(const uint64_t* sk, const V* sv, uint64_t* dk, V* dv, uint16_t* lp, size_t n, const uint32_t* dn,
                      int shift, int bits, bool gen, uint32_t* counts, hipStream_t s) {
  constexpr int NT = SAVE_LP ? kRThreads : kThreads;  // the replay sort takes 8192-key tiles
  constexpr int TILE = NT * kItems;
  const size_t nb = (n + TILE - 1) / TILE;
  const uint32_t mask = (1u << bits) - 1u;
  {
    ProfScope prof("psa_hist", s);
    if ((reinterpret_cast<uintptr_t>(sk) & 15) == 0)
      hipLaunchKernelGGL((radix_hist_kernel<uint64_t, true, NT>), dim3(grid_for(nb)), dim3(NT), 0, s, sk, n, shift,
                         mask, counts, (uint32_t)nb, dn);
    else
      hipLaunchKernelGGL((radix_hist_kernel<uint64_t, false, NT>), dim3(grid_for(nb)), dim3(NT), 0, s, sk, n, shift,
                         mask, counts, (uint32_t)nb, dn);
  }
  {
    ProfScope prof("scan", s);
    hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(kThreads), 0, s, counts, (uint32_t)nb);
  }
  {
    ProfScope prof("psa_scatter", s);
    hipLaunchKernelGGL((radix_scatter_kernel<uint64_t, V, HAS_VAL, SAVE_LP, NT>), dim3(grid_for(nb)), dim3(NT), 0, s,
                       sk, sv, dk, dv, lp, n, shift, mask, (const uint32_t*)counts, (uint32_t)nb, gen ? 1 : 0, dn);
  }
  return hipGetLastError();
}

// Sized for the smaller tile (4096 keys): the replay sort's 8192-key tiles need half of it.
size_t counts_bytes(size_t n) { return align_up(256 * ((n + kTile - 1) / kTile + 1) * sizeof(uint32_t), 256); }

template <typename V>
hipError_t radix_sort_impl(const uint64_t* keys_in, const V* vals_in, size_t n, int begin_bit, int end_bit,
                           uint64_t* k0, V* v0, uint64_t* k1, V* v1, void* tmp, hipStream_t s, bool gen_pos,
                           int* which, const uint32_t* dn) {
  *which = 0;
  if (n == 0) return hipSuccess;
  if (begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) return hipErrorInvalidValue;
  if (begin_bit == end_bit) {
    hipLaunchKernelGGL((copy_with_pos_kernel<V>), dim3(1024), dim3(256), 0, s, keys_in, vals_in, k0, v0, n,
                       gen_pos ? 1 : 0);
    return hipGetLastError();
  }
  uint32_t* counts = reinterpret_cast<uint32_t*>(tmp);
  const uint64_t* sk = keys_in;
  const V* sv = vals_in;
  int out = 0;
  bool gen = gen_pos;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    uint64_t* dk = out == 0 ? k0 : k1;
    V* dv = out == 0 ? v0 : v1;
    hipError_t e = radix_pass<V, true, false>(sk, sv, dk, dv, nullptr, n, dn, shift, bits, gen, counts, s);
    if (e != hipSuccess) return e;
    sk = dk;
    sv = dv;
    gen = false;
    *which = out;
    out ^= 1;
  }
  return hipGetLastError();
}

int n_passes(int bits) { return (bits + 7) / 8; }

}  // namespace

size_t radix_tmp_bytes(size_t n) { return counts_bytes(n) + 256; }

hipError_t radix_sort_pos(const uint64_t* keys_in, size_t n, int begin_bit, int end_bit, uint64_t* k0, uint32_t* v0,
                          uint64_t* k1, uint32_t* v1, void* tmp, hipStream_t s, int* which, const uint32_t* dn) {
  return radix_sort_impl<uint32_t>(keys_in, nullptr, n, begin_bit, end_bit, k0, v0, k1, v1, tmp, s, true, which, dn);
}

hipError_t radix_sort_u64pay(const uint64_t* keys_in, const uint64_t* vals_in, size_t n, int begin_bit, int end_bit,
                             uint64_t* k0, uint64_t* v0, uint64_t* k1, uint64_t* v1, void* tmp, hipStream_t s,
                             int* which) {
  return radix_sort_impl<uint64_t>(keys_in, vals_in, n, begin_bit, end_bit, k0, v0, k1, v1, tmp, s, false, which,
                                   nullptr);
}

// Replay record of pass p: counts at replay + p * stride, staging indices right after them.
size_t replay_bytes(size_t n, int bits) {
  return (size_t)n_passes(bits) * (counts_bytes(n) + align_up(n * sizeof(uint16_t), 256));
}

hipError_t radix_sort_keys_replay(const uint64_t* keys_in, size_t n, int begin_bit, int end_bit, uint64_t* k0,
                                  uint64_t* k1, void* replay, hipStream_t s, int* which, const uint32_t* dn) {
  *which = 0;
  if (n == 0) return hipSuccess;
  if (begin_bit < 0 || end_bit > 64 || begin_bit >= end_bit) return hipErrorInvalidValue;
  const size_t stride = counts_bytes(n) + align_up(n * sizeof(uint16_t), 256);
  const uint64_t* sk = keys_in;
  int out = 0, p = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++p) {
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    char* rec = reinterpret_cast<char*>(replay) + (size_t)p * stride;
    uint64_t* dk = out == 0 ? k0 : k1;
    hipError_t e = radix_pass<uint32_t, false, true>(sk, nullptr, dk, nullptr,
                                                     reinterpret_cast<uint16_t*>(rec + counts_bytes(n)), n, dn, shift,
                                                     bits, false, reinterpret_cast<uint32_t*>(rec), s);
    if (e != hipSuccess) return e;
    sk = dk;
    *which = out;
    out ^= 1;
  }
  return hipGetLastError();
}

hipError_t unsort_results(uint64_t* vals, uint8_t* found, size_t n, int bits, const void* replay, uint64_t* v_tmp,
                          uint8_t* f_tmp, uint64_t* out_vals, uint8_t* out_found, hipStream_t s,
                          const uint32_t* dn) {
  if (n == 0) return hipSuccess;
  const int P = n_passes(bits);
  const size_t nb = (n + kRTile - 1) / kRTile;
  const size_t stride = counts_bytes(n) + align_up(n * sizeof(uint16_t), 256);
  const uint64_t* sv = vals;
  const uint8_t* sf = found;
  for (int p = P - 1; p >= 0; --p) {
    const char* rec = reinterpret_cast<const char*>(replay) + (size_t)p * stride;
    uint64_t* dv;
    uint8_t* df;
    if (p == 0) {
      dv = out_vals, df = out_found;
    } else if (sv == vals) {
      dv = v_tmp, df = f_tmp;
    } else {
      dv = vals, df = found;
    }
    ProfScope prof("unsort", s);
    const uint16_t* lp = reinterpret_cast<const uint16_t*>(rec + counts_bytes(n));
    const uint32_t* offs = reinterpret_cast<const uint32_t*>(rec);
    if ((reinterpret_cast<uintptr_t>(dv) & 15) == 0 && (reinterpret_cast<uintptr_t>(df) & 3) == 0)
      hipLaunchKernelGGL(radix_unscatter_kernel<true>, dim3(grid_for(nb)), dim3(kUThreads), 0, s, sv, sf, dv, df, lp,
                         n, offs, (uint32_t)nb, dn);
    else
      hipLaunchKernelGGL(radix_unscatter_kernel<false>, dim3(grid_for(nb)), dim3(kUThreads), 0, s, sv, sf, dv, df, lp,
                         n, offs, (uint32_t)nb, dn);
    sv = dv;
    sf = df;
  }
  return hipGetLastError();
}

}  // namespace hm
