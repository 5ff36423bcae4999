// Hand-written stable LSD radix sort (8-bit digits) for u64 or u32 keys with a payload.
//
// The PSA step of the reference sorts (key, position) pairs with
// cub::DeviceRadixSort::SortPairs(keys, pos, n, 48, 64) before every search launch
// (H/regular_bpt/prefix-sort.cu:288-293, H/src/full-multi-prefix-sort-8thread.cu:305-309).
// That call is a stable LSD radix sort over key bits [begin_bit, end_bit); for keys in
// [0, 2^63) CUB's signed-key bit twiddle (H/cub/cub/util_type.cuh:1015-1022) does not change the
// order, so the result equals a stable sort of (key, pos) by key >> begin_bit.  The same kernels
// sort the bulk-build input (u64 value payload over every varying bit) and, keyed by the original
// position, return PSA-ordered search results to input order ("un-permute"; the last pass writes
// the caller's value and found arrays directly).
//
// One pass = histogram kernel (per-tile digit counts, digit-major) -> exclusive scan over
// (digit, tile) -> scatter kernel (wave-level digit match via 8 ballots gives each item its
// stable rank inside its wave; per-wave counters are scanned across the block; the tile is then
// staged in LDS in digit order so the global writes of each digit run are coalesced).
#include "hm_block.h"
#include "hm_internal.h"

namespace hm {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;

// Lanes (among `valid`) whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t peers = valid;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

template <typename K>
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const K* __restrict__ keys, size_t n, int shift,
                                                              uint32_t mask, uint32_t* __restrict__ counts,
                                                              uint32_t nb, const uint32_t* __restrict__ dn) {
  if (dn) n = *dn;  // device-side count (graph replays): tiles past n count nothing
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const size_t base = (size_t)blockIdx.x * kTile + (size_t)w * 64 * kItems;
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
    const bool valid = i < n;
    const uint32_t d = valid ? (uint32_t)(kk[r] >> shift) & mask : 0u;
    const uint64_t peers = match_digit(d, __ballot(valid));
    if (valid && (peers & lt) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * nb + blockIdx.x] = hist[threadIdx.x];
}

// MODE 0: write (key, value) to kout/vout.
// MODE 1: final un-permute pass: keys are pos | found << 31, the destination index is the position
//         itself; write out_vals[p] = value and out_found[p] = found.
template <typename K, typename V, int MODE>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(const K* __restrict__ kin, const V* __restrict__ vin,
                                                                 K* __restrict__ kout, V* __restrict__ vout,
                                                                 uint8_t* __restrict__ fout, size_t n, int shift,
                                                                 uint32_t mask, const uint32_t* __restrict__ offsets,
                                                                 uint32_t nb, int gen_pos,
                                                                 const uint32_t* __restrict__ dn) {
  if (dn) n = *dn;
  if ((size_t)blockIdx.x * kTile >= n) return;  // whole tile past the end (block-uniform)
  __shared__ uint32_t wcnt[kWaves][256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t boff[256];
  __shared__ uint32_t scan_sh[kWaves + 1];
  __shared__ K skeys[kTile];
  __shared__ V svals[kTile];
  for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&wcnt[0][0])[i] = 0;
  goff[threadIdx.x] = offsets[(size_t)threadIdx.x * nb + blockIdx.x];
  __syncthreads();

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const size_t base = (size_t)blockIdx.x * kTile + (size_t)w * 64 * kItems;
  K kk[kItems];
  V vv[kItems];
  uint32_t pk[kItems];  // (rank within wave << 8) | digit
  const size_t last = n - 1;  // n > 0 here
  // issue every load of the tile before the first use, so the loads overlap
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const size_t i = base + (size_t)r * 64 + lane;
    const size_t ic = i < last ? i : last;
    kk[r] = kin[ic];
    vv[r] = gen_pos ? (V)i : vin[ic];
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
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t t = wcnt[ww][d];
      wcnt[ww][d] = run;
      run += t;
    }
    uint32_t total;
    boff[d] = block_exclusive_scan<uint32_t, kThreads>(run, scan_sh, total);
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
      svals[lp] = vv[r];
    }
  }
  __syncthreads();
  // ... and write it out with consecutive lanes on consecutive addresses of each digit run
  const size_t tile0 = (size_t)blockIdx.x * kTile;
  const uint32_t cnt = (uint32_t)(n - tile0 < (size_t)kTile ? n - tile0 : (size_t)kTile);
  for (uint32_t i = threadIdx.x; i < cnt; i += kThreads) {
    const K k = skeys[i];
    const uint32_t d = (uint32_t)(k >> shift) & mask;
    const uint32_t p = goff[d] + (i - boff[d]);
    if constexpr (MODE == 0) {
      kout[p] = k;
      vout[p] = svals[i];
    } else {
      vout[p] = svals[i];
      fout[p] = (uint8_t)((uint32_t)k >> 31);
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

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// One digit pass from (sk, sv) into (dk, dv) [MODE 0] or into the result arrays [MODE 1].
template <typename K, typename V, int MODE>
hipError_t radix_pass(const K* sk, const V* sv, K* dk, V* dv, uint8_t* df, size_t n, const uint32_t* dn, int shift, int bits,
                      bool gen, void* tmp, hipStream_t s, const char* hist_stage, const char* scatter_stage) {
  const size_t nb = (n + kTile - 1) / kTile;
  uint32_t* counts = reinterpret_cast<uint32_t*>(tmp);
  void* scan_tmp = reinterpret_cast<char*>(tmp) + align_up(256 * nb * sizeof(uint32_t), 256);
  const uint32_t mask = (1u << bits) - 1u;
  {
    ProfScope prof(hist_stage, s);
    hipLaunchKernelGGL((radix_hist_kernel<K>), dim3((unsigned)nb), dim3(kThreads), 0, s, sk, n, shift, mask, counts,
                       (uint32_t)nb, dn);
  }
  hipError_t e = exclusive_scan_u32(counts, counts, 256 * nb, 0u, scan_tmp, s);
  if (e != hipSuccess) return e;
  {
    ProfScope prof(scatter_stage, s);
    hipLaunchKernelGGL((radix_scatter_kernel<K, V, MODE>), dim3((unsigned)nb), dim3(kThreads), 0, s, sk, sv, dk, dv,
                       df, n, shift, mask, (const uint32_t*)counts, (uint32_t)nb, gen ? 1 : 0, dn);
  }
  return hipGetLastError();
}

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
  const uint64_t* sk = keys_in;
  const V* sv = vals_in;
  int out = 0;
  bool gen = gen_pos;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    uint64_t* dk = out == 0 ? k0 : k1;
    V* dv = out == 0 ? v0 : v1;
    hipError_t e = radix_pass<uint64_t, V, 0>(sk, sv, dk, dv, nullptr, n, dn, shift, bits, gen, tmp, s, "psa_hist",
                                              "psa_scatter");
    if (e != hipSuccess) return e;
    sk = dk;
    sv = dv;
    gen = false;
    *which = out;
    out ^= 1;
  }
  return hipGetLastError();
}

}  // namespace

size_t radix_tmp_bytes(size_t n) {
  const size_t nb = (n + kTile - 1) / kTile + 1;
  return align_up(256 * nb * sizeof(uint32_t), 256) + scan_tmp_bytes(256 * nb) + 256;
}

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

hipError_t unpermute_results(const uint32_t* posf, const uint64_t* vals, size_t n, uint32_t* p0, uint64_t* v0,
                             uint32_t* p1, uint64_t* v1, uint64_t* out_vals, uint8_t* out_found, void* tmp,
                             hipStream_t s, const uint32_t* dn) {
  if (n == 0) return hipSuccess;
  int bits = 1;
  while (bits < 31 && ((size_t)1 << bits) < n) ++bits;  // positions fit in `bits` bits
  const uint32_t* sk = posf;
  const uint64_t* sv = vals;
  int out = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    const int b = bits - shift < 8 ? bits - shift : 8;
    hipError_t e;
    if (shift + b >= bits) {
      e = radix_pass<uint32_t, uint64_t, 1>(sk, sv, nullptr, out_vals, out_found, n, dn, shift, b, false, tmp, s,
                                            "unpermute_hist", "unpermute_scatter");
    } else {
      uint32_t* dk = out == 0 ? p0 : p1;
      uint64_t* dv = out == 0 ? v0 : v1;
      e = radix_pass<uint32_t, uint64_t, 0>(sk, sv, dk, dv, nullptr, n, dn, shift, b, false, tmp, s, "unpermute_hist",
                                            "unpermute_scatter");
      sk = dk;
      sv = dv;
      out ^= 1;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace hm
