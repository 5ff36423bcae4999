// Served-load benchmark driver: libhm_servebench.so exports hm_serve_bench_main (called in-process,
// e.g. through ctypes, so a process that already uses the GPU never has to start another program),
// and the executable hm_serve_bench wraps it.  Both link against libharmonia_amd.
//
// Builds a tree of 2^k unique keys (value = splitmix64_mix(key)), then plays an open-loop Poisson
// stream of requests against the serving loop: 85% point-lookup requests (batch size log-uniform in
// [1, 65536]; 90% hits), 10% range requests (one range, length uniform in [16, 256]), 5% insert
// batches (1..4096 new keys, log-uniform).  Every result is checked: hits are found with
// value = splitmix64_mix(key), misses (keys never inserted) are absent, and each range returns exactly
// the first `length` keys >= its start of the key set at the epoch it reports (the initial keys, read
// back from the built tree's leaves, plus the inserts published up to that epoch), with their values.
// Prints one JSON object per run.
//
// Usage: hm_serve_bench --keys-log2 27 --requests 20000 --rate 5000 [--plain] [--closed]
//                       [--inflight 4] [--budget-mb 4096] [--seed 1] [--leaf-fill 5]
//   --plain   one request per launch, one worker (the baseline path)
//   --closed  submit back to back (throughput probe) instead of Poisson arrivals
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/harmonia.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull, kM1 = 0xBF58476D1CE4E5B9ull, kM2 = 0x94D049BB133111EBull;
constexpr uint64_t kMask63 = (1ull << 63) - 1;

uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * kM1;
  z = (z ^ (z >> 27)) * kM2;
  return z ^ (z >> 31);
}
uint64_t value_of(uint64_t key) { return mix(key + kGamma); }

// A seeded bijection of [0, 2^63): xor, multiply by odd constants, xor-shift (each invertible).
uint64_t perm63(uint64_t x, uint64_t salt) {
  x = (x ^ salt) & kMask63;
  x = (x * 0x9E3779B97F4A7C15ull) & kMask63;
  x ^= x >> 31;
  x = (x * kM1) & kMask63;
  x ^= x >> 29;
  x = (x * kM2) & kMask63;
  x ^= x >> 32;
  return x;
}

struct Req {
  int type = 0;  // 0 lookup, 1 range, 2 insert
  std::vector<uint64_t> keys, vals;
  std::vector<uint32_t> lens;
  std::vector<uint8_t> hit;  // lookups: expected found
  std::vector<uint64_t> out_v, out_k;
  std::vector<uint8_t> out_f;
  std::vector<uint32_t> out_c;
  double arrival_s = 0;
  uint64_t ticket = 0;
};

double arg_d(int argc, char** argv, const char* name, double def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], name)) return std::atof(argv[i + 1]);
  return def;
}
bool arg_flag(int argc, char** argv, const char* name) {
  for (int i = 1; i < argc; ++i)
    if (!std::strcmp(argv[i], name)) return true;
  return false;
}

#define CHECK(x)                                                                 \
  do {                                                                           \
    int rc_ = (x);                                                               \
    if (rc_) {                                                                   \
      char b_[512];                                                              \
      hm_last_error(b_, sizeof b_);                                              \
      std::fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, b_);                 \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

}  // namespace

extern "C" int hm_serve_bench_main(int argc, char** argv, char* out_json, uint64_t out_len) {
  const int klog = (int)arg_d(argc, argv, "--keys-log2", 27);
  const uint64_t n_req = (uint64_t)arg_d(argc, argv, "--requests", 20000);
  const double rate = arg_d(argc, argv, "--rate", 5000);
  const bool plain = arg_flag(argc, argv, "--plain");
  const bool closed = arg_flag(argc, argv, "--closed");
  const uint32_t inflight = (uint32_t)arg_d(argc, argv, "--inflight", 4);
  const uint64_t budget = (uint64_t)(arg_d(argc, argv, "--budget-mb", 4096) * (1 << 20));
  const uint64_t seed = (uint64_t)arg_d(argc, argv, "--seed", 1);
  const uint64_t n = 1ull << klog;
  const uint64_t salt = mix(seed ^ 0x4841524D4F4E4941ull) & kMask63;
  const uint64_t base = mix(seed + 1) >> 2;
  // input space: [0, n) stored keys, [n, n + M) never-inserted misses, [n + M, ...) insert keys
  const uint64_t M = 1ull << 40;
  std::fprintf(stderr, "building %llu keys\n", (unsigned long long)n);
  std::vector<uint64_t> keys(n), vals(n);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = perm63(base + i, salt);
    if (k == kMask63) k = perm63(base + n + M + (1ull << 41), salt);  // 2^63-1 is never stored
    keys[i] = k;
    vals[i] = value_of(k);
  }
  hm_params p;
  CHECK(hm_params_init(&p));
  // 5 of 8 slots per leaf: slack for the insert stream, so few inserts overflow a leaf (a re-layout of
  // the 2^27-key tree takes ~4 ms on the update path; reads keep going meanwhile)
  p.leaf_fill = (uint32_t)arg_d(argc, argv, "--leaf-fill", 5);
  hm_tree* t = nullptr;
  auto tb = std::chrono::steady_clock::now();
  CHECK(hm_tree_build(keys.data(), vals.data(), n, &p, &t));
  const double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
  std::fprintf(stderr, "built in %.2f s\n", build_s);
  // the initial key set in order, for the range checks: the leaf chain of the built tree, whose key set
  // is checked against `keys` (same count, every key present by the lookups below, none extra)
  std::vector<uint64_t> sorted;
  {
    hm_stats st0;
    CHECK(hm_tree_stats(t, &st0));
    std::vector<uint64_t> lv(st0.n_leaves * 16);
    CHECK(hm_tree_export(t, nullptr, nullptr, lv.data()));
    sorted.reserve(n);
    for (uint64_t i = 0; i < st0.n_leaves; ++i)
      for (int s2 = 0; s2 < 8 && lv[i * 16 + s2] != ~0ull; ++s2) sorted.push_back(lv[i * 16 + s2]);
    if (sorted.size() != n || !std::is_sorted(sorted.begin(), sorted.end())) {
      std::fprintf(stderr, "leaf chain of the built tree is not the key set in order\n");
      return 1;
    }
  }

  // request stream
  std::mt19937_64 g(seed * 7919 + 17);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::exponential_distribution<double> E(rate);
  std::vector<Req> reqs(n_req);
  uint64_t miss_next = 0, ins_next = 0;
  double tnow = 0;
  for (auto& r : reqs) {
    tnow += E(g);
    r.arrival_s = tnow;
    const double c = U(g);
    if (c < 0.85) {
      r.type = 0;
      const uint32_t m = (uint32_t)std::llround(std::exp(U(g) * std::log(65536.0)));
      r.keys.resize(m);
      r.hit.resize(m);
      for (uint32_t j = 0; j < m; ++j) {
        if (U(g) < 0.9) {
          r.keys[j] = keys[g() % n];
          r.hit[j] = 1;
        } else {
          r.keys[j] = perm63(base + n + (miss_next++ % M), salt);
          r.hit[j] = 0;
        }
      }
      r.out_v.resize(m);
      r.out_f.resize(m);
    } else if (c < 0.95) {
      r.type = 1;
      r.keys = {g() & kMask63};
      r.lens = {16u + (uint32_t)(g() % 241)};
      r.out_k.resize(r.lens[0]);
      r.out_v.resize(r.lens[0]);
      r.out_c.resize(1);
    } else {
      r.type = 2;
      const uint32_t m = (uint32_t)std::llround(std::exp(U(g) * std::log(4096.0)));
      r.keys.resize(m);
      r.vals.resize(m);
      for (uint32_t j = 0; j < m; ++j) {
        r.keys[j] = perm63(base + n + M + 1 + ins_next++, salt);
        r.vals[j] = value_of(r.keys[j]);
      }
    }
  }

  hm_server_params sp{budget, 1u << 20, plain ? 1u : inflight, plain ? 0u : 1u, 0};
  hm_server* sv = nullptr;
  CHECK(hm_server_create(t, &sp, &sv));
  // warm-up outside the measurement: every request kind once (first launches load the kernels)
  const uint64_t warm_key = perm63(base + n + M + (1ull << 41) + 1, salt);
  {
    std::vector<uint64_t> wv(4096), wk(64), wv2(64);
    std::vector<uint8_t> wf(4096);
    uint32_t wlen = 64, wc = 0;
    uint64_t tk;
    CHECK(hm_server_lookup(sv, keys.data(), 4096, wv.data(), wf.data(), &tk));
    CHECK(hm_server_wait(sv, tk, nullptr, nullptr));
    CHECK(hm_server_range(sv, keys.data(), &wlen, 1, wk.data(), wv2.data(), &wc, 64, &tk));
    CHECK(hm_server_wait(sv, tk, nullptr, nullptr));
    const uint64_t uk = warm_key, uv = value_of(uk);
    CHECK(hm_server_update(sv, &uk, &uv, 1, &tk));
    CHECK(hm_server_wait(sv, tk, nullptr, nullptr));
    CHECK(hm_server_reset_stats(sv));
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (auto& r : reqs) {
    if (!closed) {
      const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                std::chrono::duration<double>(r.arrival_s));
      while (std::chrono::steady_clock::now() < due) {
      }
    }
    if (r.type == 0)
      CHECK(hm_server_lookup(sv, r.keys.data(), (uint32_t)r.keys.size(), r.out_v.data(), r.out_f.data(), &r.ticket));
    else if (r.type == 1)
      CHECK(hm_server_range(sv, r.keys.data(), r.lens.data(), 1, r.out_k.data(), r.out_v.data(), r.out_c.data(),
                            r.lens[0], &r.ticket));
    else
      CHECK(hm_server_update(sv, r.keys.data(), r.vals.data(), (uint32_t)r.keys.size(), &r.ticket));
  }
  const double submit_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  CHECK(hm_server_drain(sv));
  hm_server_stats st;
  CHECK(hm_server_get_stats(sv, &st));
  // verify every result; first the epoch of every request, and the inserted keys by key with the
  // epoch that published them
  uint64_t bad = 0;
  std::vector<uint64_t> epoch_of(reqs.size());
  std::vector<int> status_of(reqs.size());
  std::vector<std::pair<uint64_t, uint64_t>> ins{{warm_key, 0}};  // (key, publishing epoch); the warm-up
                                                                 // insert precedes every request
  for (size_t i = 0; i < reqs.size(); ++i) {
    CHECK(hm_server_wait(sv, reqs[i].ticket, &status_of[i], &epoch_of[i]));
    if (reqs[i].type == 2 && !status_of[i])
      for (uint64_t k : reqs[i].keys) ins.emplace_back(k, epoch_of[i]);
  }
  std::sort(ins.begin(), ins.end());
  for (size_t ri = 0; ri < reqs.size(); ++ri) {
    auto& r = reqs[ri];
    if (status_of[ri]) {
      ++bad;
      continue;
    }
    if (r.type == 0) {
      for (size_t j = 0; j < r.keys.size(); ++j)
        if (r.out_f[j] != r.hit[j] || (r.hit[j] && r.out_v[j] != value_of(r.keys[j]))) {
          ++bad;
          break;
        }
    } else if (r.type == 1) {
      // the first lens[0] keys >= start of the initial keys and the inserts visible at this epoch
      const uint64_t start = r.keys[0], ep = epoch_of[ri];
      auto a = std::lower_bound(sorted.begin(), sorted.end(), start);
      auto b = std::lower_bound(ins.begin(), ins.end(), std::make_pair(start, (uint64_t)0));
      uint32_t c = 0;
      bool ok = true;
      while (c < r.lens[0] && ok) {
        while (b != ins.end() && b->second > ep) ++b;  // published after this read's epoch
        const bool from_init = a != sorted.end() && (b == ins.end() || *a < b->first);
        if (!from_init && b == ins.end()) break;  // end of the key set
        const uint64_t k = from_init ? *a++ : (b++)->first;
        ok = c < r.out_c[0] && r.out_k[c] == k && r.out_v[c] == value_of(k);  // inserts carry value_of too
        ++c;
      }
      if (!ok || r.out_c[0] != c) ++bad;
    }
  }
  CHECK(hm_server_destroy(sv));
  hm_stats ts;
  CHECK(hm_tree_stats(t, &ts));
  CHECK(hm_tree_destroy(t));
  std::snprintf(out_json, out_len,
      "{\"mode\": \"%s\", \"arrivals\": \"%s\", \"rate_req_s\": %.1f, \"requests\": %llu, \"queries\": %llu, "
      "\"span_s\": %.4f, \"submit_s\": %.4f, \"queries_per_s\": %.1f, \"requests_per_s\": %.1f, \"p50_us\": %.1f, "
      "\"p99_us\": %.1f, \"mean_us\": %.1f, \"max_us\": %.1f, \"launches\": %llu, \"epochs\": %llu, "
      "\"lookups\": %llu, \"ranges\": %llu, \"updates\": %llu, \"bad_results\": %llu, \"keys\": %llu, "
      "\"inflight\": %u, \"build_s\": %.2f, \"leaf_fill\": %u, \"relayouts\": %llu, "
      "\"max_read_batch_us\": %.1f, \"max_update_us\": %.1f, \"max_prepare_us\": %.1f, "
      "\"max_commit_wait_us\": %.1f, \"p99_lookup_us\": %.1f, \"p99_range_us\": %.1f, "
      "\"p99_update_us\": %.1f}\n",
      plain ? "plain" : "served", closed ? "closed" : "poisson", rate, (unsigned long long)n_req,
      (unsigned long long)st.queries, st.span_s, submit_s, st.queries / st.span_s, st.completed / st.span_s,
      st.p50_us, st.p99_us, st.mean_us, st.max_us, (unsigned long long)st.launches, (unsigned long long)st.epochs,
      (unsigned long long)st.lookups, (unsigned long long)st.ranges, (unsigned long long)st.updates,
      (unsigned long long)bad, (unsigned long long)n, sp.max_inflight, build_s, p.leaf_fill,
      (unsigned long long)ts.relayouts, st.max_read_batch_us, st.max_update_us, st.max_prepare_us,
      st.max_commit_wait_us, st.p99_lookup_us, st.p99_range_us, st.p99_update_us);
  return bad ? 3 : 0;
}

#ifdef HM_SERVE_BENCH_EXE
int main(int argc, char** argv) {
  // one hardware queue per worker stream plus the tree's own and the null stream: a read whose
  // stream shares a queue with an update's re-layout would wait behind it (read by the HIP runtime
  // when it initializes, so before the first HIP call; an explicit setting wins)
  setenv("GPU_MAX_HW_QUEUES", "8", 0);
  char out[4096];
  out[0] = 0;
  const int rc = hm_serve_bench_main(argc, argv, out, sizeof out);
  if (out[0]) std::printf("%s\n", out);
  return rc;
}
#endif
