// A driver written against the C++ mirror (include/harmonia.hpp) the way the reference's main.cpp
// drives its BPlusTree (H/src/main.cpp:166-414): per-key insert() then build(), batched search,
// single search, range_search over closed intervals, insert_parallel, search again.
//
// usage: hm_mirror_example <n_keys> <seed> <out_file>
// Keys are the splitmix64 stream of `seed` shifted right by one bit (so below 2^63), values
// splitmix64_mix(key); queries are every key plus as many keys of the stream of seed + 1; intervals
// start at queries and span 2^52; the insert batch (an upsert) is the first n/4 outputs of the stream
// of seed + 2 with values key ^ 0x5bd1e995.  Everything the tree returns is written to
// out_file as little-endian u64 arrays, in this order: found and value per query (before the
// insert), the single-search results of the first 1000 queries, per interval its count then its
// (key, value) pairs, found and value per query after the insert.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/harmonia.hpp"

namespace {

uint64_t splitmix_next(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t value_of(uint64_t key) {
  uint64_t s = key;
  return splitmix_next(s);
}

std::vector<uint64_t> stream63(uint64_t seed, uint64_t n) {
  std::vector<uint64_t> v(n);
  for (auto& x : v) x = splitmix_next(seed) >> 1;
  return v;
}

void put(std::FILE* f, uint64_t x) { std::fwrite(&x, 8, 1, f); }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <n_keys> <seed> <out_file>\n", argv[0]);
    return 2;
  }
  const uint64_t n = std::strtoull(argv[1], nullptr, 10), seed = std::strtoull(argv[2], nullptr, 10);
  std::FILE* f = std::fopen(argv[3], "wb");
  if (!f) return 2;
  try {
    harmonia::BPlusTree tree;
    for (uint64_t k : stream63(seed, n)) tree.insert(k, value_of(k));  // staged, as the reference inserts
    tree.build();
    std::vector<uint64_t> q = stream63(seed, n), extra = stream63(seed + 1, n);
    q.insert(q.end(), extra.begin(), extra.end());
    std::vector<uint64_t> vals;
    std::vector<uint8_t> found;
    tree.search(q, vals, found);
    for (size_t i = 0; i < q.size(); ++i) {
      put(f, found[i]);
      put(f, vals[i]);
    }
    for (size_t i = 0; i < 1000 && i < q.size(); ++i) {
      const auto r = tree.search(q[i]);
      put(f, r ? 1 : 0);
      put(f, r ? *r : 0);
    }
    for (size_t i = 0; i < 200 && i < q.size(); ++i) {
      std::vector<harmonia::Record> rec;
      tree.range_search(q[i], q[i] + (1ull << 52), &rec);
      put(f, rec.size());
      for (const auto& r : rec) {
        put(f, r.r_key);
        put(f, r.val);
      }
    }
    std::vector<uint64_t> nk = stream63(seed + 2, n / 4), nv(nk.size());
    for (size_t i = 0; i < nk.size(); ++i) nv[i] = nk[i] ^ 0x5bd1e995ull;
    tree.insert_parallel(nk, nv);
    q.insert(q.end(), nk.begin(), nk.end());
    tree.search(q, vals, found);
    for (size_t i = 0; i < q.size(); ++i) {
      put(f, found[i]);
      put(f, vals[i]);
    }
    std::printf("{\"keys\": %llu, \"height\": %d, \"epoch\": %llu}\n", (unsigned long long)tree.stats().n_keys,
                tree.getHeight(), (unsigned long long)tree.stats().epoch);
  } catch (const harmonia::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    std::fclose(f);
    return 1;
  }
  std::fclose(f);
  return 0;
}
