// Thin C entry points over the reference's Harmonia tree (BPlusTree), compiled from the
// reference sources in place (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY.
//   BPlusTree::insert           H/src/bpt.cpp:357-409
//   BPlusTree::sort             H/src/traversal.cpp:35-129 (BFS relayout + prefix array)
//   BPlusTree::search           H/src/bpt.cpp:673-680 (PPI tree: prefix descent, record-list match;
//                               not found = -1)
//   BPlusTree::range_search     H/src/bpt.cpp:695-769 (keys in [start, end])
#include <cstdint>
#include <vector>
#include "bpt.h"

extern "C" {

void* ref_h_build(const int64_t* keys, const int64_t* vals, uint64_t n) {
  BPlusTree* t = new BPlusTree();
  for (uint64_t i = 0; i < n; ++i) t->insert((key_t)keys[i], (value_t)vals[i]);
  BPlusTree* sorted = nullptr;
  t->sort(sorted);
  return sorted;
}

int ref_h_height(void* t) { return ((BPlusTree*)t)->getHeight(); }

void ref_h_search(void* t, const int64_t* q, uint64_t m, int64_t* out_vals) {
  BPlusTree* tree = (BPlusTree*)t;
  for (uint64_t i = 0; i < m; ++i) out_vals[i] = (int64_t)tree->search((key_t)q[i]);
}

// search_cpu_test (H/src/main.cpp:772-812): BPlusTree::search under an OpenMP parallel for over
// the queries, `nthreads` threads.
void ref_h_search_omp(void* t, const int64_t* q, uint64_t m, int64_t* out_vals, int nthreads) {
  BPlusTree* tree = (BPlusTree*)t;
#pragma omp parallel for num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)m; ++i) out_vals[i] = (int64_t)tree->search((key_t)q[i]);
}

uint64_t ref_h_range(void* t, int64_t lo, int64_t hi, int64_t* out_keys, int64_t* out_vals, uint64_t cap) {
  std::vector<Record> v;
  ((BPlusTree*)t)->range_search((key_t)lo, (key_t)hi, &v);
  for (size_t i = 0; i < v.size() && i < cap; ++i) {
    out_keys[i] = v[i].r_key;
    out_vals[i] = v[i].val;
  }
  return v.size();
}

}  // extern "C"
