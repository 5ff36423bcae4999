// Thin C entry points over the reference's textbook B+tree (regular_bpt), compiled from the
// reference sources in place (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used to pin the
// oracle and to write golden vectors; the product never links it.
//   insert        H/regular_bpt/regular_bpt.cpp:759-804 (duplicates ignored)
//   find          H/regular_bpt/regular_bpt.cpp:383-393 (upper_bound descent, exact leaf match)
//   find_range    H/regular_bpt/regular_bpt.cpp:320-338 (all keys in [start, end], ascending)
#include <cstdint>
#include <vector>
#include "regular_bpt.h"

extern "C" {

void* ref_rbpt_build(const int64_t* keys, const int64_t* vals, uint64_t n) {
  node* root = nullptr;
  for (uint64_t i = 0; i < n; ++i) root = insert(root, (key_t)keys[i], (val_t)vals[i]);
  return root;
}

void* ref_rbpt_insert(void* root, const int64_t* keys, const int64_t* vals, uint64_t n) {
  node* r = (node*)root;
  for (uint64_t i = 0; i < n; ++i) r = insert(r, (key_t)keys[i], (val_t)vals[i]);
  return r;
}

int ref_rbpt_order(void) { return order; }

int ref_rbpt_height(void* root) { return root ? height((node*)root) : -1; }

void ref_rbpt_find(void* root, const int64_t* q, uint64_t m, int64_t* out_vals, uint8_t* out_found) {
  for (uint64_t i = 0; i < m; ++i) {
    record* r = find((node*)root, (key_t)q[i], false);
    out_found[i] = r ? 1 : 0;
    out_vals[i] = r ? (int64_t)r->value : 0;
  }
}

// search_cpu_multithread (H/regular_bpt/main.cpp:92-116): find() under an OpenMP parallel for
// over the queries, `nthreads` threads.
void ref_rbpt_find_omp(void* root, const int64_t* q, uint64_t m, int64_t* out_vals, uint8_t* out_found,
                       int nthreads) {
  node* r = (node*)root;
#pragma omp parallel for num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    record* rec = find(r, (key_t)q[i], false);
    out_found[i] = rec ? 1 : 0;
    out_vals[i] = rec ? (int64_t)rec->value : 0;
  }
}

// Keys in [lo, hi]; writes at most cap pairs, returns the full count.
uint64_t ref_rbpt_find_range(void* root, int64_t lo, int64_t hi, int64_t* out_keys, int64_t* out_vals,
                             uint64_t cap, uint64_t n_keys_total) {
  std::vector<key_t> ks(n_keys_total + 1);
  std::vector<void*> ps(n_keys_total + 1);
  int c = find_range((node*)root, (key_t)lo, (key_t)hi, false, ks.data(), ps.data());
  for (int i = 0; i < c && (uint64_t)i < cap; ++i) {
    out_keys[i] = ks[i];
    out_vals[i] = ((record*)ps[i])->value;
  }
  return (uint64_t)c;
}

}  // extern "C"
