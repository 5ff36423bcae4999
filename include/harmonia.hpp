// harmonia.hpp -- header-only C++ mirror of the reference's tree interface over the C ABI.
//
// Method names follow the reference's class BPlusTree (H/src/bpt.h:456-561): insert (464),
// insert_parallel (468), search (476), range_search (481), getHeight (486), getPrefixArray (503),
// getLevelIdx (540).  The reference's host-side drivers (`void F(BPlusTree&, std::ifstream&)`,
// H/src/ppi-bpt.h:4-98) become batched calls on a device-resident tree.  Errors throw
// harmonia::Error carrying hm_last_error() instead of the reference's exit(1).
#ifndef HARMONIA_HPP
#define HARMONIA_HPP

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "harmonia.h"

namespace harmonia {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != HM_OK) {
    char buf[512];
    hm_last_error(buf, sizeof(buf));
    throw Error(rc, std::string("harmonia: ") + buf);
  }
}

inline hm_params default_params() {
  hm_params p;
  check(hm_params_init(&p));
  return p;
}

struct Record {
  uint64_t r_key;
  uint64_t val;
};

class BPlusTree {
 public:
  explicit BPlusTree(const hm_params& p = default_params()) : params_(p) {}
  ~BPlusTree() { reset(); }
  BPlusTree(const BPlusTree&) = delete;
  BPlusTree& operator=(const BPlusTree&) = delete;
  BPlusTree(BPlusTree&& o) noexcept : params_(o.params_), t_(o.t_), keys_(std::move(o.keys_)), vals_(std::move(o.vals_)) {
    o.t_ = nullptr;
  }

  // Reference-style insert: staged until build(), an upsert (visible immediately) afterwards.
  bool insert(uint64_t key, uint64_t val) {
    if (!t_) {
      keys_.push_back(key);
      vals_.push_back(val);
      return true;
    }
    insert_parallel(std::vector<uint64_t>{key}, std::vector<uint64_t>{val});
    return true;
  }

  // Bulk build from the staged inserts (replaces the insert loop + BPlusTree::sort relayout).
  void build() {
    build(keys_.data(), vals_.data(), keys_.size());
    keys_.clear();
    vals_.clear();
  }
  // Keys/vals may be host or device pointers.
  void build(const uint64_t* keys, const uint64_t* vals, uint64_t n, bool keys_sorted = false) {
    reset();
    hm_params p = params_;
    if (keys_sorted) p.flags |= HM_FLAG_KEYS_SORTED;
    check(hm_tree_build(keys, vals, n, &p, &t_));
  }

  // Single lookup: the value, or nothing (the reference returns -1, H/src/bpt.h:272).
  std::optional<uint64_t> search(uint64_t key) {
    uint64_t v = 0;
    uint8_t f = 0;
    hm_search_opts o{0, -1, 0, 0, 0, 0, nullptr, 0, 0};
    check(hm_search_ex(handle(), &key, 1, &v, &f, &o, nullptr));
    if (!f) return std::nullopt;
    return v;
  }

  // Batched lookup in input order (host or device buffers; stream = hipStream_t or nullptr).
  void search(const uint64_t* q, uint64_t n, uint64_t* vals, uint8_t* found, void* stream = nullptr) {
    check(hm_search(handle(), q, n, vals, found, stream));
  }
  void search(const std::vector<uint64_t>& q, std::vector<uint64_t>& vals, std::vector<uint8_t>& found) {
    vals.resize(q.size());
    found.resize(q.size());
    search(q.data(), q.size(), vals.data(), found.data());
  }

  // All records with start_key <= key <= end_key, ascending (BPlusTree::range_search).
  void range_search(uint64_t start_key, uint64_t end_key, std::vector<Record>* v, uint32_t cap = 4096) {
    for (;;) {
      std::vector<uint64_t> k(cap), val(cap);
      uint32_t c = 0;
      check(hm_range_between(handle(), &start_key, &end_key, 1, cap, k.data(), val.data(), &c, nullptr));
      if (c < cap) {
        for (uint32_t i = 0; i < c; ++i) v->push_back(Record{k[i], val[i]});
        return;
      }
      cap *= 4;
    }
  }

  // Batched upsert, applied at once (BPlusTree::insert_parallel).  Returns true if the leaves
  // were re-laid out.
  bool insert_parallel(const std::vector<uint64_t>& keys, const std::vector<uint64_t>& vals) {
    if (keys.size() != vals.size()) throw Error(HM_EINVAL, "harmonia: keys/vals size mismatch");
    check(hm_update(handle(), keys.data(), vals.data(), keys.size(), nullptr));
    int rebuilt = 0;
    check(hm_epoch_publish(handle(), nullptr, &rebuilt));
    return rebuilt != 0;
  }
  // Staged form for serving loops: update() now, publish() between search epochs.
  void update(const uint64_t* keys, const uint64_t* vals, uint64_t n, void* stream = nullptr) {
    check(hm_update(handle(), keys, vals, n, stream));
  }
  bool publish(void* stream = nullptr) {
    int rebuilt = 0;
    check(hm_epoch_publish(handle(), stream, &rebuilt));
    return rebuilt != 0;
  }

  hm_stats stats() const {
    hm_stats s;
    check(hm_tree_stats(handle(), &s));
    return s;
  }
  int getHeight() const { return (int)stats().height; }
  int getLevelIdx(int level) const { return (int)stats().level_start[level]; }
  std::vector<uint32_t> getPrefixArray() const {
    std::vector<uint32_t> p(stats().n_inner);
    check(hm_tree_export(handle(), nullptr, p.data(), nullptr));
    return p;
  }
  hm_tree* handle() const {
    if (!t_) throw Error(HM_EINVAL, "harmonia: tree is not built");
    return t_;
  }
  // Binary snapshot of the published tree / reload (hm_tree_save, hm_tree_load).
  void save(const std::string& path) const { check(hm_tree_save(handle(), path.c_str())); }
  void load(const std::string& path) {
    reset();
    check(hm_tree_load(path.c_str(), &params_, &t_));
  }

 private:
  void reset() {
    if (t_) hm_tree_destroy(t_);
    t_ = nullptr;
  }
  hm_params params_;
  hm_tree* t_ = nullptr;
  std::vector<uint64_t> keys_, vals_;
};

// Keys of a text dataset in the reference's format (one decimal key per line, H/src/main.cpp:50-68).
inline std::vector<uint64_t> read_keys_text(const std::string& path) {
  uint64_t n = 0;
  check(hm_read_keys_text(path.c_str(), nullptr, 0, &n));
  std::vector<uint64_t> keys(n);
  check(hm_read_keys_text(path.c_str(), keys.data(), n, &n));
  keys.resize(n);
  return keys;
}

}  // namespace harmonia

#endif  // HARMONIA_HPP
