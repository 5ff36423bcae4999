// Text datasets in the reference's format.
//
// The reference reads one decimal key per line with getline + sscanf(TYPE_D = "%lld")
// (H/src/main.cpp:50-68, H/src/bpt.h:32-35); range files hold a start line and an end line per
// range (H/src/range.cu:348-357); its generators write keys in [1, 2^63 - 1)
// (H/script/generate-data/util.py:30-60).  Here a file is memory-mapped and parsed by several
// threads, each over a newline-aligned chunk: per chunk the lines are counted first, so every
// thread knows where its keys go.  Stricter than sscanf: a line must hold one integer in
// [0, 2^63) (surrounding blanks and a '+' allowed) -- anything else is an error naming the line,
// where sscanf would have left the key uninitialised or wrapped it.  Empty lines are skipped.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "hm_tree.h"

using namespace hm;

namespace {

struct Mapped {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap((void*)p, n);
    if (fd >= 0) close(fd);
  }
};

int map_file(const char* path, Mapped& m) {
  if (!path) return set_err(HM_EINVAL, "path is NULL");
  m.fd = open(path, O_RDONLY);
  if (m.fd < 0) return set_err(HM_EINVAL, "cannot open %s: %s", path, strerror(errno));
  struct stat st;
  if (fstat(m.fd, &st) != 0) return set_err(HM_EINVAL, "cannot stat %s", path);
  m.n = (size_t)st.st_size;
  if (m.n == 0) return HM_OK;
  void* p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
  if (p == MAP_FAILED) {
    m.n = 0;
    return set_err(HM_EINVAL, "cannot map %s", path);
  }
  m.p = (const char*)p;
  return HM_OK;
}

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Parse one line [b, e) into *v; returns 1 for a key, 0 for an empty line, -1 for an error.
int parse_line(const char* b, const char* e, uint64_t* v) {
  while (b < e && blank(*b)) ++b;
  while (e > b && blank(e[-1])) --e;
  if (b == e) return 0;
  if (*b == '+') ++b;
  if (b == e) return -1;
  uint64_t x = 0;
  for (; b < e; ++b) {
    const unsigned d = (unsigned)(*b - '0');
    if (d > 9) return -1;
    if (x > ((1ull << 63) - 1 - d) / 10) return -1;  // would leave [0, 2^63)
    x = x * 10 + d;
  }
  *v = x;
  return 1;
}

// Newline-aligned chunk boundaries of the mapped file.
std::vector<size_t> chunks(const Mapped& m, int parts) {
  std::vector<size_t> c{0};
  for (int i = 1; i < parts; ++i) {
    size_t pos = m.n * (size_t)i / (size_t)parts;
    if (pos <= c.back()) continue;
    const void* nl = memchr(m.p + pos, '\n', m.n - pos);
    pos = nl ? (size_t)((const char*)nl - m.p) + 1 : m.n;
    if (pos > c.back() && pos < m.n) c.push_back(pos);
  }
  c.push_back(m.n);
  return c;
}

int n_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::min<unsigned>(std::max<unsigned>(h, 1), 16);
}

// Parse every non-empty line of the file into out[0 .. cap); *n = number of values.  With
// out == nullptr only counts.  Lines are numbered from 1 in error messages.
int parse_file(const char* path, uint64_t* out, uint64_t cap, uint64_t* n) {
  Mapped m;
  int rc = map_file(path, m);
  if (rc) return rc;
  *n = 0;
  if (m.n == 0) return HM_OK;
  const std::vector<size_t> c = chunks(m, n_threads());
  const size_t parts = c.size() - 1;
  std::vector<uint64_t> lines(parts, 0), vals(parts, 0);
  std::vector<long long> bad(parts, -1);
  // pass 1: per chunk, lines and values
  {
    std::vector<std::thread> th;
    for (size_t k = 0; k < parts; ++k)
      th.emplace_back([&, k] {
        const char* b = m.p + c[k];
        const char* end = m.p + c[k + 1];
        uint64_t ln = 0, nv = 0;
        while (b < end) {
          const char* e = (const char*)memchr(b, '\n', (size_t)(end - b));
          if (!e) e = end;
          uint64_t v;
          const int r = parse_line(b, e, &v);
          if (r < 0 && bad[k] < 0) bad[k] = (long long)ln;
          nv += r > 0 ? 1 : 0;
          ++ln;
          b = e + 1;
        }
        lines[k] = ln;
        vals[k] = nv;
      });
    for (auto& t : th) t.join();
  }
  uint64_t line0 = 0, total = 0;
  for (size_t k = 0; k < parts; ++k) {
    if (bad[k] >= 0)
      return set_err(HM_EINVAL, "%s:%llu: not an integer in [0, 2^63)", path,
                     (unsigned long long)(line0 + (uint64_t)bad[k] + 1));
    line0 += lines[k];
    total += vals[k];
  }
  *n = total;
  if (!out) return HM_OK;
  if (total > cap)
    return set_err(HM_EINVAL, "%s holds %llu keys, buffer holds %llu", path, (unsigned long long)total,
                   (unsigned long long)cap);
  // pass 2: write each chunk's values at its offset
  std::vector<uint64_t> off(parts, 0);
  for (size_t k = 1; k < parts; ++k) off[k] = off[k - 1] + vals[k - 1];
  std::vector<std::thread> th;
  for (size_t k = 0; k < parts; ++k)
    th.emplace_back([&, k] {
      const char* b = m.p + c[k];
      const char* end = m.p + c[k + 1];
      uint64_t* o = out + off[k];
      while (b < end) {
        const char* e = (const char*)memchr(b, '\n', (size_t)(end - b));
        if (!e) e = end;
        uint64_t v;
        if (parse_line(b, e, &v) > 0) *o++ = v;
        b = e + 1;
      }
    });
  for (auto& t : th) t.join();
  return HM_OK;
}

}  // namespace

extern "C" {

int hm_read_keys_text(const char* path, uint64_t* keys, uint64_t cap, uint64_t* n) {
  if (!n) return set_err(HM_EINVAL, "n is NULL");
  if (!keys && cap) return set_err(HM_EINVAL, "keys is NULL");
  return parse_file(path, keys, cap, n);
}

int hm_read_ranges_text(const char* path, uint64_t* starts, uint64_t* ends, uint64_t cap, uint64_t* n) {
  if (!n) return set_err(HM_EINVAL, "n is NULL");
  uint64_t total = 0;
  int rc = parse_file(path, nullptr, 0, &total);
  if (rc) return rc;
  if (total % 2) return set_err(HM_EINVAL, "%s: odd number of lines (start/end pairs expected)", path);
  *n = total / 2;
  if (!starts || !ends) return cap ? set_err(HM_EINVAL, "starts/ends is NULL") : HM_OK;
  if (*n > cap)
    return set_err(HM_EINVAL, "%s holds %llu ranges, buffers hold %llu", path, (unsigned long long)*n,
                   (unsigned long long)cap);
  std::vector<uint64_t> all(total);
  rc = parse_file(path, all.data(), total, &total);
  if (rc) return rc;
  for (uint64_t i = 0; i < *n; ++i) {
    starts[i] = all[2 * i];
    ends[i] = all[2 * i + 1];
  }
  return HM_OK;
}

}  // extern "C"
