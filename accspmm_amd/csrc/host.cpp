// host.cpp — host-side utilities of libaccspmm: Matrix Market ingest, CSR
// construction and deterministic graph generators.  No GPU code here.
#include <fcntl.h>
#include <sched.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/accspmm.h"
#include "guard.hpp"
#include "rng.hpp"

extern "C" void accspmm_host_csr_free(accspmm_host_csr* c) {
  if (!c) return;
  std::free(c->row_ptr);
  std::free(c->col_idx);
  std::free(c->val);
  c->row_ptr = nullptr;
  c->col_idx = nullptr;
  c->val = nullptr;
  c->rows = c->cols = c->nnz = 0;
}

namespace {

// malloc-compatible allocation (released with std::free); blocks of 4 MiB and
// more are 2 MiB aligned and marked for transparent huge pages, which cuts
// the first-touch page faults that dominate filling fresh ingest arrays.
void* host_alloc(size_t bytes) {
  constexpr size_t kHuge = 2u << 20;
  if (bytes < 2 * kHuge) return std::malloc(bytes ? bytes : 1);
  void* p = nullptr;
  if (posix_memalign(&p, kHuge, bytes) != 0) return nullptr;
  madvise(p, bytes, MADV_HUGEPAGE);
  return p;
}

template <class T>
struct HostAllocator {
  using value_type = T;
  HostAllocator() = default;
  template <class U>
  HostAllocator(const HostAllocator<U>&) {}
  T* allocate(size_t n) {
    void* p = host_alloc(n * sizeof(T));
    if (!p) throw std::bad_alloc();
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { std::free(p); }
  template <class U>
  bool operator==(const HostAllocator<U>&) const { return true; }
  template <class U>
  bool operator!=(const HostAllocator<U>&) const { return false; }
};

template <class T>
using HostVec = std::vector<T, HostAllocator<T>>;

bool alloc_csr(accspmm_host_csr* out, uint32_t rows, uint32_t cols,
               uint64_t nnz) {
  out->rows = rows;
  out->cols = cols;
  out->nnz = (uint32_t)nnz;
  out->row_ptr = (uint32_t*)host_alloc(sizeof(uint32_t) * (rows + 1ull));
  out->col_idx = (uint32_t*)host_alloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  out->val = (float*)host_alloc(sizeof(float) * (nnz ? nnz : 1));
  if (!out->row_ptr || !out->col_idx || !out->val) {
    accspmm_host_csr_free(out);
    return false;
  }
  return true;
}

// Worker threads for host ingest: ACCSPMM_HOST_THREADS when set, else the
// CPUs this process may run on, at most 16.
int host_threads() {
  if (const char* e = std::getenv("ACCSPMM_HOST_THREADS")) {
    const int v = std::atoi(e);
    if (v >= 1) return std::min(v, 256);
  }
  int n = 0;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(n, 16));
}

// Runs f(i) for i in [0, n) on n threads (the caller's thread takes i = 0).
template <class F>
void run_parallel(int n, F&& f) {
  std::vector<std::thread> ts;
  ts.reserve(n > 1 ? n - 1 : 0);
  for (int i = 1; i < n; ++i) ts.emplace_back([&f, i] { f(i); });
  f(0);
  for (auto& t : ts) t.join();
}

// Row blocks of about equal entry counts: block i is rows [b[i], b[i+1]).
std::vector<uint32_t> row_blocks(const uint32_t* row_ptr, uint32_t rows, int parts) {
  std::vector<uint32_t> b(parts + 1, rows);
  b[0] = 0;
  const uint64_t nnz = row_ptr[rows];
  for (int i = 1; i < parts; ++i) {
    const uint64_t target = nnz * (uint64_t)i / (uint64_t)parts;
    b[i] = (uint32_t)(std::lower_bound(row_ptr, row_ptr + rows + 1, (uint32_t)target) - row_ptr);
    b[i] = std::min(std::max(b[i], b[i - 1]), rows);
  }
  return b;
}

struct ColVal {
  uint32_t c;
  float v;
};

// Sorts one row's (col, val) pairs stably by column and sums duplicates left
// to right, in place; returns the row's new length.
uint32_t canonicalize_row(uint32_t* col, float* val, uint32_t n, std::vector<ColVal>& tmp) {
  bool sorted = true;
  for (uint32_t i = 1; i < n && sorted; ++i) sorted = col[i - 1] < col[i];
  if (sorted) return n;
  if (n <= 32) {
    for (uint32_t i = 1; i < n; ++i) {
      const uint32_t c = col[i];
      const float v = val[i];
      uint32_t j = i;
      for (; j > 0 && col[j - 1] > c; --j) {
        col[j] = col[j - 1];
        val[j] = val[j - 1];
      }
      col[j] = c;
      val[j] = v;
    }
  } else {
    tmp.resize(n);
    for (uint32_t i = 0; i < n; ++i) tmp[i] = ColVal{col[i], val[i]};
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const ColVal& a, const ColVal& b) { return a.c < b.c; });
    for (uint32_t i = 0; i < n; ++i) {
      col[i] = tmp[i].c;
      val[i] = tmp[i].v;
    }
  }
  uint32_t out = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (out > 0 && col[out - 1] == col[i]) {
      val[out - 1] += val[i];
    } else {
      col[out] = col[i];
      val[out] = val[i];
      ++out;
    }
  }
  return out;
}

// Sorts each row's (col, val) pairs by column; duplicates are summed in file
// order.  Rows are sorted in parallel blocks, then compacted left to right
// when duplicates shortened any row.
void canonicalize_rows(accspmm_host_csr* c) {
  const uint32_t rows = c->rows;
  if (rows == 0) return;
  const int parts = c->nnz >= (1u << 16) ? host_threads() : 1;
  const std::vector<uint32_t> blk = row_blocks(c->row_ptr, rows, parts);
  std::vector<uint32_t> len(rows);
  std::vector<uint8_t> shrunk(parts, 0);
  run_parallel(parts, [&](int t) {
    std::vector<ColVal> tmp;
    for (uint32_t r = blk[t]; r < blk[t + 1]; ++r) {
      const uint32_t b = c->row_ptr[r], n = c->row_ptr[r + 1] - b;
      len[r] = canonicalize_row(c->col_idx + b, c->val + b, n, tmp);
      if (len[r] != n) shrunk[t] = 1;
    }
  });
  if (std::find(shrunk.begin(), shrunk.end(), 1) == shrunk.end()) return;
  uint32_t out = 0;
  for (uint32_t r = 0; r < rows; ++r) {
    const uint32_t b = c->row_ptr[r];
    if (out != b) {
      std::memmove(c->col_idx + out, c->col_idx + b, sizeof(uint32_t) * len[r]);
      std::memmove(c->val + out, c->val + b, sizeof(float) * len[r]);
    }
    c->row_ptr[r] = out;
    out += len[r];
  }
  c->row_ptr[rows] = out;
  c->nnz = out;
}

// Stable counting sort by row (the semantics of include/format.hpp:110-142).
void coo_to_csr(uint32_t rows, const std::vector<uint32_t>& R,
                const std::vector<uint32_t>& C, const std::vector<float>& V,
                accspmm_host_csr* out) {
  uint64_t nnz = R.size();
  std::fill(out->row_ptr, out->row_ptr + rows + 1, 0u);
  for (uint64_t i = 0; i < nnz; ++i) out->row_ptr[R[i] + 1]++;
  for (uint32_t i = 0; i < rows; ++i) out->row_ptr[i + 1] += out->row_ptr[i];
  std::vector<uint32_t> fill(out->row_ptr, out->row_ptr + rows);
  for (uint64_t i = 0; i < nnz; ++i) {
    uint32_t d = fill[R[i]]++;
    out->col_idx[d] = C[i];
    out->val[d] = V[i];
  }
}

// Minimal tokenizer over an in-memory file.
struct Cursor {
  const char* p;
  const char* end;
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r'))
      ++p;
  }
  bool read_u32(uint32_t* v) {
    skip_ws();
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
      x = x * 10 + (uint64_t)(*p - '0');
      if (x > 0xFFFFFFFFull) return false;
      ++p;
    }
    *v = (uint32_t)x;
    return true;
  }
  bool read_double(double* v) {
    skip_ws();
    if (p >= end) return false;
    char buf[128];
    size_t n = 0;
    while (p + n < end && n < sizeof(buf) - 1 && p[n] != ' ' && p[n] != '\t' &&
           p[n] != '\n' && p[n] != '\r')
      ++n;
    std::memcpy(buf, p, n);
    buf[n] = 0;
    char* e = nullptr;
    errno = 0;
    *v = std::strtod(buf, &e);
    if (e == buf) return false;
    p += (e - buf);
    return true;
  }
  bool read_long(long* v) {
    skip_ws();
    if (p >= end) return false;
    char buf[64];
    size_t n = 0;
    while (p + n < end && n < sizeof(buf) - 1 && p[n] != ' ' && p[n] != '\t' &&
           p[n] != '\n' && p[n] != '\r')
      ++n;
    std::memcpy(buf, p, n);
    buf[n] = 0;
    char* e = nullptr;
    *v = std::strtol(buf, &e, 10);
    if (e == buf) return false;
    p += (e - buf);
    return true;
  }
};

// ---- parallel entry parser ------------------------------------------------
//
// The entry section is cut at line starts into one chunk per thread.  Each
// chunk is parsed as whole lines "row col [value [imag]]"; a line of any other
// shape (an entry split over lines, a comment, extra tokens, an index out of
// range) or an entry count other than the size line's makes the caller parse
// the file with the sequential tokenizer instead, so the result and the error
// behaviour are those of the sequential reader on every file.

constexpr size_t kParallelMinBytes = 1u << 20;
constexpr uint64_t kHistogramBudget = 512ull << 20;  // per-group row counters

// 10^0 .. 10^22: the powers of ten that are exact doubles
constexpr double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                               1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                               1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

inline bool is_hspace(char ch) { return ch == ' ' || ch == '\t' || ch == '\r'; }
inline bool is_ws(char ch) { return is_hspace(ch) || ch == '\n'; }

// A decimal token [sign] digits [. digits] [(e|E) [sign] digits] whose
// significand is below 2^53 and whose decimal exponent is within +-22 is an
// exact double times or divided by an exact power of ten, so one IEEE
// multiply or divide gives the correctly rounded double that strtod returns
// (Clinger's fast path: W. D. Clinger, "How to read floating point numbers
// accurately", PLDI 1990).  Anything else returns false for strtod to parse.
bool fast_decimal(const char* p, const char* e, double* out) {
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  uint64_t mant = 0;
  int exp10 = 0, digits = 0;
  for (; p < e && *p >= '0' && *p <= '9'; ++p, ++digits) {
    mant = mant * 10 + (uint64_t)(*p - '0');
    if (mant >= (1ull << 53)) return false;
  }
  if (p < e && *p == '.') {
    for (++p; p < e && *p >= '0' && *p <= '9'; ++p, ++digits, --exp10) {
      mant = mant * 10 + (uint64_t)(*p - '0');
      if (mant >= (1ull << 53)) return false;
    }
  }
  if (digits == 0) return false;
  if (p < e && (*p == 'e' || *p == 'E')) {
    bool eneg = false;
    if (++p < e && (*p == '+' || *p == '-')) eneg = *p++ == '-';
    int x = 0, ed = 0;
    for (; p < e && *p >= '0' && *p <= '9'; ++p) {
      if (++ed > 4) return false;
      x = x * 10 + (*p - '0');
    }
    if (ed == 0) return false;
    exp10 += eneg ? -x : x;
  }
  if (p != e) return false;
  double v = 0.0;
  if (mant != 0) {
    if (exp10 < -22 || exp10 > 22) return false;
    v = exp10 < 0 ? (double)mant / kPow10[-exp10] : (double)mant * kPow10[exp10];
  }
  *out = neg ? -v : v;
  return true;
}

// strtod / strtol over one whitespace-delimited token, as Cursor reads it;
// false unless the whole token is consumed.
bool token_double(const char* p, const char* e, double* out) {
  if (fast_decimal(p, e, out)) return true;
  char buf[128];
  const size_t n = (size_t)(e - p);
  if (n == 0 || n >= sizeof(buf)) return false;
  std::memcpy(buf, p, n);
  buf[n] = 0;
  char* end = nullptr;
  *out = std::strtod(buf, &end);
  return end == buf + n;
}

bool token_long(const char* p, const char* e, long* out) {
  const char* q = p;
  bool neg = false;
  if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
  if (q < e && e - q <= 18) {
    long x = 0;
    for (; q < e && *q >= '0' && *q <= '9'; ++q) x = x * 10 + (*q - '0');
    if (q == e) {
      *out = neg ? -x : x;
      return true;
    }
  }
  char buf[64];
  const size_t n = (size_t)(e - p);
  if (n == 0 || n >= sizeof(buf)) return false;
  std::memcpy(buf, p, n);
  buf[n] = 0;
  char* end = nullptr;
  *out = std::strtol(buf, &end, 10);
  return end == buf + n;
}

inline bool line_u32(const char*& p, const char* e, uint32_t* v) {
  if (p >= e || *p < '0' || *p > '9') return false;
  uint64_t x = 0;
  for (; p < e && *p >= '0' && *p <= '9'; ++p) {
    x = x * 10 + (uint64_t)(*p - '0');
    if (x > 0xFFFFFFFFull) return false;
  }
  *v = (uint32_t)x;
  return true;
}

struct Part {
  HostVec<uint32_t> r, c;  // 0-based, file order
  HostVec<float> v;
  uint64_t expanded = 0;  // entries after symmetric expansion
};

bool parse_lines(const char* p, const char* e, char field, bool sym, uint32_t m,
                 uint32_t n, Part* out) {
  const size_t guess = (size_t)(e - p) / 16;
  out->r.reserve(guess);
  out->c.reserve(guess);
  out->v.reserve(guess);
  const int nvals = field == 'P' ? 0 : field == 'C' ? 2 : 1;
  while (p < e) {
    while (p < e && is_hspace(*p)) ++p;
    if (p == e) break;
    if (*p == '\n') {
      ++p;
      continue;
    }
    uint32_t r, c;
    if (!line_u32(p, e, &r) || p == e || !is_hspace(*p)) return false;
    while (p < e && is_hspace(*p)) ++p;
    if (!line_u32(p, e, &c)) return false;
    double v = 1.0;
    for (int k = 0; k < nvals; ++k) {
      if (p == e || !is_hspace(*p)) return false;
      while (p < e && is_hspace(*p)) ++p;
      const char* tb = p;
      while (p < e && !is_ws(*p)) ++p;
      if (field == 'I') {
        long iv;
        if (!token_long(tb, p, &iv)) return false;
        v = (double)iv;
      } else {
        double x;
        if (!token_double(tb, p, &x)) return false;
        if (k == 0) v = x;
      }
    }
    while (p < e && is_hspace(*p)) ++p;
    if (p < e) {
      if (*p != '\n') return false;
      ++p;
    }
    if (r == 0 || c == 0 || r > m || c > n) return false;
    out->r.push_back(r - 1);
    out->c.push_back(c - 1);
    out->v.push_back((float)v);
    out->expanded += (sym && r != c) ? 2 : 1;
  }
  return true;
}

// Parses [p, end) on `threads` threads and builds the CSR with the order of
// coo_to_csr over the sequentially expanded entries: entries are counted per
// row by groups of consecutive chunks, and group g writes row r's entries
// after those of groups before it.  Returns false (nothing allocated) when a
// chunk is irregular or the entry count differs from nz.
bool read_entries_parallel(const char* p, const char* end, char field, bool sym, uint32_t m,
                           uint32_t n, uint32_t nz, int threads, accspmm_host_csr* out,
                           accspmm_status* st) {
  std::vector<const char*> cut(threads + 1, end);
  cut[0] = p;
  const size_t len = (size_t)(end - p);
  for (int i = 1; i < threads; ++i) {
    const char* s = p + len * (size_t)i / (size_t)threads;
    s = std::max(s, cut[i - 1]);
    const char* nl = s < end ? (const char*)std::memchr(s, '\n', (size_t)(end - s)) : nullptr;
    cut[i] = nl ? nl + 1 : end;
  }
  std::vector<Part> parts(threads);
  std::vector<uint8_t> ok(threads, 0);
  run_parallel(threads, [&](int t) {
    ok[t] = parse_lines(cut[t], cut[t + 1], field, sym, m, n, &parts[t]) ? 1 : 0;
  });
  uint64_t entries = 0, total = 0;
  for (int t = 0; t < threads; ++t) {
    if (!ok[t]) return false;
    entries += parts[t].r.size();
    total += parts[t].expanded;
  }
  if (entries != nz) return false;
  if (total > 0xFFFFFFFFull) {
    *st = ACCSPMM_ERR_UNSUPPORTED;
    return true;
  }
  if (!alloc_csr(out, m, n, total)) {
    *st = ACCSPMM_ERR_OUT_OF_MEMORY;
    return true;
  }
  // groups of consecutive chunks, each with a counter per row
  const uint64_t per_group = 4ull * m + 4;
  const int groups =
      (int)std::max<uint64_t>(1, std::min<uint64_t>(threads, kHistogramBudget / per_group));
  std::vector<HostVec<uint32_t>> cnt(groups);
  auto group_parts = [&](int g, int* t0, int* t1) {
    *t0 = g * threads / groups;
    *t1 = (g + 1) * threads / groups;
  };
  run_parallel(groups, [&](int g) {
    HostVec<uint32_t>& h = cnt[g];
    h.assign(m, 0u);
    int t0, t1;
    group_parts(g, &t0, &t1);
    for (int t = t0; t < t1; ++t) {
      const Part& P = parts[t];
      for (size_t i = 0; i < P.r.size(); ++i) {
        ++h[P.r[i]];
        if (sym && P.r[i] != P.c[i]) ++h[P.c[i]];
      }
    }
  });
  // per-row group offsets and row lengths, then the row prefix
  const int rthreads = m >= (1u << 16) ? threads : 1;
  run_parallel(rthreads, [&](int t) {
    const uint32_t r0 = (uint32_t)((uint64_t)m * t / rthreads);
    const uint32_t r1 = (uint32_t)((uint64_t)m * (t + 1) / rthreads);
    for (uint32_t r = r0; r < r1; ++r) {
      uint32_t s = 0;
      for (int g = 0; g < groups; ++g) {
        const uint32_t x = cnt[g][r];
        cnt[g][r] = s;
        s += x;
      }
      out->row_ptr[r + 1] = s;
    }
  });
  out->row_ptr[0] = 0;
  for (uint32_t r = 0; r < m; ++r) out->row_ptr[r + 1] += out->row_ptr[r];
  run_parallel(groups, [&](int g) {
    HostVec<uint32_t>& h = cnt[g];
    const uint32_t* rp = out->row_ptr;
    int t0, t1;
    group_parts(g, &t0, &t1);
    for (int t = t0; t < t1; ++t) {
      const Part& P = parts[t];
      for (size_t i = 0; i < P.r.size(); ++i) {
        const uint32_t r = P.r[i], c = P.c[i];
        const uint32_t d = rp[r] + h[r]++;
        out->col_idx[d] = c;
        out->val[d] = P.v[i];
        if (sym && r != c) {
          const uint32_t d2 = rp[c] + h[c]++;
          out->col_idx[d2] = r;
          out->val[d2] = P.v[i];
        }
      }
    }
  });
  *st = ACCSPMM_OK;
  return true;
}

thread_local int g_threads_used = 0;

// The whole file, memory-mapped read-only.
struct MappedFile {
  const char* data = nullptr;
  size_t size = 0;
  void* map = nullptr;
  ~MappedFile() {
    if (map) munmap(map, size);
  }
  bool open_file(const char* path) {
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat sb;
    if (fstat(fd, &sb) != 0 || !S_ISREG(sb.st_mode)) {
      ::close(fd);
      return false;
    }
    size = (size_t)sb.st_size;
    if (size > 0) {
      void* q = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (q == MAP_FAILED) {
        ::close(fd);
        return false;
      }
      map = q;
      data = static_cast<const char*>(q);
    }
    ::close(fd);
    return true;
  }
};

}  // namespace

// Matrix Market reader.  Accepts what read_from_mtx accepts
// (include/mmio.hpp:24-93: five case-sensitive banner tokens, '%' comment
// lines before the size line, unknown storage token tolerated), expands
// symmetric/Hermitian entries as (r,c) then (c,r) when r != c
// (include/mmio.hpp:133-158), gives pattern entries 1.0 and casts each value
// double -> float.  Divergences: integer values are parsed as integers (the
// reference reads %d into a double); a malformed entry line is an error
// rather than a repeat of the previous entry.  Files with at least 1 MiB of
// entries are parsed on several threads (same result; see read_entries_parallel).
static accspmm_status read_mtx_impl(const char* path, int canonicalize,
                                    accspmm_host_csr* out, int* mtx_code) {
  MappedFile file;
  if (!file.open_file(path)) {
    if (mtx_code) *mtx_code = -1;
    return ACCSPMM_ERR_MTX;
  }
  const char* p = file.data;
  const char* end = p + file.size;
  auto next_line = [&](std::string* line) -> bool {
    if (p >= end) return false;
    const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    const char* le = nl ? nl : end;
    line->assign(p, le);
    p = nl ? nl + 1 : end;
    return true;
  };
  std::string line;
  if (!next_line(&line)) {
    if (mtx_code) *mtx_code = -2;
    return ACCSPMM_ERR_MTX;
  }
  char tok[5][64];
  if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", tok[0], tok[1],
                  tok[2], tok[3], tok[4]) != 5 ||
      std::strcmp(tok[0], "%%MatrixMarket") != 0 ||
      std::strcmp(tok[1], "matrix") != 0) {
    if (mtx_code) *mtx_code = -2;
    return ACCSPMM_ERR_MTX;
  }
  char field;
  if (!std::strcmp(tok[3], "real")) field = 'R';
  else if (!std::strcmp(tok[3], "complex")) field = 'C';
  else if (!std::strcmp(tok[3], "pattern")) field = 'P';
  else if (!std::strcmp(tok[3], "integer")) field = 'I';
  else {
    if (mtx_code) *mtx_code = -2;
    return ACCSPMM_ERR_MTX;
  }
  char sym;
  if (!std::strcmp(tok[4], "general")) sym = 'G';
  else if (!std::strcmp(tok[4], "symmetric")) sym = 'S';
  else if (!std::strcmp(tok[4], "Hermitian")) sym = 'H';
  else if (!std::strcmp(tok[4], "skew-symmetric")) sym = 'K';
  else {
    if (mtx_code) *mtx_code = -2;
    return ACCSPMM_ERR_MTX;
  }
  if (std::strcmp(tok[2], "array") == 0) {
    if (mtx_code) *mtx_code = 3;
    return ACCSPMM_ERR_UNSUPPORTED;  // dense files are not sparse input
  }
  uint32_t m = 0, n = 0, nz = 0;
  for (;;) {
    if (!next_line(&line)) {
      if (mtx_code) *mtx_code = 3;
      return ACCSPMM_ERR_MTX;
    }
    if (!line.empty() && line[0] == '%') continue;
    break;
  }
  {
    // three unsigned 32-bit integers (a value that overflows, or a sign,
    // is an error); the reference keeps scanning tokens on the following
    // lines until three integers are read
    Cursor c0{line.c_str(), line.c_str() + line.size()};
    uint32_t got[3];
    int k = 0;
    while (k < 3 && c0.read_u32(&got[k])) ++k;
    if (k < 3) {
      c0.skip_ws();
      if (c0.p < c0.end) {  // a token that is not an unsigned 32-bit integer
        if (mtx_code) *mtx_code = 3;
        return ACCSPMM_ERR_MTX;
      }
      Cursor c1{p, end};
      while (k < 3) {
        if (!c1.read_u32(&got[k])) {
          if (mtx_code) *mtx_code = 3;
          return ACCSPMM_ERR_MTX;
        }
        ++k;
      }
      p = c1.p;
    }
    m = got[0];
    n = got[1];
    nz = got[2];
  }
  bool symmetric = sym == 'S' || sym == 'H';
  if (symmetric && m != n) {
    // the mirrored entry (c, r) of a non-square matrix can fall outside it
    if (mtx_code) *mtx_code = 3;
    return ACCSPMM_ERR_MTX;
  }
  const int threads = host_threads();
  if (threads > 1 && (size_t)(end - p) >= kParallelMinBytes) {
    accspmm_status st = ACCSPMM_OK;
    if (read_entries_parallel(p, end, field, symmetric, m, n, nz, threads, out, &st)) {
      g_threads_used = threads;
      if (st == ACCSPMM_OK && canonicalize) canonicalize_rows(out);
      return st;
    }
  }
  g_threads_used = 1;
  std::vector<uint32_t> R, C;
  std::vector<float> V;
  // an entry takes at least 4 bytes ("1 1\n"): do not trust nz for the reserve
  const uint64_t fit = std::min<uint64_t>(nz, (uint64_t)(end - p) / 4 + 1);
  R.reserve(symmetric ? 2 * fit : fit);
  C.reserve(R.capacity());
  V.reserve(R.capacity());
  Cursor cur{p, end};
  for (uint32_t i = 0; i < nz; ++i) {
    uint32_t r, c;
    double v = 1.0;
    if (!cur.read_u32(&r) || !cur.read_u32(&c)) {
      if (mtx_code) *mtx_code = 3;
      return ACCSPMM_ERR_MTX;
    }
    if (field == 'R') {
      if (!cur.read_double(&v)) { if (mtx_code) *mtx_code = 3; return ACCSPMM_ERR_MTX; }
    } else if (field == 'C') {
      double im;
      if (!cur.read_double(&v) || !cur.read_double(&im)) {
        if (mtx_code) *mtx_code = 3;
        return ACCSPMM_ERR_MTX;
      }
    } else if (field == 'I') {
      long iv;
      if (!cur.read_long(&iv)) { if (mtx_code) *mtx_code = 3; return ACCSPMM_ERR_MTX; }
      v = (double)iv;
    }
    if (r == 0 || c == 0 || r > m || c > n) {
      if (mtx_code) *mtx_code = 3;
      return ACCSPMM_ERR_MTX;
    }
    R.push_back(r - 1);
    C.push_back(c - 1);
    V.push_back((float)v);
    if (symmetric && r != c) {
      R.push_back(c - 1);
      C.push_back(r - 1);
      V.push_back((float)v);
    }
  }
  if (R.size() > 0xFFFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  if (!alloc_csr(out, m, n, R.size())) return ACCSPMM_ERR_OUT_OF_MEMORY;
  coo_to_csr(m, R, C, V, out);
  if (canonicalize) canonicalize_rows(out);
  return ACCSPMM_OK;
}

extern "C" accspmm_status accspmm_read_mtx(const char* path, int canonicalize,
                                           accspmm_host_csr* out, int* mtx_code) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!path || !out) return ACCSPMM_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    if (mtx_code) *mtx_code = 0;
    g_threads_used = 0;
    try {
      const accspmm_status st = read_mtx_impl(path, canonicalize, out, mtx_code);
      if (st != ACCSPMM_OK) accspmm_host_csr_free(out);
      return st;
    } catch (const std::exception&) {  // allocation or thread creation failed
      accspmm_host_csr_free(out);
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    }
  });
}

extern "C" int accspmm_read_mtx_threads_used(void) { return g_threads_used; }

extern "C" int accspmm_csr_is_canonical(const accspmm_host_csr* c) {
  if (!c || !c->row_ptr) return 0;
  if (c->row_ptr[0] != 0 || c->row_ptr[c->rows] != c->nnz) return 0;
  for (uint32_t r = 0; r < c->rows; ++r) {
    if (c->row_ptr[r + 1] < c->row_ptr[r]) return 0;
    for (uint32_t e = c->row_ptr[r]; e < c->row_ptr[r + 1]; ++e) {
      if (c->col_idx[e] >= c->cols) return 0;
      if (e > c->row_ptr[r] && c->col_idx[e] <= c->col_idx[e - 1]) return 0;
    }
  }
  return 1;
}

namespace {

inline void put_u32(std::string& s, uint32_t x) {
  char buf[10];
  int n = 0;
  do {
    buf[n++] = (char)('0' + x % 10u);
    x /= 10u;
  } while (x);
  while (n) s.push_back(buf[--n]);
}

// One entry line "row col value" (1-based), the value with 9 significant
// digits, which reads back to the same float.
inline void put_entry(std::string& s, uint32_t r, uint32_t c, float v) {
  put_u32(s, r + 1u);
  s.push_back(' ');
  put_u32(s, c + 1u);
  char buf[32];
  const int n = std::snprintf(buf, sizeof(buf), " %.9g\n", (double)v);
  s.append(buf, (size_t)n);
}

}  // namespace

// Matrix Market writer (coordinate real).  general: every entry, row-major.
// symmetric: the lower triangle in column-major order — for column c the
// entries (r, c), r >= c, taken from row c of the CSR (so the CSR must hold a
// symmetric matrix) — the order in which accspmm_read_mtx's expansion yields
// sorted rows.  Rows are formatted on several threads and written in order.
extern "C" accspmm_status accspmm_write_mtx(const accspmm_host_csr* c, const char* path,
                                            int symmetric) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!c || !path || !c->row_ptr || (c->nnz && (!c->col_idx || !c->val)))
      return ACCSPMM_ERR_INVALID_ARG;
    if (symmetric && c->rows != c->cols) return ACCSPMM_ERR_SHAPE;
    try {
      const int parts = c->nnz >= (1u << 16) ? host_threads() : 1;
      const std::vector<uint32_t> blk = row_blocks(c->row_ptr, c->rows, parts);
      std::vector<std::string> text(parts);
      std::vector<uint64_t> count(parts, 0);
      run_parallel(parts, [&](int t) {
        std::string& s = text[t];
        s.reserve((size_t)(c->row_ptr[blk[t + 1]] - c->row_ptr[blk[t]]) * 24);
        for (uint32_t r = blk[t]; r < blk[t + 1]; ++r) {
          for (uint32_t e = c->row_ptr[r]; e < c->row_ptr[r + 1]; ++e) {
            const uint32_t col = c->col_idx[e];
            if (symmetric) {
              if (col < r) continue;
              put_entry(s, col, r, c->val[e]);  // (col, r) below the diagonal
            } else {
              put_entry(s, r, col, c->val[e]);
            }
            ++count[t];
          }
        }
      });
      uint64_t total = 0;
      for (uint64_t x : count) total += x;
      std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), &std::fclose);
      if (!f) return ACCSPMM_ERR_FILE;
      std::string head = std::string("%%MatrixMarket matrix coordinate real ") +
                         (symmetric ? "symmetric\n" : "general\n");
      head += std::to_string(c->rows) + " " + std::to_string(c->cols) + " " +
              std::to_string(total) + "\n";
      bool ok = std::fwrite(head.data(), 1, head.size(), f.get()) == head.size();
      for (const std::string& s : text)
        ok = ok && std::fwrite(s.data(), 1, s.size(), f.get()) == s.size();
      ok = (std::fclose(f.release()) == 0) && ok;
      return ok ? ACCSPMM_OK : ACCSPMM_ERR_FILE;
    } catch (const std::exception&) {
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    }
  });
}

// ------------------------------------------------------------------------
// Generators
// ------------------------------------------------------------------------

extern "C" void accspmm_gen_dense(uint64_t seed, uint64_t count, float* out) {
  accspmm::Xoshiro256 rng(seed);
  for (uint64_t i = 0; i < count; ++i) out[i] = rng.sym_unit();
}

// Chung-Lu expected-degree graph, sampled with the Miller-Hagberg skip
// method (J. C. Miller, A. Hagberg, "Efficient generation of networks with
// given expected degrees", WAW 2011): weights w_i ∝ (i+1)^(-1/(γ-1)) scaled
// to mean d̄ and capped at sqrt(Σw) (scale and cap iterated to a fixed
// point), each unordered pair {u<v} drawn at most once with probability
// min(1, w_u w_v / Σw), both directions emitted, a self loop on every node,
// then a seeded random relabel and per-row column sort.
extern "C" accspmm_status accspmm_gen_chung_lu(uint32_t n, double dbar,
                                               double gamma, uint64_t seed,
                                               int value_mode,
                                               accspmm_host_csr* out) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!out || n == 0 || !(gamma > 1.0) || !(dbar >= 0.0))
      return ACCSPMM_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    std::vector<double> w(n);
    const double ex = -1.0 / (gamma - 1.0);
    for (uint32_t i = 0; i < n; ++i) w[i] = std::pow((double)(i + 1), ex);
    for (int it = 0; it < 30; ++it) {
      double s = 0.0;
      for (double x : w) s += x;
      double scale = dbar * n / s;
      for (double& x : w) x *= scale;
      s = dbar * n;
      double cap = std::sqrt(s);
      for (double& x : w) x = std::min(x, cap);
    }
    double S = 0.0;
    for (double x : w) S += x;
    accspmm::Xoshiro256 rng(seed);
    std::vector<uint32_t> eu, ev;
    eu.reserve((size_t)(S / 2 * 1.02) + 16);
    ev.reserve(eu.capacity());
    if (S > 0) {
      for (uint32_t u = 0; u + 1 < n; ++u) {
        uint64_t v = u + 1;
        double p = std::min(w[u] * w[v] / S, 1.0);
        while (v < n && p > 0) {
          if (p != 1.0) {
            double r = rng.open_unit();
            double skip = std::floor(std::log(r) / std::log1p(-p));
            if (skip > (double)n) break;
            v += (uint64_t)skip;
          }
          if (v < n) {
            double q = std::min(w[u] * w[v] / S, 1.0);
            double r = rng.open_unit();
            if (r < q / p) {
              eu.push_back(u);
              ev.push_back((uint32_t)v);
            }
            p = q;
            ++v;
          }
        }
      }
    }
    // relabel with a Fisher-Yates permutation
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint32_t i = n - 1; i > 0; --i) {
      uint32_t j = (uint32_t)(rng.next() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    const uint64_t E = eu.size();
    const uint64_t nnz = 2 * E + n;
    if (nnz > 0xFFFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    if (!alloc_csr(out, n, n, nnz)) return ACCSPMM_ERR_OUT_OF_MEMORY;
    std::vector<float> evals;
    std::vector<float> svals;
    if (value_mode == 0) {
      accspmm::Xoshiro256 vr(seed ^ 0x5bd1e995ull);
      evals.resize(E);
      for (uint64_t i = 0; i < E; ++i) evals[i] = vr.pos_unit();
      svals.resize(n);
      for (uint32_t i = 0; i < n; ++i) svals[i] = vr.pos_unit();
    }
    uint32_t* rp = out->row_ptr;
    std::fill(rp, rp + n + 1, 0u);
    for (uint64_t i = 0; i < E; ++i) {
      rp[perm[eu[i]] + 1]++;
      rp[perm[ev[i]] + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) rp[i + 1] += 1;  // self loops
    for (uint32_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<uint32_t> fill(rp, rp + n);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t r = perm[i];
      uint32_t d = fill[r]++;
      out->col_idx[d] = r;
      out->val[d] = value_mode == 0 ? svals[i] : 0.0f;
    }
    for (uint64_t i = 0; i < E; ++i) {
      uint32_t a = perm[eu[i]], b = perm[ev[i]];
      float v = value_mode == 0 ? evals[i] : 0.0f;
      uint32_t d = fill[a]++;
      out->col_idx[d] = b;
      out->val[d] = v;
      d = fill[b]++;
      out->col_idx[d] = a;
      out->val[d] = v;
    }
    eu.clear(); eu.shrink_to_fit();
    ev.clear(); ev.shrink_to_fit();
    // per-row sort (columns are unique: no duplicate pairs, one self loop);
    // rows are independent, so the result does not depend on the split
    {
      unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      std::vector<std::thread> pool;
      for (unsigned ti = 0; ti < nt; ++ti)
        pool.emplace_back([&, ti]() {
          std::vector<std::pair<uint32_t, float>> tmp;
          for (uint64_t r = ti; r < n; r += nt) {
            uint32_t b = rp[r], e = rp[r + 1];
            tmp.resize(e - b);
            for (uint32_t i = b; i < e; ++i) tmp[i - b] = {out->col_idx[i], out->val[i]};
            std::sort(tmp.begin(), tmp.end(),
                      [](const std::pair<uint32_t, float>& x,
                         const std::pair<uint32_t, float>& y) { return x.first < y.first; });
            for (uint32_t i = b; i < e; ++i) {
              out->col_idx[i] = tmp[i - b].first;
              out->val[i] = tmp[i - b].second;
            }
          }
        });
      for (auto& th : pool) th.join();
    }
    if (value_mode == 1) {
      for (uint32_t r = 0; r < n; ++r) {
        float v = (float)(1.0 / (double)(rp[r + 1] - rp[r]));
        for (uint32_t e = rp[r]; e < rp[r + 1]; ++e) out->val[e] = v;
      }
    } else if (value_mode == 2) {
      accspmm::Xoshiro256 vr(seed ^ 0x5bd1e995ull);
      for (uint64_t e = 0; e < nnz; ++e) out->val[e] = vr.sym_unit();
    }
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_gen_stencil27(uint32_t g, uint64_t seed,
                                                accspmm_host_csr* out) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!out || g == 0) return ACCSPMM_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    uint64_t n = (uint64_t)g * g * g;
    uint64_t per = 3ull * g - 2;
    uint64_t nnz = per * per * per;
    if (n > 0xFFFFFFFFull || nnz > 0xFFFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    if (!alloc_csr(out, (uint32_t)n, (uint32_t)n, nnz))
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    accspmm::Xoshiro256 vr(seed);
    uint64_t k = 0;
    for (uint64_t id = 0; id < n; ++id) {
      out->row_ptr[id] = (uint32_t)k;
      int64_t x = id % g, y = (id / g) % g, z = id / ((uint64_t)g * g);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int64_t X = x + dx, Y = y + dy, Z = z + dz;
            if (X < 0 || Y < 0 || Z < 0 || X >= g || Y >= g || Z >= g) continue;
            out->col_idx[k] = (uint32_t)(X + (int64_t)g * Y + (int64_t)g * g * Z);
            out->val[k] = vr.sym_unit();
            ++k;
          }
    }
    out->row_ptr[n] = (uint32_t)k;
    return ACCSPMM_OK;
  });
}

// Reverse Cuthill-McKee ordering of a square pattern (Cuthill and McKee,
// "Reducing the bandwidth of sparse symmetric matrices", 1969; reversed as
// George proposed; start nodes by the pseudo-peripheral search of Gibbs,
// Poole and Stockmeyer as George and Liu give it, "Computer Solution of
// Large Sparse Positive Definite Systems", 1981, sec. 4.3).  Row i's
// columns are its neighbours; every connected component is ordered from a
// pseudo-peripheral start node, breadth first, each node's unvisited
// neighbours in increasing degree (ties: smaller index).  The whole order is
// then reversed.
namespace {

struct RcmGraph {
  uint32_t n;
  const uint32_t* rp;
  const uint32_t* ci;
  uint32_t deg(uint32_t v) const { return rp[v + 1] - rp[v]; }
};

// Breadth-first level structure from r inside r's component (mark holds
// `stamp` for visited nodes, 1 for nodes already placed): returns the
// number of levels and leaves the last level's nodes in
// [last_begin, order.size()).
uint32_t rcm_levels(const RcmGraph& g, uint32_t r, std::vector<uint32_t>& mark, uint32_t stamp,
                    std::vector<uint32_t>& order, size_t* last_begin) {
  order.clear();
  order.push_back(r);
  mark[r] = stamp;
  size_t begin = 0;
  uint32_t levels = 0;
  while (begin < order.size()) {
    const size_t end = order.size();
    *last_begin = begin;
    ++levels;
    for (size_t i = begin; i < end; ++i) {
      const uint32_t v = order[i];
      for (uint32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
        const uint32_t w = g.ci[e];
        // mark 1: already placed (another component, reached through an
        // unsymmetric entry)
        if (mark[w] != stamp && mark[w] != 1) {
          mark[w] = stamp;
          order.push_back(w);
        }
      }
    }
    begin = end;
  }
  return levels;
}

}  // namespace

extern "C" accspmm_status accspmm_order_rcm(uint32_t n, const uint32_t* row_ptr,
                                            const uint32_t* col_idx, uint32_t* perm) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!row_ptr || !perm || n == 0) return ACCSPMM_ERR_INVALID_ARG;
    if (row_ptr[0] != 0) return ACCSPMM_ERR_INVALID_ARG;
    for (uint32_t v = 0; v < n; ++v)
      if (row_ptr[v + 1] < row_ptr[v]) return ACCSPMM_ERR_INVALID_ARG;
    if (row_ptr[n] && !col_idx) return ACCSPMM_ERR_INVALID_ARG;
    for (uint64_t e = 0; e < row_ptr[n]; ++e)
      if (col_idx[e] >= n) return ACCSPMM_ERR_INVALID_ARG;
    try {
      const RcmGraph g{n, row_ptr, col_idx};
      // stamps: 1 = placed in the final order; >= 2 the search passes
      std::vector<uint32_t> mark(n, 0), order, nbr;
      order.reserve(n);
      std::vector<uint32_t> out;
      out.reserve(n);
      // component roots are taken in increasing degree (ties: index)
      std::vector<uint32_t> by_deg(n);
      std::iota(by_deg.begin(), by_deg.end(), 0u);
      std::stable_sort(by_deg.begin(), by_deg.end(),
                       [&](uint32_t a, uint32_t b) { return g.deg(a) < g.deg(b); });
      uint32_t stamp = 1;
      // a fresh search stamp; before the counter wraps, clear every search
      // mark (placed nodes keep 1) so an old stamp never reads as current
      auto next_stamp = [&]() {
        if (stamp == 0xFFFFFFFFu) {
          for (uint32_t& m : mark)
            if (m != 1) m = 0;
          stamp = 1;
        }
        return ++stamp;
      };
      for (uint32_t root : by_deg) {
        if (mark[root] == 1) continue;
        // pseudo-peripheral node: move to a smallest-degree node of the last
        // level while that deepens the level structure (bounded passes)
        uint32_t r = root;
        size_t last = 0;
        uint32_t depth = rcm_levels(g, r, mark, next_stamp(), order, &last);
        for (int pass = 0; pass < 8; ++pass) {
          uint32_t best = order[last];
          for (size_t i = last; i < order.size(); ++i)
            if (g.deg(order[i]) < g.deg(best) || (g.deg(order[i]) == g.deg(best) && order[i] < best))
              best = order[i];
          size_t last2 = 0;
          std::vector<uint32_t> order2;
          order2.reserve(order.size());
          const uint32_t d2 = rcm_levels(g, best, mark, next_stamp(), order2, &last2);
          if (d2 <= depth) break;
          r = best;
          depth = d2;
          order.swap(order2);
          last = last2;
        }
        // Cuthill-McKee from r over the component
        const size_t first = out.size();
        out.push_back(r);
        mark[r] = 1;
        for (size_t i = first; i < out.size(); ++i) {
          const uint32_t v = out[i];
          nbr.clear();
          for (uint32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
            const uint32_t w = g.ci[e];
            if (mark[w] != 1) {
              mark[w] = 1;
              nbr.push_back(w);
            }
          }
          std::sort(nbr.begin(), nbr.end(), [&](uint32_t a, uint32_t b) {
            return g.deg(a) != g.deg(b) ? g.deg(a) < g.deg(b) : a < b;
          });
          out.insert(out.end(), nbr.begin(), nbr.end());
        }
      }
      std::reverse(out.begin(), out.end());
      std::copy(out.begin(), out.end(), perm);
    } catch (const std::bad_alloc&) {
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    }
    return ACCSPMM_OK;
  });
}
