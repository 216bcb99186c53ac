// host.cpp — host-side utilities of libaccspmm: Matrix Market ingest, CSR
// construction and deterministic graph generators.  No GPU code here.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/accspmm.h"
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

bool alloc_csr(accspmm_host_csr* out, uint32_t rows, uint32_t cols,
               uint64_t nnz) {
  out->rows = rows;
  out->cols = cols;
  out->nnz = (uint32_t)nnz;
  out->row_ptr = (uint32_t*)std::malloc(sizeof(uint32_t) * (rows + 1ull));
  out->col_idx = (uint32_t*)std::malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  out->val = (float*)std::malloc(sizeof(float) * (nnz ? nnz : 1));
  if (!out->row_ptr || !out->col_idx || !out->val) {
    accspmm_host_csr_free(out);
    return false;
  }
  return true;
}

// Sorts each row's (col, val) pairs by column; duplicates are summed.
void canonicalize_rows(accspmm_host_csr* c) {
  std::vector<uint32_t> perm;
  std::vector<uint32_t> ctmp;
  std::vector<float> vtmp;
  uint32_t out = 0;
  uint32_t start = 0;
  for (uint32_t r = 0; r < c->rows; ++r) {
    uint32_t b = start, e = c->row_ptr[r + 1];
    start = e;
    uint32_t n = e - b;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) {
      return c->col_idx[b + x] < c->col_idx[b + y];
    });
    ctmp.resize(n);
    vtmp.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      ctmp[i] = c->col_idx[b + perm[i]];
      vtmp[i] = c->val[b + perm[i]];
    }
    c->row_ptr[r] = out;
    for (uint32_t i = 0; i < n; ++i) {
      if (out > c->row_ptr[r] && c->col_idx[out - 1] == ctmp[i]) {
        c->val[out - 1] += vtmp[i];
      } else {
        c->col_idx[out] = ctmp[i];
        c->val[out] = vtmp[i];
        ++out;
      }
    }
  }
  c->row_ptr[c->rows] = out;
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

}  // namespace

// Matrix Market reader.  Accepts what read_from_mtx accepts
// (include/mmio.hpp:24-93: five case-sensitive banner tokens, '%' comment
// lines before the size line, unknown storage token tolerated), expands
// symmetric/Hermitian entries as (r,c) then (c,r) when r != c
// (include/mmio.hpp:133-158), gives pattern entries 1.0 and casts each value
// double -> float.  Divergences: integer values are parsed as integers (the
// reference reads %d into a double); a malformed entry line is an error
// rather than a repeat of the previous entry.
extern "C" accspmm_status accspmm_read_mtx(const char* path, int canonicalize,
                                           accspmm_host_csr* out,
                                           int* mtx_code) {
  if (!path || !out) return ACCSPMM_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  if (mtx_code) *mtx_code = 0;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    if (mtx_code) *mtx_code = -1;
    return ACCSPMM_ERR_MTX;
  }
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string text((size_t)(sz > 0 ? sz : 0), '\0');
  if (sz > 0 && std::fread(&text[0], 1, (size_t)sz, f) != (size_t)sz) {
    std::fclose(f);
    if (mtx_code) *mtx_code = -1;
    return ACCSPMM_ERR_MTX;
  }
  std::fclose(f);
  const char* p = text.data();
  const char* end = p + text.size();
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
  if (std::sscanf(line.c_str(), "%u %u %u", &m, &n, &nz) != 3) {
    // the reference keeps scanning tokens until three integers are read
    Cursor c0{line.c_str(), line.c_str() + line.size()};
    uint32_t got[3];
    int k = 0;
    while (k < 3 && c0.read_u32(&got[k])) ++k;
    Cursor c1{p, end};
    while (k < 3) {
      if (!c1.read_u32(&got[k])) {
        if (mtx_code) *mtx_code = 3;
        return ACCSPMM_ERR_MTX;
      }
      ++k;
    }
    p = c1.p;
    m = got[0];
    n = got[1];
    nz = got[2];
  }
  bool symmetric = sym == 'S' || sym == 'H';
  std::vector<uint32_t> R, C;
  std::vector<float> V;
  R.reserve(symmetric ? 2ull * nz : nz);
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
  if (!alloc_csr(out, m, n, R.size())) return ACCSPMM_ERR_OUT_OF_MEMORY;
  coo_to_csr(m, R, C, V, out);
  if (canonicalize) canonicalize_rows(out);
  return ACCSPMM_OK;
}

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
}

extern "C" accspmm_status accspmm_gen_stencil27(uint32_t g, uint64_t seed,
                                                accspmm_host_csr* out) {
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
}
