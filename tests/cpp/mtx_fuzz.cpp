// Random mutations of Matrix Market text through accspmm_read_mtx, built
// with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitized.py.  Any status is acceptable; a file that is
// accepted must give a consistent CSR (row_ptr from 0 to nnz and monotone,
// every column below cols, rows sorted and duplicate-free when canonicalized;
// rows are walked when there are at most 2^24).
// Small files take the sequential tokenizer, one in 25 is large enough for
// the line-parallel parser.  Exit 0 when every case held.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "accspmm.h"

namespace {

bool consistent(const accspmm_host_csr& c, bool canonical) {
  if (c.row_ptr[0] != 0 || c.row_ptr[c.rows] != c.nnz) return false;
  for (uint32_t e = 0; e < c.nnz; ++e)
    if (c.col_idx[e] >= c.cols) return false;
  if (c.rows > (1u << 24)) return true;  // a mutated size line; rows checked below
  for (uint32_t r = 0; r < c.rows; ++r) {
    if (c.row_ptr[r + 1] < c.row_ptr[r]) return false;
    for (uint32_t e = c.row_ptr[r]; e < c.row_ptr[r + 1]; ++e) {
      if (c.col_idx[e] >= c.cols) return false;
      if (canonical && e > c.row_ptr[r] && c.col_idx[e] <= c.col_idx[e - 1]) return false;
    }
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int iters = std::atoi(argv[1]);
  const char* path = argv[2];
  std::mt19937_64 rng(12345);
  const std::string small =
      "%%MatrixMarket matrix coordinate real symmetric\n% comment\n4 4 5\n"
      "1 1 1.0\n2 1 -2.5e3\n3 3 7\n4 2 0.5\n4 4 1e-7\n";
  std::string big = "%%MatrixMarket matrix coordinate real general\n500 400 60000\n";
  for (int i = 0; i < 60000; ++i)
    big += std::to_string(rng() % 500 + 1) + " " + std::to_string(rng() % 400 + 1) + " " +
           std::to_string((double)(rng() % 100000) / 977.0) + "\n";
  const char* alphabet = "0123456789 \n\t.-+eE%abcxyz\r";
  const size_t na = std::strlen(alphabet);
  int parsed = 0, rejected = 0;
  for (int it = 0; it < iters; ++it) {
    std::string s = it % 25 == 0 ? big : small;
    const int edits = 1 + (int)(rng() % 6);
    for (int e = 0; e < edits && !s.empty(); ++e) {
      const size_t pos = rng() % s.size();
      switch (rng() % 5) {
        case 0: s[pos] = alphabet[rng() % na]; break;
        case 1: s.erase(pos, 1 + rng() % 8); break;
        case 2: s.insert(pos, 1, alphabet[rng() % na]); break;
        case 3: s.resize(pos); break;
        default: s.insert(pos, std::to_string(rng())); break;
      }
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return 3;
    std::fwrite(s.data(), 1, s.size(), f);
    std::fclose(f);
    accspmm_host_csr c;
    std::memset(&c, 0, sizeof c);
    int code = 0;
    const bool canon = (it & 1) != 0;
    if (accspmm_read_mtx(path, canon ? 1 : 0, &c, &code) == ACCSPMM_OK) {
      ++parsed;
      const bool ok = consistent(c, canon);
      accspmm_host_csr_free(&c);
      if (!ok) {
        std::printf("inconsistent CSR from mutation %d\n", it);
        return 1;
      }
    } else {
      ++rejected;
    }
  }
  std::printf("fuzz done: %d parsed, %d rejected\n", parsed, rejected);
  return 0;
}
