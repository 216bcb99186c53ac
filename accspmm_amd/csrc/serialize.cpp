// serialize.cpp — binary save and load of a resident METCF(+AdpBME) graph.
//
// No reference counterpart: the reference rebuilds the format from the
// Matrix Market file on every run (src/mma_tf32.cu:195-226).  A saved graph
// lets a server warm-start its resident set without rebuilding.
//
// File layout (little-endian):
//   "ACCSPMTC" (8 bytes), format version (u32, = 3; versions 1 and 2 are
//     read too)
//   header (u32 each): rows, cols, nnz, row_window, W, T, G, ibd, balanced,
//     max_window_tiles, has_adpbme, the 8 build options, tuning-key length,
//     from version 2 the reorder_rows option (1 when the file holds a row
//     order), from version 3 the symmetric_order option (then the row order
//     is the column order too)
//   tuning key (bytes)
//   rowWindowOffset[W+1] u32, tcOffset[T+1] u32, tcLocalBit[T*RW/8] u64,
//   sparseA2B[8T] u32, data[nnz] f32, and with AdpBME groupOffset[G+1] u32,
//   rowIndices[T] u32 — the arrays accspmm_metcf_export returns — and with
//   reorder_rows the row order[rows] u32 (accspmm_metcf_row_order)
//   FNV-1a 64 of every byte before it (u64)
// Loading checks the header, every size, the checksum and the offsets'
// invariants (monotone, popcounts match tcOffset), then rebuilds the
// kernel-side padding the builder adds (padding tile, value zeros).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "internal.hpp"

using namespace accspmm;

namespace {

constexpr char kMagic[8] = {'A', 'C', 'C', 'S', 'P', 'M', 'T', 'C'};
constexpr uint32_t kVersion = 3;
constexpr uint32_t kHeaderWordsV1 = 19;
constexpr uint32_t kHeaderWordsV2 = 20;
constexpr uint32_t kHeaderWords = 21;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
};

struct Writer {
  FILE* f = nullptr;
  Fnv sum;
  bool ok = true;
  void put(const void* p, size_t n) {
    if (!ok || n == 0) return;
    sum.add(p, n);
    ok = std::fwrite(p, 1, n, f) == n;
  }
};

struct Reader {
  FILE* f = nullptr;
  Fnv sum;
  bool ok = true;
  void get(void* p, size_t n) {
    if (!ok || n == 0) return;
    ok = std::fread(p, 1, n, f) == n;
    if (ok) sum.add(p, n);
  }
};

template <class T>
accspmm_status to_host(std::vector<T>& v, const DevBuf& d, size_t n) {
  v.resize(n);
  if (n) ACC_HIP(hipMemcpy(v.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost));
  return ACCSPMM_OK;
}

template <class T>
accspmm_status to_device(DevBuf& d, const std::vector<T>& v, size_t alloc_elems) {
  ACC_TRY(d.alloc(std::max<size_t>(alloc_elems, 1) * sizeof(T)));
  ACC_HIP(hipMemset(d.p, 0, std::max<size_t>(alloc_elems, 1) * sizeof(T)));
  if (!v.empty())
    ACC_HIP(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return ACCSPMM_OK;
}

bool monotone(const std::vector<uint32_t>& v, uint32_t last) {
  if (v.empty() || v[0] != 0 || v.back() != last) return false;
  for (size_t i = 1; i < v.size(); ++i)
    if (v[i] < v[i - 1]) return false;
  return true;
}

}  // namespace

extern "C" accspmm_status accspmm_metcf_save(const accspmm_metcf* h, const char* path) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !path) return ACCSPMM_ERR_INVALID_ARG;
    const uint32_t T = h->T, W = h->W, words = h->words;
    // the loader's bound on sparseA2B (8T entries counted in 32 bits)
    if ((uint64_t)T * kColWindow > 0xFFFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    const bool adp = h->grp.p != nullptr;
    std::vector<uint32_t> rwo, tco, a2b, grp, rix, order;
    std::vector<unsigned long long> bits;
    std::vector<float> data;
    const bool reordered = h->row_perm.p != nullptr;
    ACC_HIP(hipSetDevice(h->device));
    ACC_TRY(to_host(rwo, h->rwo, (size_t)W + 1));
    ACC_TRY(to_host(tco, h->tco, (size_t)T + 1));
    ACC_TRY(to_host(bits, h->bits, (size_t)T * words));
    ACC_TRY(to_host(a2b, h->a2b, (size_t)T * kColWindow));
    ACC_TRY(to_host(data, h->data, (size_t)h->nnz));
    if (adp) {
      ACC_TRY(to_host(grp, h->grp, (size_t)h->G + 1));
      ACC_TRY(to_host(rix, h->rix, (size_t)T));
    }
    if (reordered) ACC_TRY(to_host(order, h->row_perm, (size_t)h->rows));
    std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), &std::fclose);
    if (!f) return ACCSPMM_ERR_FILE;
    Writer w;
    w.f = f.get();
    const uint32_t hdr[kHeaderWords] = {
        h->rows, h->cols, h->nnz, h->rw, W, T, h->G, h->ibd, h->balanced,
        h->max_window_tiles, adp ? 1u : 0u, h->opts.row_window, h->opts.col_window,
        h->opts.t_load, h->opts.t_comp, h->opts.t_write, h->opts.target_num,
        h->opts.imbalance_threshold, (uint32_t)h->hist_key.size(), reordered ? 1u : 0u,
        h->col_perm.p ? h->opts.symmetric_order : 0u};
    w.put(kMagic, sizeof(kMagic));
    w.put(&kVersion, 4);
    w.put(hdr, sizeof(hdr));
    w.put(h->hist_key.data(), h->hist_key.size());
    w.put(rwo.data(), rwo.size() * 4);
    w.put(tco.data(), tco.size() * 4);
    w.put(bits.data(), bits.size() * 8);
    w.put(a2b.data(), a2b.size() * 4);
    w.put(data.data(), data.size() * 4);
    if (adp) {
      w.put(grp.data(), grp.size() * 4);
      w.put(rix.data(), rix.size() * 4);
    }
    if (reordered) w.put(order.data(), order.size() * 4);
    const uint64_t sum = w.sum.h;
    if (w.ok) w.ok = std::fwrite(&sum, 1, 8, f.get()) == 8;
    return w.ok ? ACCSPMM_OK : ACCSPMM_ERR_FILE;
  });
}

namespace {

accspmm_status load_impl(const char* path, accspmm_metcf** out) {
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), &std::fclose);
  if (!f) return ACCSPMM_ERR_FILE;
  Reader r;
  r.f = f.get();
  char magic[8];
  uint32_t version = 0, hdr[kHeaderWords] = {};
  r.get(magic, 8);
  r.get(&version, 4);
  if (!r.ok || std::memcmp(magic, kMagic, 8) != 0 || version < 1 || version > kVersion)
    return ACCSPMM_ERR_FILE;
  r.get(hdr, 4 * (version == 1 ? kHeaderWordsV1 : version == 2 ? kHeaderWordsV2 : kHeaderWords));
  const uint32_t rows = hdr[0], cols = hdr[1], nnz = hdr[2], rw = hdr[3], W = hdr[4],
                 T = hdr[5], G = hdr[6], adp = hdr[10], key_len = hdr[18],
                 reorder = hdr[19], sym = hdr[20];
  if (!r.ok || (rw != 8 && rw != 16) || rows == 0 || cols == 0 || nnz > kMaxNnz ||
      W != (rows + rw - 1) / rw || adp > 1 || reorder > 1 || key_len > 4096 ||
      (uint64_t)T * kColWindow > 0xFFFFFFFFull || sym > 2 ||
      (sym && (!reorder || rows != cols)))
    return ACCSPMM_ERR_FILE;
  const uint32_t words = rw / 8;
  // the header fixes the file's length: check it before sizing any array
  // from the header, so a corrupt count cannot allocate gigabytes
  {
    const uint64_t hw = version == 1 ? kHeaderWordsV1 : version == 2 ? kHeaderWordsV2
                                                                     : kHeaderWords;
    const uint64_t expect = 8 + 4 + 4 * hw + key_len + 4 * ((uint64_t)W + 1) +
                            4 * ((uint64_t)T + 1) + 8 * (uint64_t)T * words +
                            4 * (uint64_t)T * kColWindow + 4 * (uint64_t)nnz +
                            (adp ? 4 * ((uint64_t)G + 1) + 4 * (uint64_t)T : 0) +
                            (reorder ? 4 * (uint64_t)rows : 0) + 8;
    const long here = std::ftell(f.get());
    if (here < 0 || std::fseek(f.get(), 0, SEEK_END) != 0) return ACCSPMM_ERR_FILE;
    const long size = std::ftell(f.get());
    if (size < 0 || (uint64_t)size != expect || std::fseek(f.get(), here, SEEK_SET) != 0)
      return ACCSPMM_ERR_FILE;
  }
  std::string key(key_len, '\0');
  r.get(key.data(), key_len);
  std::vector<uint32_t> rwo((size_t)W + 1), tco((size_t)T + 1), a2b((size_t)T * kColWindow);
  std::vector<unsigned long long> bits((size_t)T * words);
  std::vector<float> data(nnz);
  std::vector<uint32_t> grp, rix, order;
  r.get(rwo.data(), rwo.size() * 4);
  r.get(tco.data(), tco.size() * 4);
  r.get(bits.data(), bits.size() * 8);
  r.get(a2b.data(), a2b.size() * 4);
  r.get(data.data(), data.size() * 4);
  if (adp) {
    grp.resize((size_t)G + 1);
    rix.resize(T);
    r.get(grp.data(), grp.size() * 4);
    r.get(rix.data(), rix.size() * 4);
  }
  if (reorder) {
    order.resize(rows);
    r.get(order.data(), order.size() * 4);
  }
  uint64_t sum = 0;
  const uint64_t expect = r.sum.h;
  const bool tail_ok = r.ok && std::fread(&sum, 1, 8, f.get()) == 8;
  char extra;
  if (!tail_ok || sum != expect || std::fread(&extra, 1, 1, f.get()) != 0)
    return ACCSPMM_ERR_FILE;
  // invariants of the format
  if (!monotone(rwo, T) || !monotone(tco, nnz)) return ACCSPMM_ERR_FILE;
  if (adp && !monotone(grp, T)) return ACCSPMM_ERR_FILE;
  if (reorder) {  // a permutation of the rows
    std::vector<uint8_t> seen(rows, 0);
    for (uint32_t x : order) {
      if (x >= rows || seen[x]) return ACCSPMM_ERR_FILE;
      seen[x] = 1;
    }
  }
  for (uint32_t t = 0; t < T; ++t) {
    uint32_t pc = 0;
    for (uint32_t k = 0; k < words; ++k) pc += (uint32_t)__builtin_popcountll(bits[(size_t)t * words + k]);
    if (pc != tco[t + 1] - tco[t]) return ACCSPMM_ERR_FILE;
    for (uint32_t j = 0; j < kColWindow; ++j) {
      const uint32_t c = a2b[(size_t)t * kColWindow + j];
      if (c != kPad && c >= cols) return ACCSPMM_ERR_FILE;
    }
  }
  auto h = std::make_unique<accspmm_metcf>();
  ACC_HIP(hipGetDevice(&h->device));
  h->rows = rows;
  h->cols = cols;
  h->nnz = nnz;
  h->rw = rw;
  h->words = words;
  h->W = W;
  h->T = T;
  h->G = G;
  h->ibd = hdr[7];
  h->balanced = hdr[8];
  h->max_window_tiles = hdr[9];
  h->opts = accspmm_build_opts{hdr[11], hdr[12], hdr[13], hdr[14], hdr[15], hdr[16], hdr[17],
                               adp, sym ? 0u : reorder, sym};
  h->hist_key = key;
  h->h_rwo = rwo;
  // device arrays with the kernels' padding: tcOffset[T+1..T+3] repeat
  // nnz, the all-padding tile T (kPad columns, empty bitmap), a spare empty
  // bitmap and zeros after the values
  std::vector<uint32_t> tco_pad = tco;
  tco_pad.insert(tco_pad.end(), 3, nnz);
  std::vector<uint32_t> a2b_pad = a2b;
  a2b_pad.resize((size_t)(T + 1) * kColWindow, kPad);
  ACC_TRY(to_device(h->rwo, rwo, rwo.size()));
  ACC_TRY(to_device(h->tco, tco_pad, tco_pad.size()));
  ACC_TRY(to_device(h->bits, bits, (size_t)(T + 2) * words));
  ACC_TRY(to_device(h->a2b, a2b_pad, a2b_pad.size()));
  ACC_TRY(to_device(h->data, data, (size_t)nnz + kValuePad));
  if (adp) {
    ACC_TRY(to_device(h->grp, grp, grp.size()));
    ACC_TRY(to_device(h->rix, rix, rix.size()));
  }
  if (reorder) ACC_TRY(to_device(h->row_perm, order, order.size()));
  if (sym) ACC_TRY(to_device(h->col_perm, order, order.size()));
  ACC_HIP(hipDeviceSynchronize());
  ACC_TRY(measure_group_values(h.get(), nullptr));
  *out = h.release();
  return ACCSPMM_OK;
}

}  // namespace

extern "C" accspmm_status accspmm_metcf_load(const char* path, accspmm_metcf** out,
                                             void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!path || !out) return ACCSPMM_ERR_INVALID_ARG;
    *out = nullptr;
    (void)stream;
    try {
      return load_impl(path, out);
    } catch (const std::bad_alloc&) {
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    }
  });
}
