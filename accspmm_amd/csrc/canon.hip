// canon.hip — canonical CSR on the device: every row sorted by column
// (stable), duplicate columns summed left to right in float32, rows
// compacted.  The result equals the host loader's canonicalisation
// (accspmm_read_mtx with canonicalize = 1) bit for bit.
//
// Why it exists: the METCF builder stores a tile's values in CSR order while
// the SpMM kernels index them by bitmap popcount (include/format.hpp:238-269
// against include/mma_tf32.cuh:74-78), so both agree only on rows that are
// column-sorted and duplicate-free.  Graphs that arrive as device arrays
// (a caller's own pipeline) are made canonical here without a host round
// trip.
//
// Steps: validate row_ptr; key every entry by (row, column) — rows take the
// high bits, so one stable LSD radix sort over the row and column bits
// orders each row by column and keeps equal columns in their input order —
// then mark the first entry of each (row, column) run, scan the marks into
// output positions, and let the run heads add their runs in order.
//
// Also here: the row reordering the builder applies on request (rows by
// decreasing length, a stable radix sort of the complemented lengths) and
// the symmetric permutation P A P^T of its locality orderings.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"

using namespace accspmm;

namespace {

constexpr uint32_t kThreads = 256;
enum : uint32_t { kBadRowPtr = 1u, kBadCol = 2u, kUnsorted = 4u };

inline uint32_t bit_width(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

inline unsigned blocks_for(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, (n + kThreads - 1) / kThreads);
}

__global__ __launch_bounds__(kThreads) void k_check_rows(const uint32_t* rp, uint32_t rows,
                                                         uint32_t nnz, uint32_t* flags) {
  const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  bool bad = false;
  if (r <= rows) {
    const uint32_t x = rp[r];
    if (r == 0 && x != 0u) bad = true;
    if (r == rows && x != nnz) bad = true;
    if (r < rows && rp[r + 1] < x) bad = true;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flags, kBadRowPtr);
}

// Largest r in [0, rows) with rp[r] <= i: the row holding entry i when rp is
// monotone (reads stay inside rp[0, rows) whatever rp holds).
__device__ inline uint32_t find_row(const uint32_t* rp, uint32_t rows, uint32_t i) {
  uint32_t lo = 0, hi = rows;
  while (hi - lo > 1u) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (rp[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <class Key>
__global__ __launch_bounds__(kThreads) void k_make_keys(const uint32_t* rp, const uint32_t* col,
                                                        uint32_t rows, uint32_t cols, uint32_t nnz,
                                                        uint32_t col_bits, Key* keys,
                                                        uint32_t* flags) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  bool bad_col = false, unsorted = false;
  if (i < nnz) {
    const uint32_t r = find_row(rp, rows, (uint32_t)i);
    const uint32_t c = col[i];
    bad_col = c >= cols;
    unsorted = i > rp[r] && col[i - 1] >= c;
    keys[i] = ((Key)r << col_bits) | (Key)c;
  }
  const bool any_bad = __syncthreads_or(bad_col);
  const bool any_unsorted = __syncthreads_or(unsorted);
  if (threadIdx.x == 0 && (any_bad || any_unsorted))
    atomicOr(flags, (any_bad ? kBadCol : 0u) | (any_unsorted ? kUnsorted : 0u));
}

// heads[i] = 1 where a (row, column) run starts; heads[nnz] = 0 so that the
// exclusive scan's last element is the new nonzero count.
template <class Key>
__global__ __launch_bounds__(kThreads) void k_heads(const Key* keys, uint32_t nnz,
                                                    uint32_t* heads) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < nnz) heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
  else if (i == nnz) heads[i] = 0u;
}

// Each run head adds its run left to right and writes one entry.
template <class Key>
__global__ __launch_bounds__(kThreads) void k_compact(const Key* keys, const float* vals,
                                                      const uint32_t* heads, const uint32_t* pos,
                                                      uint32_t nnz, Key col_mask, uint32_t* col,
                                                      float* val) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nnz || !heads[i]) return;
  float s = vals[i];
  for (uint64_t j = i + 1; j < nnz && !heads[j]; ++j) s += vals[j];
  const uint32_t d = pos[i];
  col[d] = (uint32_t)(keys[i] & col_mask);
  val[d] = s;
}

// Rows keep their places in the sorted order, so a row's new start is the
// number of run heads before its old start.
__global__ __launch_bounds__(kThreads) void k_remap_rows(uint32_t* rp, uint32_t rows,
                                                         const uint32_t* pos) {
  const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (r <= rows) rp[r] = pos[rp[r]];
}

template <class Key>
accspmm_status sort_and_compact(uint32_t rows, uint32_t nnz, uint32_t col_bits,
                                uint32_t key_bits, uint32_t* rp, uint32_t* col, float* val,
                                const DevBuf& keys_in, uint32_t* nnz_out, hipStream_t s) {
  DevBuf keys_out, vals_out, heads, pos, temp;
  ACC_TRY(keys_out.alloc(sizeof(Key) * (size_t)nnz));
  ACC_TRY(vals_out.alloc(sizeof(float) * (size_t)nnz));
  size_t temp_bytes = 0;
  ACC_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_in.as<Key>(), keys_out.as<Key>(),
                                    val, vals_out.as<float>(), nnz, 0u, key_bits, s));
  ACC_TRY(temp.alloc(temp_bytes));
  ACC_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys_in.as<Key>(), keys_out.as<Key>(),
                                    val, vals_out.as<float>(), nnz, 0u, key_bits, s));
  ACC_TRY(heads.alloc(sizeof(uint32_t) * ((size_t)nnz + 1)));
  ACC_TRY(pos.alloc(sizeof(uint32_t) * ((size_t)nnz + 1)));
  k_heads<Key><<<blocks_for((uint64_t)nnz + 1), kThreads, 0, s>>>(keys_out.as<Key>(), nnz,
                                                                   heads.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(heads.as<uint32_t>(), pos.as<uint32_t>(), (uint64_t)nnz + 1, s));
  const Key col_mask = col_bits ? (((Key)1 << col_bits) - 1) : (Key)0;
  k_compact<Key><<<blocks_for(nnz), kThreads, 0, s>>>(
      keys_out.as<Key>(), vals_out.as<float>(), heads.as<uint32_t>(), pos.as<uint32_t>(), nnz,
      col_mask, col, val);
  ACC_HIP(hipGetLastError());
  k_remap_rows<<<blocks_for((uint64_t)rows + 1), kThreads, 0, s>>>(rp, rows,
                                                                    pos.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_HIP(hipMemcpyAsync(nnz_out, pos.as<uint32_t>() + nnz, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}

template <class Key>
accspmm_status canonicalize(uint32_t rows, uint32_t cols, uint32_t nnz, uint32_t* rp,
                            uint32_t* col, float* val, uint32_t* nnz_out, hipStream_t s) {
  const uint32_t col_bits = bit_width(cols - 1u);
  const uint32_t key_bits = std::max(1u, col_bits + bit_width(rows - 1u));
  DevBuf flags, keys;
  ACC_TRY(flags.alloc(sizeof(uint32_t)));
  ACC_HIP(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), s));
  k_check_rows<<<blocks_for((uint64_t)rows + 1), kThreads, 0, s>>>(rp, rows, nnz,
                                                                    flags.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(keys.alloc(sizeof(Key) * (size_t)std::max(nnz, 1u)));
  if (nnz) {
    k_make_keys<Key><<<blocks_for(nnz), kThreads, 0, s>>>(rp, col, rows, cols, nnz, col_bits,
                                                          keys.as<Key>(), flags.as<uint32_t>());
    ACC_HIP(hipGetLastError());
  }
  uint32_t h_flags = 0;
  ACC_HIP(hipMemcpyAsync(&h_flags, flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  if (h_flags & (kBadRowPtr | kBadCol)) return ACCSPMM_ERR_INVALID_ARG;
  if (!(h_flags & kUnsorted)) {
    *nnz_out = nnz;
    return ACCSPMM_OK;
  }
  return sort_and_compact<Key>(rows, nnz, col_bits, key_bits, rp, col, val, keys, nnz_out, s);
}

// keys[r] = ~length(r) (ascending order = decreasing length), ids[r] = r
__global__ __launch_bounds__(kThreads) void k_length_keys(const uint32_t* rp, uint32_t rows,
                                                          uint32_t* keys, uint32_t* ids) {
  const uint64_t r = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (r < rows) {
    keys[r] = ~(rp[r + 1] - rp[r]);
    ids[r] = (uint32_t)r;
  }
}

// len[i] = length of input row perm[i]; len[rows] = 0 (scanned into row_ptr)
__global__ __launch_bounds__(kThreads) void k_perm_lengths(const uint32_t* rp,
                                                           const uint32_t* perm, uint32_t rows,
                                                           uint32_t* len) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < rows) len[i] = rp[perm[i] + 1] - rp[perm[i]];
  else if (i == rows) len[i] = 0u;
}

// One wave per output row copies the input row's columns and values.
__global__ __launch_bounds__(kThreads) void k_copy_rows(const uint32_t* rp, const uint32_t* ci,
                                                        const float* va, const uint32_t* perm,
                                                        const uint32_t* nrp, uint32_t rows,
                                                        uint32_t* nci, float* nva) {
  const uint64_t i = (uint64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
  if (i >= rows) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t src = rp[perm[i]], dst = nrp[i], len = nrp[i + 1] - dst;
  for (uint32_t j = lane; j < len; j += 64u) {
    nci[dst + j] = ci[src + j];
    nva[dst + j] = va[src + j];
  }
}

// flags kBadCol / kUnsorted for a column >= cols or a row that is not
// strictly increasing (row_ptr already checked)
__global__ __launch_bounds__(kThreads) void k_check_canonical(const uint32_t* rp,
                                                              const uint32_t* col, uint32_t rows,
                                                              uint32_t cols, uint32_t nnz,
                                                              uint32_t* flags) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  bool bad_col = false, unsorted = false;
  if (i < nnz) {
    const uint32_t r = find_row(rp, rows, (uint32_t)i);
    const uint32_t c = col[i];
    bad_col = c >= cols;
    unsorted = i > rp[r] && col[i - 1] >= c;
  }
  const bool any_bad = __syncthreads_or(bad_col);
  const bool any_unsorted = __syncthreads_or(unsorted);
  if (threadIdx.x == 0 && (any_bad || any_unsorted))
    atomicOr(flags, (any_bad ? kBadCol : 0u) | (any_unsorted ? kUnsorted : 0u));
}

// inv[perm[i]] = i
__global__ __launch_bounds__(kThreads) void k_invert(const uint32_t* perm, uint32_t n,
                                                     uint32_t* inv) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) inv[perm[i]] = (uint32_t)i;
}

// One wave per output row: input row perm[i] with its columns relabelled
// through inv (the row comes out unsorted).
__global__ __launch_bounds__(kThreads) void k_copy_rows_relabel(
    const uint32_t* rp, const uint32_t* ci, const float* va, const uint32_t* perm,
    const uint32_t* inv, const uint32_t* nrp, uint32_t rows, uint32_t* nci, float* nva) {
  const uint64_t i = (uint64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
  if (i >= rows) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t src = rp[perm[i]], dst = nrp[i], len = nrp[i + 1] - dst;
  for (uint32_t j = lane; j < len; j += 64u) {
    nci[dst + j] = inv[ci[src + j]];
    nva[dst + j] = va[src + j];
  }
}

// perm = rows by decreasing length (stable); perm must hold `rows` words
accspmm_status length_order(const uint32_t* row_ptr, uint32_t rows, DevBuf& perm,
                            hipStream_t s) {
  DevBuf keys, keys_out, ids, temp;
  ACC_TRY(keys.alloc(sizeof(uint32_t) * (size_t)rows));
  ACC_TRY(keys_out.alloc(sizeof(uint32_t) * (size_t)rows));
  ACC_TRY(ids.alloc(sizeof(uint32_t) * (size_t)rows));
  k_length_keys<<<blocks_for(rows), kThreads, 0, s>>>(row_ptr, rows, keys.as<uint32_t>(),
                                                      ids.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  size_t temp_bytes = 0;
  ACC_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys.as<uint32_t>(),
                                    keys_out.as<uint32_t>(), ids.as<uint32_t>(),
                                    perm.as<uint32_t>(), rows, 0u, 32u, s));
  ACC_TRY(temp.alloc(temp_bytes));
  ACC_HIP(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys.as<uint32_t>(),
                                    keys_out.as<uint32_t>(), ids.as<uint32_t>(),
                                    perm.as<uint32_t>(), rows, 0u, 32u, s));
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}

}  // namespace

namespace accspmm {

accspmm_status symmetric_permute(const accspmm_csr_view* in, uint32_t order, DevBuf& rp,
                                 DevBuf& ci, DevBuf& va, DevBuf& perm, hipStream_t s) {
  const uint32_t n = in->rows, nnz = in->nnz;
  if (in->rows != in->cols || (order != 1 && order != 2)) return ACCSPMM_ERR_UNSUPPORTED;
  // the input must be canonical, as for any build (the relabelled rows are
  // sorted below, which would otherwise hide duplicates or disorder)
  DevBuf flags;
  ACC_TRY(flags.alloc(sizeof(uint32_t)));
  ACC_HIP(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), s));
  k_check_rows<<<blocks_for((uint64_t)n + 1), kThreads, 0, s>>>(in->row_ptr, n, nnz,
                                                                 flags.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  uint32_t h_flags = 0;
  ACC_HIP(hipMemcpyAsync(&h_flags, flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  if (h_flags) return ACCSPMM_ERR_INVALID_ARG;
  if (nnz) {
    k_check_canonical<<<blocks_for(nnz), kThreads, 0, s>>>(in->row_ptr, in->col_idx, n, n, nnz,
                                                           flags.as<uint32_t>());
    ACC_HIP(hipGetLastError());
    ACC_HIP(hipMemcpyAsync(&h_flags, flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ACC_HIP(hipStreamSynchronize(s));
    if (h_flags & kBadCol) return ACCSPMM_ERR_INVALID_ARG;
    if (h_flags & kUnsorted) return ACCSPMM_ERR_NOT_CANONICAL;
  }
  ACC_TRY(perm.alloc(sizeof(uint32_t) * (size_t)n));
  if (order == 2) {
    ACC_TRY(length_order(in->row_ptr, n, perm, s));
  } else {
    // reverse Cuthill-McKee on the host over the pattern
    std::vector<uint32_t> hrp((size_t)n + 1), hci(nnz), hp(n);
    ACC_HIP(hipMemcpyAsync(hrp.data(), in->row_ptr, hrp.size() * 4, hipMemcpyDeviceToHost, s));
    if (nnz)
      ACC_HIP(hipMemcpyAsync(hci.data(), in->col_idx, (size_t)nnz * 4, hipMemcpyDeviceToHost, s));
    ACC_HIP(hipStreamSynchronize(s));
    ACC_TRY(accspmm_order_rcm(n, hrp.data(), hci.data(), hp.data()));
    ACC_HIP(hipMemcpyAsync(perm.p, hp.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    ACC_HIP(hipStreamSynchronize(s));
  }
  DevBuf inv, len;
  ACC_TRY(inv.alloc(sizeof(uint32_t) * (size_t)n));
  k_invert<<<blocks_for(n), kThreads, 0, s>>>(perm.as<uint32_t>(), n, inv.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(len.alloc(sizeof(uint32_t) * ((size_t)n + 1)));
  ACC_TRY(rp.alloc(sizeof(uint32_t) * ((size_t)n + 1)));
  k_perm_lengths<<<blocks_for((uint64_t)n + 1), kThreads, 0, s>>>(
      in->row_ptr, perm.as<uint32_t>(), n, len.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(len.as<uint32_t>(), rp.as<uint32_t>(), (uint64_t)n + 1, s));
  ACC_TRY(ci.alloc(sizeof(uint32_t) * (size_t)std::max(nnz, 1u)));
  ACC_TRY(va.alloc(sizeof(float) * (size_t)std::max(nnz, 1u)));
  const uint64_t waves_per_block = kThreads / 64;
  k_copy_rows_relabel<<<(unsigned)std::max<uint64_t>(1, (n + waves_per_block - 1) /
                                                            waves_per_block),
                        kThreads, 0, s>>>(in->row_ptr, in->col_idx, in->val,
                                          perm.as<uint32_t>(), inv.as<uint32_t>(),
                                          rp.as<uint32_t>(), n, ci.as<uint32_t>(),
                                          va.as<float>());
  ACC_HIP(hipGetLastError());
  // sort each relabelled row (a permutation creates no duplicates, so the
  // count stays nnz)
  uint32_t nnz_out = 0;
  ACC_TRY(accspmm_csr_canonicalize(n, n, nnz, rp.as<uint32_t>(), ci.as<uint32_t>(),
                                   va.as<float>(), &nnz_out, s));
  if (nnz_out != nnz) return ACCSPMM_ERR_INVALID_ARG;
  return ACCSPMM_OK;
}

accspmm_status reorder_rows_by_length(const accspmm_csr_view* in, DevBuf& rp, DevBuf& ci,
                                      DevBuf& va, DevBuf& perm, hipStream_t s) {
  const uint32_t rows = in->rows, nnz = in->nnz;
  // the copies below index by row_ptr: it must run from 0 to nnz, in order
  DevBuf flags, len;
  ACC_TRY(flags.alloc(sizeof(uint32_t)));
  ACC_HIP(hipMemsetAsync(flags.p, 0, sizeof(uint32_t), s));
  k_check_rows<<<blocks_for((uint64_t)rows + 1), kThreads, 0, s>>>(in->row_ptr, rows, nnz,
                                                                    flags.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  uint32_t h_flags = 0;
  ACC_HIP(hipMemcpyAsync(&h_flags, flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  if (h_flags) return ACCSPMM_ERR_INVALID_ARG;

  ACC_TRY(perm.alloc(sizeof(uint32_t) * (size_t)rows));
  ACC_TRY(length_order(in->row_ptr, rows, perm, s));
  ACC_TRY(len.alloc(sizeof(uint32_t) * ((size_t)rows + 1)));
  ACC_TRY(rp.alloc(sizeof(uint32_t) * ((size_t)rows + 1)));
  k_perm_lengths<<<blocks_for((uint64_t)rows + 1), kThreads, 0, s>>>(
      in->row_ptr, perm.as<uint32_t>(), rows, len.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(len.as<uint32_t>(), rp.as<uint32_t>(), (uint64_t)rows + 1, s));
  ACC_TRY(ci.alloc(sizeof(uint32_t) * (size_t)std::max(nnz, 1u)));
  ACC_TRY(va.alloc(sizeof(float) * (size_t)std::max(nnz, 1u)));
  const uint64_t waves_per_block = kThreads / 64;
  k_copy_rows<<<(unsigned)std::max<uint64_t>(1, (rows + waves_per_block - 1) / waves_per_block),
                kThreads, 0, s>>>(in->row_ptr, in->col_idx, in->val, perm.as<uint32_t>(),
                                  rp.as<uint32_t>(), rows, ci.as<uint32_t>(), va.as<float>());
  ACC_HIP(hipGetLastError());
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}

}  // namespace accspmm

extern "C" accspmm_status accspmm_csr_canonicalize(uint32_t rows, uint32_t cols, uint32_t nnz,
                                                   uint32_t* row_ptr, uint32_t* col_idx,
                                                   float* val, uint32_t* nnz_out,
                                                   void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!row_ptr || !nnz_out || rows == 0 || cols == 0 || (nnz && (!col_idx || !val)))
      return ACCSPMM_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t bits = bit_width(cols - 1u) + bit_width(rows - 1u);
    if (bits <= 32u) return canonicalize<uint32_t>(rows, cols, nnz, row_ptr, col_idx, val, nnz_out, s);
    return canonicalize<uint64_t>(rows, cols, nnz, row_ptr, col_idx, val, nnz_out, s);
  });
}
