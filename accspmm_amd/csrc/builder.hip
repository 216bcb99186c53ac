// builder.hip — GPU construction of the METCF bitmap-tile format, its AdpBME
// grouping and the imbalance test, bit-exact with the reference's CPU
// builders:
//   METCFBit::convertFromCSR   include/format.hpp:215-274
//   TensorCoreFormat::extractUniqueColumns / padListLength  :179-199, :23-30
//   AdpBME::CSR2AdpBME         include/format.hpp:340-429
//   balanceStrategy            include/utils.hpp:75-85
//
// Pipeline (one thread per nonzero unless noted; canonical CSR input):
//   1. validate        rows strictly increasing in column, columns < cols
//   2. merge_rank      rank of each nonzero in its window's merged column
//                      list = own offset + per-other-row binary-search counts
//                      (ties broken by row), scattered to `merged`
//   3. first_flags     merged[p] != merged[p-1]  -> flag (window-local)
//   4. scan            inclusive scan of flags = condensed column index + 1
//   5. window_tiles    unique columns per window -> tiles -> rowWindowOffset
//   6. place           sparseA2B entries, tile bitmaps (atomic OR)
//   7. tile counts     popcount per tile -> exclusive scan -> tcOffset
//   8. values          data[tcOffset[t] + rank of the bit in the tile]
// For canonical rows the reference's CSR-order placement of values inside a
// tile equals bit order, so step 8 reproduces its `data` exactly.
#include <algorithm>
#include <cstring>
#include <string>

#include "internal.hpp"

namespace accspmm {
namespace {

constexpr int kThreads = 256;

inline uint32_t blocks_for(uint64_t n, uint32_t per = kThreads) {
  uint64_t b = (n + per - 1) / per;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 30));
}

// first index i in [0, n) with a[i] > key (n if none)
__device__ inline uint32_t ub(const uint32_t* __restrict__ a, uint32_t n,
                              uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ inline uint32_t lb(const uint32_t* __restrict__ a, uint32_t n,
                              uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// row containing nonzero e: last r with row_ptr[r] <= e
__device__ inline uint32_t row_of(const uint32_t* __restrict__ row_ptr,
                                  uint32_t rows, uint32_t e) {
  return ub(row_ptr, rows + 1, e) - 1;
}

__global__ void k_validate_rows(const uint32_t* __restrict__ row_ptr,
                                uint32_t rows, uint32_t nnz,
                                uint32_t* __restrict__ err) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = row_ptr[r], b = row_ptr[r + 1];
    if (a > b || b > nnz) atomicOr(err, 1u);
  }
}

__global__ void k_validate_cols(const uint32_t* __restrict__ row_ptr,
                                const uint32_t* __restrict__ col,
                                uint32_t rows, uint32_t cols, uint32_t nnz,
                                uint32_t* __restrict__ err) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t r = row_of(row_ptr, rows, (uint32_t)e);
    uint32_t c = col[e];
    bool bad = c >= cols;
    if (e > row_ptr[r] && col[e - 1] >= c) bad = true;
    if (bad) atomicOr(err, 2u);
  }
}

__global__ void k_merge_rank(const uint32_t* __restrict__ row_ptr,
                             const uint32_t* __restrict__ col, uint32_t rows,
                             uint32_t nnz, uint32_t rw,
                             uint32_t* __restrict__ pos,
                             uint32_t* __restrict__ merged) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = row_of(row_ptr, rows, (uint32_t)e);
    const uint32_t w0 = r - r % rw;
    const uint32_t w1 = min(w0 + rw, rows);
    const uint32_t c = col[e];
    uint32_t rank = (uint32_t)e - row_ptr[r];
    for (uint32_t q = w0; q < w1; ++q) {
      if (q == r) continue;
      const uint32_t b = row_ptr[q], n = row_ptr[q + 1] - b;
      rank += q < r ? ub(col + b, n, c) : lb(col + b, n, c);
    }
    pos[e] = rank;
    merged[row_ptr[w0] + rank] = c;
  }
}

__global__ void k_first_flags(const uint32_t* __restrict__ row_ptr,
                              uint32_t rows, uint32_t nnz, uint32_t rw,
                              const uint32_t* __restrict__ pos,
                              const uint32_t* __restrict__ merged,
                              uint32_t* __restrict__ flag) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = row_of(row_ptr, rows, (uint32_t)e);
    const uint32_t base = row_ptr[r - r % rw];
    const uint32_t p = base + pos[e];
    flag[p] = (pos[e] == 0 || merged[p] != merged[p - 1]) ? 1u : 0u;
  }
}

// per window: number of tiles = ceil(unique columns / 8)
__global__ void k_window_tiles(const uint32_t* __restrict__ row_ptr,
                               uint32_t rows, uint32_t W, uint32_t rw,
                               const uint32_t* __restrict__ incl,
                               uint32_t* __restrict__ tiles) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t b = row_ptr[w * rw], e = row_ptr[min((uint32_t)(w * rw + rw), rows)];
    uint32_t u = 0;
    if (e > b) u = incl[e - 1] - (b > 0 ? incl[b - 1] : 0u);
    tiles[w] = (u + kColWindow - 1) / kColWindow;
  }
}

// sparseA2B entries + bitmap bits.  Afterwards `pos[e]` holds the bit index
// inside the tile and `tile_of[e]` the tile (the merged buffer is reused).
__global__ void k_place(const uint32_t* __restrict__ row_ptr,
                        const uint32_t* __restrict__ col, uint32_t rows,
                        uint32_t nnz, uint32_t rw,
                        const uint32_t* __restrict__ rwo,
                        const uint32_t* __restrict__ flag,
                        const uint32_t* __restrict__ incl,
                        uint32_t* __restrict__ pos,
                        uint32_t* __restrict__ tile_of,
                        uint32_t* __restrict__ a2b,
                        unsigned long long* __restrict__ bits,
                        uint32_t words) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = row_of(row_ptr, rows, (uint32_t)e);
    const uint32_t w = r / rw;
    const uint32_t base = row_ptr[w * rw];
    const uint32_t p = base + pos[e];
    const uint32_t gbase = base > 0 ? incl[base - 1] : 0u;
    const uint32_t ci = incl[p] - gbase - 1;
    const uint32_t t = rwo[w] + ci / kColWindow;
    if (flag[p]) a2b[(uint64_t)rwo[w] * kColWindow + ci] = col[e];
    const uint32_t local = (r - w * rw) * kColWindow + ci % kColWindow;
    atomicOr(&bits[(uint64_t)t * words + local / 64], 1ull << (local % 64));
    pos[e] = local;
    tile_of[e] = t;
  }
}

__global__ void k_tile_popcount(const unsigned long long* __restrict__ bits,
                                uint32_t T, uint32_t words,
                                uint32_t* __restrict__ cnt) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
    for (uint32_t k = 0; k < words; ++k) c += __popcll(bits[t * words + k]);
    cnt[t] = c;
  }
}

__global__ void k_values(uint32_t nnz, const float* __restrict__ val,
                         const uint32_t* __restrict__ pos,
                         const uint32_t* __restrict__ tile_of,
                         const unsigned long long* __restrict__ bits,
                         uint32_t words, const uint32_t* __restrict__ tco,
                         float* __restrict__ data) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = tile_of[e], local = pos[e];
    uint32_t rank = 0;
    for (uint32_t k = 0; k < local / 64; ++k) rank += __popcll(bits[(uint64_t)t * words + k]);
    const unsigned long long wd = bits[(uint64_t)t * words + local / 64];
    const uint32_t sh = local % 64;
    rank += sh ? __popcll(wd & ((1ull << sh) - 1ull)) : 0u;
    data[tco[t] + rank] = val[e];
  }
}

// balanceStrategy (include/utils.hpp:75-85): Σ over windows 0..W-2 of
// |int(u32(T_w - avg))|, summed modulo 2^32.
__global__ void k_ibd_sum(const uint32_t* __restrict__ rwo, uint32_t W,
                          uint32_t avg, unsigned long long* __restrict__ sum,
                          uint32_t* __restrict__ max_tiles) {
  unsigned long long acc = 0;
  uint32_t mx = 0;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t cur = rwo[w + 1] - rwo[w];
    mx = max(mx, cur);
    if (w + 1 < W) {
      int32_t d = (int32_t)(uint32_t)(cur - avg);
      acc += (uint32_t)(d < 0 ? -d : d);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc += __shfl_down(acc, off, 64);
    mx = max(mx, (uint32_t)__shfl_down(mx, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(sum, acc);
    atomicMax(max_tiles, mx);
  }
}

// Values of the fullest 4-tile group: groups start at each window's first
// tile and every 4 tiles after it (the SpMM kernels' groups; work-unit
// pieces start on them too).  One wave per window, lanes over its groups.
__global__ void k_group_values(const uint32_t* __restrict__ rwo,
                               const uint32_t* __restrict__ tco, uint32_t W,
                               uint32_t* __restrict__ max_values) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  uint32_t mx = 0;
  for (uint64_t w = blockIdx.x * (uint64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); w < W;
       w += waves) {
    const uint32_t tb = rwo[w], te = rwo[w + 1];
    for (uint64_t t = tb + 4ull * lane; t < te; t += 256) {
      const uint32_t e = (uint32_t)min<uint64_t>(t + 4, te);
      mx = max(mx, tco[e] - tco[t]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_down(mx, off, 64));
  if (lane == 0) atomicMax(max_values, mx);
}

// Windows whose condensed columns span less than a 1/8 of B's rows, and
// non-empty windows (counts[0], counts[1]).  A window's tiles are sorted by
// column and padding sits at the end of its last tile, so the span is the
// last tile's last real column minus the first tile's first.
__global__ void k_window_spans(const uint32_t* __restrict__ rwo,
                               const uint32_t* __restrict__ a2b, uint32_t W, uint32_t cols,
                               uint32_t* __restrict__ counts) {
  uint32_t narrow = 0, used = 0;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t tb = rwo[w], te = rwo[w + 1];
    if (tb == te) continue;
    const uint32_t lo = a2b[(uint64_t)tb * kColWindow];
    uint32_t hi = lo;
    for (int j = kColWindow - 1; j >= 0; --j) {
      const uint32_t c = a2b[(uint64_t)(te - 1) * kColWindow + j];
      if (c != kPad) {
        hi = c;
        break;
      }
    }
    ++used;
    narrow += (uint64_t)(hi - lo) * 8 < cols ? 1u : 0u;
  }
  if (used) atomicAdd(&counts[1], used);
  if (narrow) atomicAdd(&counts[0], narrow);
}

// ---- AdpBME grouping (include/format.hpp:346,366-382,418-423) ----------
// Tile t costs cf = t_load+t_comp+t_write when it is the first tile of its
// window, else co = t_load+t_comp.  A group starting at tile i always takes
// i and then extends while the running cost stays <= target.

__global__ void k_tile_window(const uint32_t* __restrict__ rwo, uint32_t W,
                              uint32_t T, uint32_t rw,
                              uint32_t* __restrict__ rix,
                              uint8_t* __restrict__ first) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t w = ub(rwo, W + 1, (uint32_t)t) - 1;
    rix[t] = w * rw;
    first[t] = rwo[w] == t;
  }
}

__global__ void k_group_next(const uint8_t* __restrict__ first, uint32_t T,
                             uint32_t cf, uint32_t co, uint32_t target,
                             uint32_t* __restrict__ next) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < T;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t acc = first[i] ? cf : co;
    uint64_t j = i + 1;
    while (j < T) {
      uint32_t c = first[j] ? cf : co;
      if (acc + c > target) break;
      acc += c;
      ++j;
    }
    next[i] = (uint32_t)j;
  }
}

// For chunk k (tiles [kC, (k+1)C)) and entry offset e in [0, L): follow the
// chain from kC+e until it leaves the chunk; record the exit offset into the
// next chunk and the number of group starts visited.
__global__ void k_chunk_paths(const uint32_t* __restrict__ next, uint32_t T,
                              uint32_t C, uint32_t L, uint32_t nchunks,
                              uint32_t* __restrict__ exit_off,
                              uint32_t* __restrict__ count) {
  uint64_t id = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (id >= (uint64_t)nchunks * L) return;
  uint32_t k = (uint32_t)(id / L), e = (uint32_t)(id % L);
  uint64_t s = (uint64_t)k * C + e;
  uint64_t end = min((uint64_t)(k + 1) * C, (uint64_t)T);
  uint32_t n = 0;
  while (s < end) {
    s = next[s];
    ++n;
  }
  exit_off[id] = (uint32_t)(s >= (uint64_t)(k + 1) * C ? s - (uint64_t)(k + 1) * C : 0);
  count[id] = n;
}

// Sequential composition across chunks (one thread): entry and group base of
// every chunk.
__global__ void k_chunk_compose(const uint32_t* __restrict__ exit_off,
                                const uint32_t* __restrict__ count,
                                uint32_t nchunks, uint32_t L,
                                uint32_t* __restrict__ entry,
                                uint32_t* __restrict__ gbase) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint32_t e = 0, g = 0;
  for (uint32_t k = 0; k < nchunks; ++k) {
    entry[k] = e;
    gbase[k] = g;
    uint32_t id = k * L + e;
    g += count[id];
    e = exit_off[id];
  }
  gbase[nchunks] = g;
}

__global__ void k_group_write(const uint32_t* __restrict__ next, uint32_t T,
                              uint32_t C, uint32_t nchunks,
                              const uint32_t* __restrict__ entry,
                              const uint32_t* __restrict__ gbase,
                              uint32_t* __restrict__ grp) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nchunks) return;
  uint64_t s = (uint64_t)k * C + entry[k];
  uint64_t end = min((uint64_t)(k + 1) * C, (uint64_t)T);
  uint32_t g = gbase[k];
  while (s < end) {
    grp[g++] = (uint32_t)s;
    s = next[s];
  }
  if (k == nchunks - 1) grp[g] = T;
}

}  // namespace

// h->max_group_values from the device rowWindowOffset and tcOffset
accspmm_status measure_group_values(accspmm_metcf* h, hipStream_t s) {
  h->max_group_values = 0;
  h->banded = false;
  if (h->W == 0 || h->T == 0) return ACCSPMM_OK;
  DevBuf mx;
  ACC_TRY(mx.alloc(4));
  ACC_HIP(hipMemsetAsync(mx.p, 0, 4, s));
  const uint32_t blocks = std::min<uint32_t>((h->W + 3) / 4, 4096);
  hipLaunchKernelGGL(k_group_values, dim3(blocks), dim3(256), 0, s, h->rwo.as<uint32_t>(),
                     h->tco.as<uint32_t>(), h->W, mx.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  uint32_t v = 0;
  ACC_HIP(hipMemcpyAsync(&v, mx.p, 4, hipMemcpyDeviceToHost, s));
  // banded: most non-empty windows narrow (the work-unit mapping keeps
  // neighbouring windows on one XCD for such graphs)
  DevBuf cnt;
  ACC_TRY(cnt.alloc(8));
  ACC_HIP(hipMemsetAsync(cnt.p, 0, 8, s));
  hipLaunchKernelGGL(k_window_spans, dim3(std::min<uint32_t>((h->W + 255) / 256, 1024)),
                     dim3(256), 0, s, h->rwo.as<uint32_t>(), h->a2b.as<uint32_t>(), h->W,
                     h->cols, cnt.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  uint32_t counts[2] = {0, 0};
  ACC_HIP(hipMemcpyAsync(counts, cnt.p, 8, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  h->max_group_values = v;
  h->banded = counts[1] && 2ull * counts[0] > counts[1];
  return ACCSPMM_OK;
}
}  // namespace accspmm

using namespace accspmm;

// mean row length from which the automatic build choices reorder rows and
// use 16-row windows
constexpr uint32_t kLongRows = 128;

extern "C" void accspmm_build_opts_default(accspmm_build_opts* o) {
  if (!o) return;
  o->row_window = 8;
  o->col_window = 8;
  o->t_load = 3;
  o->t_comp = 1;
  o->t_write = 2;
  o->target_num = 16;
  o->imbalance_threshold = 8;
  // the SpMM kernels run work units, not AdpBME groups: groups are built
  // only for callers that export them (the AdpBME drop-in class)
  o->build_adpbme = 0;
  o->reorder_rows = 0;
  o->symmetric_order = 0;
}

namespace {

accspmm_status build_impl(const accspmm_csr_view* csr,
                          const accspmm_build_opts& o, accspmm_metcf* h,
                          hipStream_t s) {
  const uint32_t rows = csr->rows, nnz = csr->nnz, rw = o.row_window;
  h->rows = rows;
  h->cols = csr->cols;
  h->nnz = nnz;
  h->rw = rw;
  h->words = rw / 8;
  h->opts = o;
  ACC_HIP(hipGetDevice(&h->device));
  const uint32_t W = (rows + rw - 1) / rw;
  h->W = W;
  const uint32_t* rp = csr->row_ptr;
  const uint32_t* ci = csr->col_idx;

  // 1. validation
  uint32_t ends[2];
  ACC_HIP(hipMemcpyAsync(&ends[0], rp, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipMemcpyAsync(&ends[1], rp + rows, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  if (ends[0] != 0 || ends[1] != nnz) return ACCSPMM_ERR_INVALID_ARG;
  DevBuf err;
  ACC_TRY(err.alloc(4));
  ACC_HIP(hipMemsetAsync(err.p, 0, 4, s));
  hipLaunchKernelGGL(k_validate_rows, dim3(blocks_for(rows)), dim3(kThreads), 0,
                     s, rp, rows, nnz, err.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  uint32_t herr = 0;
  ACC_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  if (herr) return ACCSPMM_ERR_INVALID_ARG;
  if (nnz) {
    hipLaunchKernelGGL(k_validate_cols, dim3(blocks_for(nnz)), dim3(kThreads),
                       0, s, rp, ci, rows, csr->cols, nnz, err.as<uint32_t>());
    ACC_HIP(hipGetLastError());
    ACC_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    ACC_HIP(hipStreamSynchronize(s));
    if (herr) return ACCSPMM_ERR_NOT_CANONICAL;
  }

  // 2-5. ranks, flags, condensed indices, tiles per window
  DevBuf pos, merged, incl, wt;
  ACC_TRY(pos.alloc((uint64_t)nnz * 4));
  ACC_TRY(merged.alloc((uint64_t)nnz * 4));
  ACC_TRY(incl.alloc((uint64_t)nnz * 4));
  ACC_TRY(wt.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(h->rwo.alloc(((uint64_t)W + 1) * 4));
  if (nnz) {
    hipLaunchKernelGGL(k_merge_rank, dim3(blocks_for(nnz)), dim3(kThreads), 0,
                       s, rp, ci, rows, nnz, rw, pos.as<uint32_t>(),
                       merged.as<uint32_t>());
    ACC_HIP(hipGetLastError());
    // flags go to `incl`, then are scanned in place
    hipLaunchKernelGGL(k_first_flags, dim3(blocks_for(nnz)), dim3(kThreads), 0,
                       s, rp, rows, nnz, rw, pos.as<uint32_t>(),
                       merged.as<uint32_t>(), incl.as<uint32_t>());
    ACC_HIP(hipGetLastError());
  }
  // keep flags: copy to `merged` (merged no longer needed after flags)
  if (nnz)
    ACC_HIP(hipMemcpyAsync(merged.p, incl.p, (uint64_t)nnz * 4,
                           hipMemcpyDeviceToDevice, s));
  ACC_TRY(inclusive_scan_u32(incl.as<uint32_t>(), incl.as<uint32_t>(), nnz, s));
  ACC_HIP(hipMemsetAsync(wt.p, 0, ((uint64_t)W + 1) * 4, s));
  hipLaunchKernelGGL(k_window_tiles, dim3(blocks_for(W)), dim3(kThreads), 0, s,
                     rp, rows, W, rw, incl.as<uint32_t>(), wt.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(wt.as<uint32_t>(), h->rwo.as<uint32_t>(),
                             (uint64_t)W + 1, s));
  h->h_rwo.resize((size_t)W + 1);
  ACC_HIP(hipMemcpyAsync(h->h_rwo.data(), h->rwo.p, ((uint64_t)W + 1) * 4,
                         hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  const uint32_t T = h->h_rwo[W];
  h->T = T;
  {
    // tuning key: log2 buckets of tiles per window as 10% steps, plus the
    // log2 of the window count
    uint32_t buckets[24] = {0};
    for (uint32_t w = 0; w < W; ++w) {
      uint32_t t = h->h_rwo[w + 1] - h->h_rwo[w], b = 0;
      while (b < 23 && (1u << b) <= t) ++b;
      buckets[b]++;
    }
    uint32_t lw = 0;
    while ((2ull << lw) <= W) ++lw;
    std::string key = "rw" + std::to_string(rw) + "-w" + std::to_string(lw) + "-h";
    for (int b = 0; b < 24; ++b) {
      uint32_t pct = W ? (uint32_t)((1000ull * buckets[b] / W + 50) / 100) : 0;
      key += (b ? "." : "") + std::to_string(pct);
    }
    h->hist_key = key;
  }

  // 6. sparseA2B + bitmaps
  const uint32_t words = h->words;
  // one extra all-padding tile T (columns kPad, empty bitmap, tcOffset[T] =
  // nnz): the SpMM kernels read it for group slots past a range's last tile
  ACC_TRY(h->a2b.alloc(((uint64_t)T + 1) * kColWindow * 4));
  // a spare empty tile and three spare tcOffset entries (= nnz): the
  // kernels read 16 bytes of bitmaps (8-row windows: tiles t and t+1) and
  // of tcOffset (t..t+3) for every tile t <= T (T is the all-padding tile)
  ACC_TRY(h->bits.alloc(((uint64_t)T + 2) * words * 8));
  ACC_TRY(h->tco.alloc(((uint64_t)T + 4) * 4));
  // kValuePad zeros past the values (not part of the exported array): the
  // SpMM kernels load a 4-tile group's values as one 256-float span that
  // may run past the last value
  ACC_TRY(h->data.alloc(((uint64_t)nnz + kValuePad) * 4));
  ACC_HIP(hipMemsetAsync(static_cast<char*>(h->data.p) + (uint64_t)nnz * 4, 0,
                         (uint64_t)kValuePad * 4, s));
  ACC_HIP(hipMemsetAsync(h->a2b.p, 0xFF, ((uint64_t)T + 1) * kColWindow * 4, s));
  ACC_HIP(hipMemsetAsync(h->bits.p, 0, ((uint64_t)T + 2) * words * 8, s));
  DevBuf tile_of;
  ACC_TRY(tile_of.alloc((uint64_t)nnz * 4));
  if (nnz) {
    hipLaunchKernelGGL(k_place, dim3(blocks_for(nnz)), dim3(kThreads), 0, s, rp,
                       ci, rows, nnz, rw, h->rwo.as<uint32_t>(),
                       merged.as<uint32_t>(), incl.as<uint32_t>(),
                       pos.as<uint32_t>(), tile_of.as<uint32_t>(),
                       h->a2b.as<uint32_t>(),
                       h->bits.as<unsigned long long>(), words);
    ACC_HIP(hipGetLastError());
  }
  // 7. tcOffset
  DevBuf cnt;
  ACC_TRY(cnt.alloc(((uint64_t)T + 1) * 4));
  ACC_HIP(hipMemsetAsync(cnt.p, 0, ((uint64_t)T + 1) * 4, s));
  if (T) {
    hipLaunchKernelGGL(k_tile_popcount, dim3(blocks_for(T)), dim3(kThreads), 0,
                       s, h->bits.as<unsigned long long>(), T, words,
                       cnt.as<uint32_t>());
    ACC_HIP(hipGetLastError());
  }
  ACC_TRY(exclusive_scan_u32(cnt.as<uint32_t>(), h->tco.as<uint32_t>(),
                             (uint64_t)T + 1, s));
  for (uint32_t k = 1; k <= 3; ++k)
    ACC_HIP(hipMemcpyAsync(h->tco.as<uint32_t>() + T + k, h->tco.as<uint32_t>() + T, 4,
                           hipMemcpyDeviceToDevice, s));
  // 8. values
  if (nnz) {
    hipLaunchKernelGGL(k_values, dim3(blocks_for(nnz)), dim3(kThreads), 0, s,
                       nnz, csr->val, pos.as<uint32_t>(),
                       tile_of.as<uint32_t>(),
                       h->bits.as<unsigned long long>(), words,
                       h->tco.as<uint32_t>(), h->data.as<float>());
    ACC_HIP(hipGetLastError());
  }
  pos.release();
  merged.release();
  incl.release();
  tile_of.release();
  cnt.release();

  // imbalance degree on the GPU
  {
    DevBuf acc;
    ACC_TRY(acc.alloc(16));
    ACC_HIP(hipMemsetAsync(acc.p, 0, 16, s));
    uint32_t avg = W ? T / W : 0;
    hipLaunchKernelGGL(k_ibd_sum, dim3(std::min<uint32_t>(blocks_for(W), 1024)),
                       dim3(kThreads), 0, s, h->rwo.as<uint32_t>(), W, avg,
                       acc.as<unsigned long long>(),
                       reinterpret_cast<uint32_t*>(acc.as<char>() + 8));
    ACC_HIP(hipGetLastError());
    unsigned long long hacc[2] = {0, 0};
    ACC_HIP(hipMemcpyAsync(hacc, acc.p, 16, hipMemcpyDeviceToHost, s));
    ACC_HIP(hipStreamSynchronize(s));
    uint32_t sum32 = (uint32_t)(hacc[0] & 0xFFFFFFFFull);
    h->ibd = W ? sum32 / W : 0;
    h->balanced = h->ibd > o.imbalance_threshold;
    h->max_window_tiles = (uint32_t)(hacc[1] & 0xFFFFFFFFull);
  }
  ACC_TRY(measure_group_values(h, s));

  // AdpBME groups
  if (o.build_adpbme) {
    const uint32_t cf = o.t_load + o.t_comp + o.t_write;
    const uint32_t co = o.t_load + o.t_comp;
    const uint32_t target = o.target_num * (o.t_load + o.t_comp) + o.t_write;
    ACC_TRY(h->rix.alloc((uint64_t)T * 4));
    if (T == 0) {
      ACC_TRY(h->grp.alloc(4));
      ACC_HIP(hipMemsetAsync(h->grp.p, 0, 4, s));
      h->G = 0;
    } else {
      DevBuf first, next;
      ACC_TRY(first.alloc(T));
      ACC_TRY(next.alloc((uint64_t)T * 4));
      hipLaunchKernelGGL(k_tile_window, dim3(blocks_for(T)), dim3(kThreads), 0,
                         s, h->rwo.as<uint32_t>(), W, T, rw,
                         h->rix.as<uint32_t>(), first.as<uint8_t>());
      ACC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_group_next, dim3(blocks_for(T)), dim3(kThreads), 0,
                         s, first.as<uint8_t>(), T, cf, co, target,
                         next.as<uint32_t>());
      ACC_HIP(hipGetLastError());
      const uint32_t L = target / co + 1;  // longest possible group
      const uint32_t C = std::max<uint32_t>(2048, 4 * L);
      const uint32_t nchunks = (T + C - 1) / C;
      DevBuf exit_off, count, entry, gbase;
      ACC_TRY(exit_off.alloc((uint64_t)nchunks * L * 4));
      ACC_TRY(count.alloc((uint64_t)nchunks * L * 4));
      ACC_TRY(entry.alloc((uint64_t)nchunks * 4));
      ACC_TRY(gbase.alloc(((uint64_t)nchunks + 1) * 4));
      hipLaunchKernelGGL(k_chunk_paths, dim3(blocks_for((uint64_t)nchunks * L)),
                         dim3(kThreads), 0, s, next.as<uint32_t>(), T, C, L,
                         nchunks, exit_off.as<uint32_t>(), count.as<uint32_t>());
      ACC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_chunk_compose, dim3(1), dim3(64), 0, s,
                         exit_off.as<uint32_t>(), count.as<uint32_t>(),
                         nchunks, L, entry.as<uint32_t>(),
                         gbase.as<uint32_t>());
      ACC_HIP(hipGetLastError());
      uint32_t G = 0;
      ACC_HIP(hipMemcpyAsync(&G, gbase.as<uint32_t>() + nchunks, 4,
                             hipMemcpyDeviceToHost, s));
      ACC_HIP(hipStreamSynchronize(s));
      h->G = G;
      ACC_TRY(h->grp.alloc(((uint64_t)G + 1) * 4));
      hipLaunchKernelGGL(k_group_write, dim3(blocks_for(nchunks)),
                         dim3(kThreads), 0, s, next.as<uint32_t>(), T, C,
                         nchunks, entry.as<uint32_t>(), gbase.as<uint32_t>(),
                         h->grp.as<uint32_t>());
      ACC_HIP(hipGetLastError());
      ACC_HIP(hipStreamSynchronize(s));
    }
  }
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}

}  // namespace

// Tiles of the METCF of a canonical CSR (already validated by a build) at
// row window rw: the builder's steps 2-5 without the format arrays.
static accspmm_status count_tiles(const accspmm_csr_view* c, uint32_t rw, hipStream_t s,
                                  uint64_t* T) {
  const uint32_t rows = c->rows, nnz = c->nnz;
  const uint32_t W = (rows + rw - 1) / rw;
  *T = 0;
  if (nnz == 0) return ACCSPMM_OK;
  const uint32_t* rp = c->row_ptr;
  DevBuf pos, merged, incl, wt, off;
  ACC_TRY(pos.alloc((uint64_t)nnz * 4));
  ACC_TRY(merged.alloc((uint64_t)nnz * 4));
  ACC_TRY(incl.alloc((uint64_t)nnz * 4));
  ACC_TRY(wt.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(off.alloc(((uint64_t)W + 1) * 4));
  hipLaunchKernelGGL(k_merge_rank, dim3(blocks_for(nnz)), dim3(kThreads), 0, s, rp,
                     c->col_idx, rows, nnz, rw, pos.as<uint32_t>(), merged.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_first_flags, dim3(blocks_for(nnz)), dim3(kThreads), 0, s, rp, rows,
                     nnz, rw, pos.as<uint32_t>(), merged.as<uint32_t>(), incl.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(inclusive_scan_u32(incl.as<uint32_t>(), incl.as<uint32_t>(), nnz, s));
  ACC_HIP(hipMemsetAsync(wt.p, 0, ((uint64_t)W + 1) * 4, s));
  hipLaunchKernelGGL(k_window_tiles, dim3(blocks_for(W)), dim3(kThreads), 0, s, rp, rows, W,
                     rw, incl.as<uint32_t>(), wt.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(wt.as<uint32_t>(), off.as<uint32_t>(), (uint64_t)W + 1, s));
  uint32_t t = 0;
  ACC_HIP(hipMemcpyAsync(&t, off.as<uint32_t>() + W, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  *T = t;
  return ACCSPMM_OK;
}

extern "C" accspmm_status accspmm_metcf_build(const accspmm_csr_view* csr,
                                              const accspmm_build_opts* opts,
                                              accspmm_metcf** out,
                                              void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!csr || !out) return ACCSPMM_ERR_INVALID_ARG;
    *out = nullptr;
    accspmm_build_opts o;
    if (opts) o = *opts; else accspmm_build_opts_default(&o);
    if (o.row_window != 0 && o.row_window != 8 && o.row_window != 16)
      return ACCSPMM_ERR_UNSUPPORTED;
    if (o.col_window != 8) return ACCSPMM_ERR_UNSUPPORTED;
    if (o.build_adpbme && o.t_load + o.t_comp == 0) return ACCSPMM_ERR_UNSUPPORTED;
    if (o.reorder_rows > 2 || o.symmetric_order > 2) return ACCSPMM_ERR_UNSUPPORTED;
    if (o.symmetric_order && csr->rows != csr->cols) return ACCSPMM_ERR_UNSUPPORTED;
    if (csr->rows == 0 || csr->cols == 0) return ACCSPMM_ERR_INVALID_ARG;
    if (csr->nnz > kMaxNnz) return ACCSPMM_ERR_UNSUPPORTED;
    // automatic choices (row_window 0, reorder_rows 2): 16-row windows and
    // rows by length only for long rows (mean >= kLongRows nonzeros), where
    // they were measured faster (232,965-node power-law graph, mean 493:
    // 3.74 -> 3.09 ms per split-BF16 step at K=128); on graphs with short
    // rows (2M nodes, mean 33; 500-50,000 nodes, means 4-64) both were slower
    const bool long_rows = (uint64_t)csr->nnz >= (uint64_t)kLongRows * csr->rows;
    const bool auto_window = o.row_window == 0;
    if (o.row_window == 0) o.row_window = long_rows ? 16 : 8;
    if (o.reorder_rows == 2) o.reorder_rows = long_rows ? 1 : 0;
    // a symmetric ordering fixes the rows' order itself
    if (o.symmetric_order) o.reorder_rows = 0;
    if (!csr->row_ptr || (csr->nnz && (!csr->col_idx || !csr->val)))
      return ACCSPMM_ERR_INVALID_ARG;
    auto* h = new (std::nothrow) accspmm_metcf();
    if (!h) return ACCSPMM_ERR_OUT_OF_MEMORY;
    // reorder_rows: build from the rows in order of decreasing length;
    // symmetric_order: build from P A P^T.  The handle keeps the permutation
    // (rows; columns for P A P^T): the kernels write C rows through it and
    // each call permutes B through it.
    DevBuf prp, pci, pva, rperm, cperm;
    accspmm_csr_view pv;
    const accspmm_csr_view* src = csr;
    accspmm_status st = ACCSPMM_OK;
    if (o.reorder_rows || o.symmetric_order) {
      st = o.symmetric_order
               ? symmetric_permute(csr, o.symmetric_order, prp, pci, pva, rperm,
                                   (hipStream_t)stream)
               : reorder_rows_by_length(csr, prp, pci, pva, rperm, (hipStream_t)stream);
      if (st == ACCSPMM_OK && o.symmetric_order) {
        st = cperm.alloc((size_t)csr->cols * 4);
        if (st == ACCSPMM_OK &&
            hipMemcpyAsync(cperm.p, rperm.p, (size_t)csr->cols * 4, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream) != hipSuccess)
          st = ACCSPMM_ERR_HIP;
      }
      pv = accspmm_csr_view{csr->rows, csr->cols, csr->nnz, prp.as<uint32_t>(),
                            pci.as<uint32_t>(), pva.as<float>()};
      src = &pv;
    }
    if (st == ACCSPMM_OK) st = build_impl(src, o, h, (hipStream_t)stream);
    // automatic choice on short rows of a large graph (>= 64 * 8192 8-row
    // tiles): 16-row windows when they hold at most 92% of the 8-row tiles,
    // i.e. rows of a window share columns as in banded matrices (27-point
    // stencils: 0.88-0.90; large random and power-law graphs: 0.95-0.99).
    // 27-point stencil 100^3, split-BF16 at K=128: 0.591 ms against 0.674 ms;
    // FP32 0.712 against 0.822 ms.  Small graphs keep 8-row windows (many
    // small dense ones share columns too, unmeasured there).
    // The 8-row build validated the CSR, which the tile count needs.  The
    // count is optional: if it fails the 8-row graph stands.  The 8-row
    // arrays are released before the 16-row build (peak memory: one graph
    // and the count's scratch, never two graphs); should that build fail,
    // the 8-row graph is built again.
    if (st == ACCSPMM_OK && auto_window && o.row_window == 8 && !o.reorder_rows &&
        (uint64_t)h->T >= 64ull * 8192) {
      uint64_t t16 = 0;
      const uint64_t t8 = h->T;
      if (count_tiles(src, 16, (hipStream_t)stream, &t16) == ACCSPMM_OK &&
          t16 * 100 <= t8 * 92) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        delete h;
        h = new (std::nothrow) accspmm_metcf();
        if (!h) return ACCSPMM_ERR_OUT_OF_MEMORY;
        accspmm_build_opts o16 = o;
        o16.row_window = 16;
        st = build_impl(src, o16, h, (hipStream_t)stream);
        if (st != ACCSPMM_OK) {
          (void)hipStreamSynchronize((hipStream_t)stream);
          delete h;
          h = new (std::nothrow) accspmm_metcf();
          if (!h) return ACCSPMM_ERR_OUT_OF_MEMORY;
          st = build_impl(src, o, h, (hipStream_t)stream);
        }
      } else {
        (void)hipGetLastError();
      }
    }
    if (st != ACCSPMM_OK) {
      (void)hipStreamSynchronize((hipStream_t)stream);
      delete h;
      return st;
    }
    h->row_perm.swap(rperm);
    h->col_perm.swap(cperm);
    h->opts.symmetric_order = o.symmetric_order;
    *out = h;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_metcf_build_host(const accspmm_csr_view* hc,
                                                   const accspmm_build_opts* opts,
                                                   accspmm_metcf** out) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!hc || !out || !hc->row_ptr || (hc->nnz && (!hc->col_idx || !hc->val)))
      return ACCSPMM_ERR_INVALID_ARG;
    DevBuf rp, ci, va;
    ACC_TRY(rp.alloc(((uint64_t)hc->rows + 1) * 4));
    ACC_TRY(ci.alloc((uint64_t)hc->nnz * 4));
    ACC_TRY(va.alloc((uint64_t)hc->nnz * 4));
    ACC_HIP(hipMemcpy(rp.p, hc->row_ptr, ((uint64_t)hc->rows + 1) * 4, hipMemcpyHostToDevice));
    if (hc->nnz) {
      ACC_HIP(hipMemcpy(ci.p, hc->col_idx, (uint64_t)hc->nnz * 4, hipMemcpyHostToDevice));
      ACC_HIP(hipMemcpy(va.p, hc->val, (uint64_t)hc->nnz * 4, hipMemcpyHostToDevice));
    }
    accspmm_csr_view dv{hc->rows, hc->cols, hc->nnz, rp.as<uint32_t>(), ci.as<uint32_t>(),
                        va.as<float>()};
    return accspmm_metcf_build(&dv, opts, out, nullptr);
  });
}

extern "C" accspmm_status accspmm_metcf_destroy(accspmm_metcf* h) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h) return ACCSPMM_ERR_INVALID_ARG;
    (void)hipDeviceSynchronize();
    delete h;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_metcf_get_info(const accspmm_metcf* h,
                                                 accspmm_metcf_info* info) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !info) return ACCSPMM_ERR_INVALID_ARG;
    info->rows = h->rows;
    info->cols = h->cols;
    info->nnz = h->nnz;
    info->row_window = h->rw;
    info->num_windows = h->W;
    info->num_tiles = h->T;
    info->num_groups = h->G;
    info->ibd = h->ibd;
    info->balanced = h->balanced;
    info->max_window_tiles = h->max_window_tiles;
    info->device_bytes = h->device_bytes();
    info->reordered = h->row_perm.p ? 1u : 0u;
    info->max_group_values = h->max_group_values;
    info->symmetric_order = h->col_perm.p ? h->opts.symmetric_order : 0u;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_metcf_row_order(const accspmm_metcf* h,
                                                  uint32_t* row_order) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !row_order) return ACCSPMM_ERR_INVALID_ARG;
    if (!h->row_perm.p) {
      for (uint32_t i = 0; i < h->rows; ++i) row_order[i] = i;
      return ACCSPMM_OK;
    }
    ACC_HIP(hipSetDevice(h->device));
    ACC_HIP(hipMemcpy(row_order, h->row_perm.p, (size_t)h->rows * 4, hipMemcpyDeviceToHost));
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_metcf_col_order(const accspmm_metcf* h,
                                                  uint32_t* col_order) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !col_order) return ACCSPMM_ERR_INVALID_ARG;
    if (!h->col_perm.p) {
      for (uint32_t j = 0; j < h->cols; ++j) col_order[j] = j;
      return ACCSPMM_OK;
    }
    ACC_HIP(hipSetDevice(h->device));
    ACC_HIP(hipMemcpy(col_order, h->col_perm.p, (size_t)h->cols * 4, hipMemcpyDeviceToHost));
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_metcf_export(
    const accspmm_metcf* h, uint32_t* rowWindowOffset, uint32_t* tcOffset,
    uint64_t* tcLocalBit, uint32_t* sparseA2B, float* data,
    uint32_t* groupOffset, uint32_t* rowIndices) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h) return ACCSPMM_ERR_INVALID_ARG;
    if ((groupOffset || rowIndices) && !h->opts.build_adpbme)
      return ACCSPMM_ERR_INVALID_ARG;
    ACC_HIP(hipSetDevice(h->device));
    ACC_HIP(hipDeviceSynchronize());
    auto cp = [](void* dst, const void* src, uint64_t bytes) -> hipError_t {
      if (!dst || bytes == 0) return hipSuccess;
      return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    };
    ACC_HIP(cp(rowWindowOffset, h->rwo.p, ((uint64_t)h->W + 1) * 4));
    ACC_HIP(cp(tcOffset, h->tco.p, ((uint64_t)h->T + 1) * 4));
    ACC_HIP(cp(tcLocalBit, h->bits.p, (uint64_t)h->T * h->words * 8));
    ACC_HIP(cp(sparseA2B, h->a2b.p, (uint64_t)h->T * kColWindow * 4));
    ACC_HIP(cp(data, h->data.p, (uint64_t)h->nnz * 4));
    ACC_HIP(cp(groupOffset, h->grp.p, ((uint64_t)h->G + 1) * 4));
    ACC_HIP(cp(rowIndices, h->rix.p, (uint64_t)h->T * 4));
    return ACCSPMM_OK;
  });
}
