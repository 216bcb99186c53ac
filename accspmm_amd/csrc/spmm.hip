// spmm.hip — METCF SpMM kernels for gfx950 (CDNA4, wave64, MFMA).
//
// Replaces the four TF32 kernels of include/mma_tf32.cuh (tf32_mma_kernel
// :13-228, tf32_mma_g512_kernel :230-550, tf32_mma_balance_kernel :552-792,
// tf32_mma_balance_g128_kernel :794-1042) and their launcher
// TF32Compute::compute / adpBalanceCompute (include/tf32_comp.hpp:52-113).
//
// Work decomposition.  One wave processes one work unit (a contiguous tile
// range of one row window) for one feature chunk of KC = 16*NB columns.  The
// unit's tiles are consumed in groups of four: quarter-wave q (lanes
// 16q..16q+15) owns tile q of the group.  Lane n = lane%16 of a quarter
// gathers, for each of the tile's 8 condensed columns, NB consecutive
// features [k0 + NB*n, k0 + NB*n + NB) of that B row straight into VGPRs
// (16 lanes x NB x 4 B = one contiguous segment of the row), and decodes the
// A values of its tile row from the bitmap with popcount.  Every lane
// consumes exactly what it loaded, so no LDS is involved; loads of group
// g+1 are issued before the MFMAs of group g (register double buffer).
//
// MFMA mapping (D[m][n] += sum_k A[m][k] * B[k][n], 16x16 output):
//   rows m     = tile rows (RW=8: rows 0-7; 8-15 carry A_lo in the split
//                form and are zero otherwise)
//   columns n  = feature NB*n + b for block b in [0, NB)
//   reduction  = the 32 condensed columns of the 4-tile group
// FP32 form:   v_mfma_f32_16x16x4_f32, k = tile q, one MFMA per column slot.
// BF16 forms:  v_mfma_f32_16x16x32_bf16, k = 8q + column slot.
// Split-BF16:  A = [A_hi; A_lo] against B_hi and B_lo (2 MFMAs per block)
//              gives all four products; rows r and r+8 are added at the end.
//
// Split windows: a window with more than `split_tiles` tiles is cut into
// balanced pieces; each piece writes a partial, and the last piece to arrive
// (per-window atomic counter) sums all partials in piece order and writes
// C, so the result bits do not depend on scheduling.  The counter is reset
// by that last piece, ready for the next launch.
#include <algorithm>
#include <cstring>

#include "internal.hpp"

#include "spmm_device.hpp"

namespace accspmm {
namespace {
using namespace dev;

template <int FORM, int NB, int RW>
__global__ __launch_bounds__(512) void spmm_kernel(SpmmArgs g) {
  const uint32_t wpb = blockDim.x >> 6;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t unit, kc;
  if (g.xcd_slice) {
    // blocks round-robin over the 8 XCDs: keep feature chunk kc on XCDs
    // x with x % n_kc == kc so each XCD's L2 holds one column slice of B
    const uint32_t x = blockIdx.x & 7u, per = 8u / g.n_kc;
    kc = x % g.n_kc;
    unit = ((blockIdx.x >> 3) * per + x / g.n_kc) * wpb + wave;
  } else {
    const uint32_t tau = blockIdx.x * wpb + wave;
    unit = tau / g.n_kc;
    kc = tau % g.n_kc;
  }
  if (unit >= g.num_units) return;
  const WorkUnit u = g.units[unit];
  run_unit<FORM, NB, RW>(g, u, kc, g.B, g.ldb, g.C, g.ldc, g.K, g.rows);
}

// B conversion for the BF16 forms: zero-padded to ldx >= K columns.
__global__ void convert_b_split(const float* __restrict__ B, uint32_t N,
                                uint32_t K, uint32_t ldb,
                                uint32_t* __restrict__ out, uint32_t ldx) {
  for (uint32_t r = blockIdx.x; r < N; r += gridDim.x)
    for (uint32_t k = threadIdx.x; k < ldx; k += blockDim.x) {
      uint32_t v = 0;
      if (k < K) {
        float x = B[(uint64_t)r * ldb + k];
        uint32_t hp = pack_bf16(x, 0.f) & 0xFFFFu;
        float lo = x - bf16_lo_to_f32(hp);
        uint32_t lp = pack_bf16(lo, 0.f) & 0xFFFFu;
        v = hp | (lp << 16);
      }
      out[(uint64_t)r * ldx + k] = v;
    }
}

__global__ void convert_b_bf16(const float* __restrict__ B, uint32_t N,
                               uint32_t K, uint32_t ldb,
                               uint16_t* __restrict__ out, uint32_t ldx) {
  for (uint32_t r = blockIdx.x; r < N; r += gridDim.x)
    for (uint32_t k = threadIdx.x; k < ldx; k += blockDim.x) {
      uint16_t v = 0;
      if (k < K) v = (uint16_t)(pack_bf16(B[(uint64_t)r * ldb + k], 0.f) & 0xFFFFu);
      out[(uint64_t)r * ldx + k] = v;
    }
}

// ---- schedule construction (one thread per window) ----------------------
__global__ void k_unit_counts(const uint32_t* __restrict__ rwo, uint32_t W,
                              uint32_t split, uint32_t* __restrict__ nunits,
                              uint32_t* __restrict__ npieces,
                              uint32_t* __restrict__ nsplit) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t tw = rwo[w + 1] - rwo[w];
    uint32_t nu = tw > split ? (tw + split - 1) / split : 1u;
    nunits[w] = nu;
    npieces[w] = nu > 1 ? nu : 0u;
    nsplit[w] = nu > 1 ? 1u : 0u;
  }
}

__global__ void k_unit_write(const uint32_t* __restrict__ rwo, uint32_t W,
                             const uint32_t* __restrict__ unit_off,
                             const uint32_t* __restrict__ piece_off,
                             const uint32_t* __restrict__ split_off,
                             WorkUnit* __restrict__ units,
                             uint32_t* __restrict__ piece_info) {
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t tb = rwo[w], te = rwo[w + 1];
    const uint32_t nu = unit_off[w + 1] - unit_off[w];
    const uint32_t base = unit_off[w];
    if (nu == 1) {
      units[base] = WorkUnit{(uint32_t)w, tb, te, -1};
      continue;
    }
    // balanced pieces, boundaries on 4-tile groups where possible
    const uint32_t tw = te - tb;
    const uint32_t first = piece_off[w];
    for (uint32_t p = 0; p < nu; ++p) {
      uint32_t a = (uint32_t)((uint64_t)tw * p / nu);
      uint32_t b = (uint32_t)((uint64_t)tw * (p + 1) / nu);
      a = p == 0 ? 0 : min(tw, (a + 3u) & ~3u);
      b = p + 1 == nu ? tw : min(tw, (b + 3u) & ~3u);
      units[base + p] = WorkUnit{(uint32_t)w, tb + a, tb + max(a, b),
                                 (int32_t)(first + p)};
      piece_info[3ull * (first + p)] = first;
      piece_info[3ull * (first + p) + 1] = nu;
      piece_info[3ull * (first + p) + 2] = split_off[w];
    }
  }
}

}  // namespace

accspmm_status get_schedule(const accspmm_metcf* h, uint32_t split_tiles,
                            hipStream_t s, const Schedule** out) {
  if (split_tiles == 0) split_tiles = 0xFFFFFFFFu;
  auto it = h->schedules.find(split_tiles);
  if (it != h->schedules.end()) {
    *out = it->second.get();
    return ACCSPMM_OK;
  }
  auto sch = std::make_unique<Schedule>();
  sch->split_tiles = split_tiles;
  const uint32_t W = h->W;
  DevBuf nunits, npieces, nsplit, uoff, poff, soff;
  ACC_TRY(nunits.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(npieces.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(nsplit.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(uoff.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(poff.alloc(((uint64_t)W + 1) * 4));
  ACC_TRY(soff.alloc(((uint64_t)W + 1) * 4));
  ACC_HIP(hipMemsetAsync(nunits.p, 0, ((uint64_t)W + 1) * 4, s));
  ACC_HIP(hipMemsetAsync(npieces.p, 0, ((uint64_t)W + 1) * 4, s));
  ACC_HIP(hipMemsetAsync(nsplit.p, 0, ((uint64_t)W + 1) * 4, s));
  uint32_t nb = (uint32_t)std::min<uint64_t>((W + 255) / 256, 65535);
  hipLaunchKernelGGL(k_unit_counts, dim3(std::max(nb, 1u)), dim3(256), 0, s,
                     h->rwo.as<uint32_t>(), W, split_tiles,
                     nunits.as<uint32_t>(), npieces.as<uint32_t>(),
                     nsplit.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_TRY(exclusive_scan_u32(nunits.as<uint32_t>(), uoff.as<uint32_t>(), (uint64_t)W + 1, s));
  ACC_TRY(exclusive_scan_u32(npieces.as<uint32_t>(), poff.as<uint32_t>(), (uint64_t)W + 1, s));
  ACC_TRY(exclusive_scan_u32(nsplit.as<uint32_t>(), soff.as<uint32_t>(), (uint64_t)W + 1, s));
  uint32_t tot[3];
  ACC_HIP(hipMemcpyAsync(&tot[0], uoff.as<uint32_t>() + W, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipMemcpyAsync(&tot[1], poff.as<uint32_t>() + W, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipMemcpyAsync(&tot[2], soff.as<uint32_t>() + W, 4, hipMemcpyDeviceToHost, s));
  ACC_HIP(hipStreamSynchronize(s));
  sch->num_units = tot[0];
  sch->num_pieces = tot[1];
  sch->num_split_windows = tot[2];
  ACC_TRY(sch->units.alloc((uint64_t)tot[0] * sizeof(WorkUnit)));
  ACC_TRY(sch->piece_first.alloc((uint64_t)std::max(tot[1], 1u) * 12));
  hipLaunchKernelGGL(k_unit_write, dim3(std::max(nb, 1u)), dim3(256), 0, s,
                     h->rwo.as<uint32_t>(), W, uoff.as<uint32_t>(),
                     poff.as<uint32_t>(), soff.as<uint32_t>(),
                     sch->units.as<WorkUnit>(), sch->piece_first.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  ACC_HIP(hipStreamSynchronize(s));
  *out = sch.get();
  h->schedules[split_tiles] = std::move(sch);
  return ACCSPMM_OK;
}

namespace {

template <int FORM, int NB, int RW>
void launch_t(const SpmmArgs& g, uint32_t blocks, uint32_t wpb, hipStream_t s) {
  hipLaunchKernelGGL((spmm_kernel<FORM, NB, RW>), dim3(blocks), dim3(64 * wpb),
                     0, s, g);
}

template <int FORM, int RW>
accspmm_status launch_nb(uint32_t nb, const SpmmArgs& g, uint32_t blocks,
                         uint32_t wpb, hipStream_t s) {
  switch (nb) {
    case 1: launch_t<FORM, 1, RW>(g, blocks, wpb, s); break;
    case 2: launch_t<FORM, 2, RW>(g, blocks, wpb, s); break;
    case 4: launch_t<FORM, 4, RW>(g, blocks, wpb, s); break;
    case 8: launch_t<FORM, 8, RW>(g, blocks, wpb, s); break;
    default: return ACCSPMM_ERR_UNSUPPORTED;
  }
  return ACCSPMM_OK;
}

accspmm_status launch_form(int form, uint32_t rw, uint32_t nb,
                           const SpmmArgs& g, uint32_t blocks, uint32_t wpb,
                           hipStream_t s) {
  if (rw == 8) {
    if (form == kFP32) return launch_nb<kFP32, 8>(nb, g, blocks, wpb, s);
    if (form == kSPLIT) return launch_nb<kSPLIT, 8>(nb, g, blocks, wpb, s);
    return launch_nb<kBF16, 8>(nb, g, blocks, wpb, s);
  }
  if (form == kFP32) return launch_nb<kFP32, 16>(nb, g, blocks, wpb, s);
  if (form == kSPLIT) return launch_nb<kSPLIT, 16>(nb, g, blocks, wpb, s);
  return launch_nb<kBF16, 16>(nb, g, blocks, wpb, s);
}

}  // namespace

// Default tile heuristic: widest feature chunk whose alignment the operands
// allow, capped so small K does not idle lanes.
void default_tile(const accspmm_metcf* h, uint32_t K, int form,
                  accspmm_tile* t) {
  uint32_t nb = K >= 128 ? 4 : K >= 64 ? 4 : K >= 32 ? 2 : 1;
  (void)form;
  t->feat_per_lane = nb;
  t->waves_per_block = 4;
  // split windows much longer than the mean (keeps the tail short)
  uint32_t mean = h->W ? h->T / h->W : 0;
  uint32_t s = std::max<uint32_t>(64, 4 * std::max<uint32_t>(mean, 16));
  t->split_tiles = h->max_window_tiles > s ? s : 0;
  t->xcd_slice = 0;
}

// phases: bit 0 = convert B (BF16 forms), bit 1 = SpMM kernel
accspmm_status spmm_launch(const accspmm_metcf* h, const float* B, uint32_t ldb,
                           float* C, uint32_t ldc, uint32_t K, int form,
                           const accspmm_tile& tile, hipStream_t s,
                           int phases) {
  uint32_t nb = tile.feat_per_lane;
  if (nb != 1 && nb != 2 && nb != 4 && nb != 8) return ACCSPMM_ERR_UNSUPPORTED;
  uint32_t wpb = tile.waves_per_block ? tile.waves_per_block : 4;
  if (wpb > 8) return ACCSPMM_ERR_UNSUPPORTED;
  const uint32_t KC = 16 * nb;
  const uint32_t n_kc = (K + KC - 1) / KC;
  // FP32 reads the caller's B with NB-wide vector loads: needs alignment
  if (form == kFP32) {
    if ((ldb % nb) != 0 || ((uintptr_t)B % (4u * nb)) != 0) return ACCSPMM_ERR_UNSUPPORTED;
  }
  if ((phases & 2) && ((ldc % nb) != 0 || ((uintptr_t)C % (4u * nb)) != 0))
    return ACCSPMM_ERR_UNSUPPORTED;

  std::lock_guard<std::mutex> lock(h->mu);
  const Schedule* sch = nullptr;
  ACC_TRY(get_schedule(h, tile.split_tiles, s, &sch));

  SpmmArgs g{};
  g.tco = h->tco.as<uint32_t>();
  g.bits = h->bits.as<unsigned long long>();
  g.a2b = h->a2b.as<uint32_t>();
  g.data = h->data.as<float>();
  g.units = sch->units.as<WorkUnit>();
  g.piece_info = sch->piece_first.as<uint32_t>();
  g.C = C;
  g.ldc = ldc;
  g.rows = h->rows;
  g.K = K;
  g.num_units = sch->num_units;
  g.n_kc = n_kc;
  g.kc_padded = n_kc * KC;
  g.xcd_slice = (tile.xcd_slice && (8 % n_kc) == 0) ? 1u : 0u;

  if (sch->num_pieces) {
    const uint64_t pbytes = (uint64_t)sch->num_pieces * h->rw * g.kc_padded * 4;
    ACC_TRY(h->partials.reserve(pbytes));
    // arrival counters: one per (split window, feature chunk); zero at rest
    const uint64_t cbytes = (uint64_t)sch->num_split_windows * n_kc * 4;
    auto* mut = const_cast<Schedule*>(sch);
    if (mut->counters.bytes < cbytes) {
      ACC_TRY(mut->counters.alloc(cbytes));
      ACC_HIP(hipMemsetAsync(mut->counters.p, 0, cbytes, s));
    }
    g.partials = h->partials.as<float>();
    g.counters = mut->counters.as<uint32_t>();
  }

  if (form == kFP32) {
    g.B = B;
    g.ldb = ldb;
  } else {
    const uint32_t ldx = n_kc * KC;
    const uint64_t es = form == kSPLIT ? 4 : 2;
    ACC_TRY(h->bconv.reserve((uint64_t)h->cols * ldx * es));
    uint32_t blocks = std::min<uint32_t>(h->cols, 65535u * 4);
    if (!(phases & 1)) {
      // converted by an earlier prepare call
    } else if (form == kSPLIT)
      hipLaunchKernelGGL(convert_b_split, dim3(blocks), dim3(256), 0, s, B,
                         h->cols, K, ldb, h->bconv.as<uint32_t>(), ldx);
    else
      hipLaunchKernelGGL(convert_b_bf16, dim3(blocks), dim3(256), 0, s, B,
                         h->cols, K, ldb, h->bconv.as<uint16_t>(), ldx);
    ACC_HIP(hipGetLastError());
    g.B = h->bconv.p;
    g.ldb = ldx;
  }
  if (!(phases & 2)) return ACCSPMM_OK;

  uint64_t blocks;
  if (g.xcd_slice) {
    const uint64_t ub = (sch->num_units + wpb - 1) / wpb;
    blocks = (ub * n_kc + 7) / 8 * 8;
  } else {
    blocks = ((uint64_t)sch->num_units * n_kc + wpb - 1) / wpb;
  }
  if (blocks == 0) return ACCSPMM_OK;
  if (blocks > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  ACC_TRY(launch_form(form, h->rw, nb, g, (uint32_t)blocks, wpb, s));
  ACC_HIP(hipGetLastError());
  return ACCSPMM_OK;
}

}  // namespace accspmm

using namespace accspmm;

namespace accspmm {
// Tile for (h, K, form): the caller's if given, else the tuned/default pick;
// then narrowed until the NB-wide vector loads/stores fit the operands'
// alignment.  Shared by the single-graph and batched paths so a problem gets
// the same tile (and so the same result bits) in both.
accspmm_status spmm_fit_tile(const accspmm_metcf* h, const float* B,
                             uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                             int form, const accspmm_tile* tile,
                             accspmm_tile* out) {
  accspmm_tile t{};
  if (tile && tile->feat_per_lane) {
    t = *tile;
  } else {
    ACC_TRY(accspmm_tile_select(h, K, (accspmm_form)form, &t));
  }
  while (t.feat_per_lane > 1 &&
         ((form == kFP32 &&
           ((ldb % t.feat_per_lane) || ((uintptr_t)B % (4u * t.feat_per_lane)))) ||
          (C && ((ldc % t.feat_per_lane) || ((uintptr_t)C % (4u * t.feat_per_lane))))))
    t.feat_per_lane /= 2;
  *out = t;
  return ACCSPMM_OK;
}
}  // namespace accspmm

static accspmm_status spmm_entry(const accspmm_metcf* h, const float* B,
                                 uint32_t ldb, float* C, uint32_t ldc,
                                 uint32_t K, accspmm_form form,
                                 const accspmm_tile* tile, void* stream,
                                 int phases) {
  if (!h || !B || K == 0) return ACCSPMM_ERR_INVALID_ARG;
  if ((phases & 2) && !C) return ACCSPMM_ERR_INVALID_ARG;
  if (ldb < K || ((phases & 2) && ldc < K)) return ACCSPMM_ERR_SHAPE;
  if ((int)form < 0 || (int)form > 2) return ACCSPMM_ERR_INVALID_ARG;
  accspmm_tile t{};
  ACC_TRY(spmm_fit_tile(h, B, ldb, (phases & 2) ? C : nullptr, ldc, K, (int)form,
                        tile, &t));
  return spmm_launch(h, B, ldb, C, ldc, K, (int)form, t, (hipStream_t)stream,
                     phases);
}

extern "C" accspmm_status accspmm_spmm(const accspmm_metcf* h, const float* B,
                                       uint32_t ldb, float* C, uint32_t ldc,
                                       uint32_t K, accspmm_form form,
                                       const accspmm_tile* tile,
                                       void* stream) {
  return spmm_entry(h, B, ldb, C, ldc, K, form, tile, stream, 3);
}

extern "C" accspmm_status accspmm_spmm_prepare_b(const accspmm_metcf* h,
                                                 const float* B, uint32_t ldb,
                                                 uint32_t K, accspmm_form form,
                                                 const accspmm_tile* tile,
                                                 void* stream) {
  return spmm_entry(h, B, ldb, nullptr, K, K, form, tile, stream, 1);
}

extern "C" accspmm_status accspmm_spmm_prepared(const accspmm_metcf* h,
                                                const float* B, uint32_t ldb,
                                                float* C, uint32_t ldc,
                                                uint32_t K, accspmm_form form,
                                                const accspmm_tile* tile,
                                                void* stream) {
  return spmm_entry(h, B, ldb, C, ldc, K, form, tile, stream, 2);
}
