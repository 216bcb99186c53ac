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
//   rows m     = tile rows (RW=8: rows 0-7; 8-15 carry a second A term in
//                the split form and are zero otherwise)
//   columns n  = feature NB*n + b for block b in [0, NB)
//   reduction  = the 32 condensed columns of the 4-tile group
// FP32 form:   v_mfma_f32_16x16x4_f32, k = tile q, one MFMA per column slot.
// BF16 forms:  v_mfma_f32_16x16x32_bf16, k = 8q + column slot.
// Split-BF16:  A and B split exactly into three bf16 terms in registers
//              (bf16x6): RW=8 [A0; A1] against B0, B1, B2 and [A2; 0]
//              against B0 (4 MFMAs per block; rows r and r+8 are added at
//              the end), RW=16 the six products A_i B_j, i + j <= 2.
//
// Split windows: a window with more than `split_tiles` tiles is cut into
// balanced pieces; each piece writes a partial, and k_reduce_split sums a
// window's partials in piece order and writes C, so the result bits do not
// depend on scheduling.  The batched kernel runs the same second phase.
//
// Column panels: with col_panels P > 1, long windows are also cut where
// their (column-sorted) tiles cross P equal column ranges of B, and the
// pieces run range by range over all windows.  A graph whose B does not fit
// the 4 MiB L2 of an XCD otherwise gathers every B row from the MALL/HBM at
// random; panel by panel the rows in flight span 1/P of B and are re-read
// from L2.
#include <algorithm>
#include <cstring>

#include "internal.hpp"

#include "spmm_device.hpp"

namespace accspmm {
// whether a launch over g gathers B by buffer loads (every offset fits)
bool gather_by_buffer(const dev::SpmmArgs& g, int form) {
  return dev::buffer_ok(g.b_rows, (uint64_t)g.ldb * (form == dev::kBF16 ? 2 : 4));
}

// The kernel form of a call: split-BF16 with b_split 2 splits the caller's
// FP32 B in registers (kSPLITF), otherwise it gathers the converted copy.
int kernel_form(int form, const accspmm_tile& tile) {
  return form == dev::kSPLIT && tile.b_split == 2 ? (int)dev::kSPLITF : form;
}

namespace {
using namespace dev;

// NB = 8 holds 2 x 64 words of gathered B per lane in flight: blocks of at
// most 4 waves, so the compiler may give a wave all 512 registers of its
// SIMD lane (VGPRs and AGPRs).
template <int FORM, int NB, int RW, bool PACKED, bool BUF>
__global__ __launch_bounds__(NB == 8 ? 256 : 512) void spmm_kernel(SpmmArgs g) {
  const uint32_t wpb = blockDim.x >> 6;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t unit, kc;
  if (g.xcd_slice == 1) {
    // blocks round-robin over the 8 XCDs: keep feature chunk kc on XCDs
    // x with x % n_kc == kc so each XCD's L2 holds one column slice of B
    const uint32_t x = blockIdx.x & 7u, per = 8u / g.n_kc;
    kc = x % g.n_kc;
    unit = ((blockIdx.x >> 3) * per + x / g.n_kc) * wpb + wave;
  } else if (g.xcd_slice == 2) {
    // XCD x (blocks x, x+8, ...) runs the x-th eighth of the work units in
    // order: neighbouring windows (banded matrices) or one range of column
    // panels (panel-major schedules) share that XCD's L2
    const uint32_t per = gridDim.x >> 3;
    const uint32_t tau = ((blockIdx.x & 7u) * per + (blockIdx.x >> 3)) * wpb + wave;
    unit = tau / g.n_kc;
    kc = tau % g.n_kc;
  } else if (g.xcd_slice == 3) {
    // pinned schedule: XCD x runs its own unit list [x * pin_len, ...), the
    // column panels p = x (mod 8), so each B row is cached by one XCD's L2
    const uint32_t i = (blockIdx.x >> 3) * wpb + wave;
    const uint32_t slot = i / g.n_kc;
    if (slot >= g.pin_len) return;
    kc = i % g.n_kc;
    unit = (blockIdx.x & 7u) * g.pin_len + slot;
  } else {
    const uint32_t tau = blockIdx.x * wpb + wave;
    unit = tau / g.n_kc;
    kc = tau % g.n_kc;
  }
  if (unit >= g.num_units) return;
  // per-wave value buffers, sized by the launch (wpb * wave_lds floats): a
  // fixed 8-wave array would hold 16.5 KiB per 1-wave block and cap the CU
  // at 9 resident blocks
  extern __shared__ __align__(16) float lds_vals[];
  float* lv = lds_vals + wave * wave_lds<RW>();
  if ((threadIdx.x & 63u) == 0) lv[group_values<RW>()] = 0.f;
  const WorkUnit u = g.units[unit];
  run_unit<FORM, NB, RW, PACKED, BUF>(g, u, kc, g.B, g.ldb, g.C, g.ldc, g.K, g.rows, lv);
}

// B conversion for the converted forms (split-BF16 by pass, BF16):
// zero-padded to ldx >= K columns, plus one all-zero row N that padded tile
// columns read: grid-stride over the (N + 1) * ldx / 4 four-feature chunks
// (convert_chunk).
template <int FORM>
__global__ __launch_bounds__(256) void convert_b(const float* __restrict__ B, uint32_t N,
                                                 uint32_t K, uint32_t ldb, void* __restrict__ out,
                                                 uint32_t ldx, uint32_t vec,
                                                 const uint32_t* __restrict__ perm) {
  const uint64_t total = (uint64_t)(N + 1) * (ldx / 4);
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total;
       c += (uint64_t)gridDim.x * blockDim.x)
    convert_chunk<FORM>(B, N, K, ldb, out, ldx, vec, c, perm);
}

// For window w and panel boundary p in [1, P): the number of the window's
// tiles whose first condensed column lies below column p * pc.  Tiles of a
// window are in ascending column order, so this is a lower bound search.
__global__ void k_panel_bounds(const uint32_t* __restrict__ rwo,
                               const uint32_t* __restrict__ a2b, uint32_t W,
                               uint32_t P, uint32_t pc,
                               uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)W * (P - 1)) return;
  const uint32_t w = (uint32_t)(i / (P - 1)), p = (uint32_t)(i % (P - 1)) + 1;
  const uint64_t key = (uint64_t)p * pc;
  const uint32_t base = rwo[w];
  uint32_t lo = base, hi = rwo[w + 1];
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (a2b[8ull * mid] < key) lo = mid + 1; else hi = mid;
  }
  out[i] = lo - base;
}

// Second phase for split windows (one block per split window).
template <int RW>
__global__ __launch_bounds__(256) void k_reduce_split(
    const uint32_t* __restrict__ sinfo, const float* __restrict__ partials,
    uint32_t kcp, uint32_t K, float* __restrict__ C, uint32_t ldc, uint32_t rows,
    const uint32_t* __restrict__ rmap) {
  reduce_split_window<RW>(sinfo, blockIdx.x, partials, kcp, K, C, ldc, rows, rmap);
}

// Shortest piece a column panel cut may leave (shorter ones are merged
// with their neighbour: a piece pays a partial write and read).
constexpr uint32_t kPanelMinTiles = 8;

}  // namespace

// Work-unit schedule, built on the host in one pass over rowWindowOffset
// (cached per handle and (split, pack, panels)):
//   * with col_panels P > 1, a window of at least 2*kPanelMinTiles tiles is
//     cut where its tiles cross the column boundaries p*ceil(cols/P)
//     (rounded to 4-tile groups, short pieces merged); its pieces are
//     scheduled panel by panel across all windows, so the GPU gathers from
//     one column range of B at a time;
//   * a window (or panel piece) with more than `split_tiles` tiles becomes
//     ceil(T/split) balanced pieces (boundaries on 4-tile groups);
//   * each piece has a partial slot; the pieces of a window have
//     consecutive slots, summed in slot order by the second phase;
//   * other windows are packed, in order, into units of at least
//     `pack_tiles` tiles (a heavy window closes the open unit first); runs
//     of empty windows form units of their own, which write zero rows.
// Unit order: whole-window units first, then panel 0's pieces, panel 1's...
// plan_schedule fills the host side (h_units, h_sinfo, counts); the panel
// cut search runs on the GPU into `bounds` (W * (P - 1) words of device
// memory the caller provides).
uint64_t panel_bound_words(const accspmm_metcf* h, const accspmm_tile& tile) {
  const uint32_t P = std::max<uint32_t>(1, std::min<uint32_t>(tile.col_panels, 1024));
  return P > 1 && h->W > 0 && h->T > 0 ? (uint64_t)h->W * (P - 1) : 0;
}

std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> schedule_key(const accspmm_tile& tile) {
  const uint32_t split_tiles = tile.split_tiles ? tile.split_tiles : 0xFFFFFFFFu;
  const uint32_t P = std::max<uint32_t>(1, std::min<uint32_t>(tile.col_panels, 1024));
  return std::make_tuple(split_tiles, tile.pack_tiles, P, tile.xcd_slice == 3 && P > 1 ? 1u : 0u);
}

accspmm_status plan_schedule(const accspmm_metcf* h, const accspmm_tile& tile, hipStream_t s,
                             uint32_t* bounds_dev, Schedule* sch) {
  const uint32_t split_tiles = tile.split_tiles ? tile.split_tiles : 0xFFFFFFFFu;
  const uint32_t pack_tiles = tile.pack_tiles;
  const uint32_t P = std::max<uint32_t>(1, std::min<uint32_t>(tile.col_panels, 1024));
  sch->split_tiles = split_tiles;
  sch->pack_tiles = pack_tiles;
  sch->col_panels = P;
  const std::vector<uint32_t>& rwo = h->h_rwo;

  // panel cut positions (tiles below each boundary), per window
  std::vector<uint32_t> bounds;
  const uint32_t pc = (uint32_t)(((uint64_t)h->cols + P - 1) / P);
  const uint64_t nb = panel_bound_words(h, tile);
  if (nb) {
    const uint64_t blocks = (nb + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_panel_bounds, dim3((uint32_t)blocks), dim3(256), 0, s,
                       h->rwo.as<uint32_t>(), h->a2b.as<uint32_t>(), h->W, P,
                       pc, bounds_dev);
    ACC_HIP(hipGetLastError());
    bounds.resize(nb);
    ACC_HIP(hipMemcpyAsync(bounds.data(), bounds_dev, nb * 4, hipMemcpyDeviceToHost, s));
    ACC_HIP(hipStreamSynchronize(s));
  }

  std::vector<WorkUnit> units;
  std::vector<std::vector<WorkUnit>> by_panel(P);
  std::vector<uint32_t> sinfo;  // per split window: window, first slot, count
  units.reserve(h->W / 2 + 16);
  WorkUnit open{0, 0, 0, 0, -1, 0, 0, 0};
  auto close_open = [&]() {
    if (open.nwin) units.push_back(open);
    open.nwin = 0;
  };
  uint32_t pieces = 0, split_windows = 0;
  struct Cut { uint32_t a, b, panel; };
  std::vector<Cut> cuts;
  std::vector<uint32_t> edge(P + 1);
  for (uint32_t w = 0; w < h->W; ++w) {
    const uint32_t tb = rwo[w], te = rwo[w + 1], tw = te - tb;
    cuts.clear();
    if (P > 1 && tw >= 2 * kPanelMinTiles) {
      edge[0] = 0;
      for (uint32_t p = 1; p < P; ++p) {
        const uint32_t e = (bounds[(uint64_t)w * (P - 1) + p - 1] + 2u) & ~3u;
        edge[p] = std::min(tw, std::max(edge[p - 1], e));
      }
      edge[P] = tw;
      uint32_t a = 0;
      for (uint32_t p = 1; p <= P; ++p) {
        if (edge[p] - a < kPanelMinTiles && p < P) continue;
        if (edge[p] > a) {
          uint32_t panel = 0;
          while (panel + 1 < P && edge[panel + 1] <= a) ++panel;
          cuts.push_back(Cut{a, edge[p], panel});
        }
        a = edge[p];
      }
      if (cuts.size() >= 2 && cuts.back().b - cuts.back().a < kPanelMinTiles) {
        cuts[cuts.size() - 2].b = cuts.back().b;
        cuts.pop_back();
      }
      if (cuts.size() < 2) cuts.clear();
    }
    if (cuts.empty() && tw > split_tiles) cuts.push_back(Cut{0, tw, 0});
    if (!cuts.empty()) {
      close_open();
      // sub-split long pieces; count the window's pieces first
      auto npieces = [&](uint32_t len) {
        return (uint32_t)(((uint64_t)len + split_tiles - 1) / split_tiles);
      };
      uint32_t nwp = 0;
      for (const Cut& c : cuts) nwp += npieces(c.b - c.a);
      uint32_t slot = pieces;
      for (const Cut& c : cuts) {
        const uint32_t len = c.b - c.a;
        const uint32_t nu = npieces(len);
        std::vector<WorkUnit>& dst = P > 1 && cuts.size() >= 2 ? by_panel[c.panel] : units;
        for (uint32_t p = 0; p < nu; ++p) {
          uint32_t a = (uint32_t)((uint64_t)len * p / nu);
          uint32_t b = (uint32_t)((uint64_t)len * (p + 1) / nu);
          a = p == 0 ? 0 : std::min(len, (a + 3u) & ~3u);
          b = p + 1 == nu ? len : std::min(len, (b + 3u) & ~3u);
          b = std::max(a, b);
          dst.push_back(WorkUnit{w, 1, tb + c.a + a, tb + c.a + b, (int32_t)slot,
                                 (b - a + 3) / 4, c.panel, 0});
          ++slot;
        }
      }
      sinfo.push_back(w);
      sinfo.push_back(pieces);
      sinfo.push_back(nwp);
      pieces += nwp;
      ++split_windows;
      continue;
    }
    // empty windows never sit inside a unit that has tiles
    if (open.nwin && ((tw == 0) != (open.tile_end == open.tile_begin))) close_open();
    if (open.nwin == 0) open = WorkUnit{w, 0, tb, tb, -1, 0, 0, 0};
    open.nwin++;
    open.tile_end = te;
    open.ngroups += (tw + 3) / 4;
    if (open.tile_end - open.tile_begin >= pack_tiles) close_open();
  }
  close_open();
  sch->pin_len = 0;
  if (tile.xcd_slice == 3 && P > 1) {
    // pinned: XCD x gets the pieces of the panels p = x (mod 8) in window
    // order, then whole-window units where its tile count is lowest; the
    // eight lists are padded with no-op units to one length
    std::vector<std::vector<WorkUnit>> lists(8);
    std::vector<uint64_t> load(8, 0);
    for (uint32_t p = 0; p < P; ++p)
      for (const WorkUnit& u : by_panel[p]) {
        lists[p % 8].push_back(u);
        load[p % 8] += u.tile_end - u.tile_begin;
      }
    for (const WorkUnit& u : units) {
      const uint32_t x = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
      lists[x].push_back(u);
      load[x] += (u.tile_end - u.tile_begin) + u.nwin;
    }
    size_t len = 1;
    for (const auto& l : lists) len = std::max(len, l.size());
    if (len * 8 > 0xFFFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    std::vector<WorkUnit> pinned(len * 8, WorkUnit{0, 0, 0, 0, -1, 0, 0, 0});
    for (uint32_t x = 0; x < 8; ++x)
      std::copy(lists[x].begin(), lists[x].end(), pinned.begin() + (size_t)x * len);
    units = std::move(pinned);
    sch->pin_len = (uint32_t)len;
  } else {
    for (auto& v : by_panel) units.insert(units.end(), v.begin(), v.end());
  }
  sch->num_units = (uint32_t)units.size();
  sch->packed = false;
  for (const WorkUnit& u : units) sch->packed |= (u.slot < 0 && u.nwin > 1 && u.ngroups > 0);
  sch->h_units = std::move(units);
  sch->h_sinfo = std::move(sinfo);
  sch->num_pieces = pieces;
  sch->num_split_windows = split_windows;
  return ACCSPMM_OK;
}

// Device bytes of a schedule's work units and split-window records.
uint64_t schedule_unit_bytes(const Schedule& sch) {
  return (uint64_t)std::max<size_t>(sch.h_units.size(), 1) * sizeof(WorkUnit);
}
uint64_t schedule_sinfo_bytes(const Schedule& sch) {
  return (uint64_t)std::max<size_t>(sch.h_sinfo.size(), 3) * 4;
}

// Copies a schedule's work units and split-window records to device memory
// (units: schedule_unit_bytes, sinfo: schedule_sinfo_bytes); asynchronous,
// from the schedule's host vectors.
accspmm_status upload_schedule(const Schedule& sch, void* units, void* sinfo, hipStream_t s) {
  if (!sch.h_units.empty())
    ACC_HIP(hipMemcpyAsync(units, sch.h_units.data(), sch.h_units.size() * sizeof(WorkUnit),
                           hipMemcpyHostToDevice, s));
  if (!sch.h_sinfo.empty())
    ACC_HIP(hipMemcpyAsync(sinfo, sch.h_sinfo.data(), sch.h_sinfo.size() * 4,
                           hipMemcpyHostToDevice, s));
  return ACCSPMM_OK;
}

// The handle's cached schedule for the tile's (split, pack, panels), built
// and uploaded to the handle's own device buffers on first use.
accspmm_status get_schedule(const accspmm_metcf* h, const accspmm_tile& tile,
                            hipStream_t s, const Schedule** out) {
  const auto key = schedule_key(tile);
  auto it = h->schedules.find(key);
  if (it != h->schedules.end()) {
    *out = it->second.get();
    return ACCSPMM_OK;
  }
  auto sch = std::make_unique<Schedule>();
  {
    DevBuf bounds;
    const uint64_t nb = panel_bound_words(h, tile);
    if (nb) ACC_TRY(bounds.alloc(nb * 4));
    ACC_TRY(plan_schedule(h, tile, s, bounds.as<uint32_t>(), sch.get()));
  }
  ACC_TRY(sch->units.alloc(schedule_unit_bytes(*sch)));
  ACC_TRY(sch->split_info.alloc(schedule_sinfo_bytes(*sch)));
  ACC_TRY(upload_schedule(*sch, sch->units.p, sch->split_info.p, s));
  ACC_HIP(hipStreamSynchronize(s));
  *out = sch.get();
  h->schedules[key] = std::move(sch);
  return ACCSPMM_OK;
}

// The handle's cached schedule for the tile, or null (no allocation).
const Schedule* find_schedule(const accspmm_metcf* h, const accspmm_tile& tile) {
  auto it = h->schedules.find(schedule_key(tile));
  return it == h->schedules.end() ? nullptr : it->second.get();
}

namespace {

template <int FORM, int NB, int RW>
void launch_t(const SpmmArgs& g, uint32_t blocks, uint32_t wpb, bool packed,
              hipStream_t s) {
  const bool buf = gather_by_buffer(g, FORM);
  const dim3 grid(blocks), blk(64 * wpb);
  const size_t lds = (size_t)wpb * wave_lds<RW>() * sizeof(float);
  if (packed) {
    if (buf) hipLaunchKernelGGL((spmm_kernel<FORM, NB, RW, true, true>), grid, blk, lds, s, g);
    else hipLaunchKernelGGL((spmm_kernel<FORM, NB, RW, true, false>), grid, blk, lds, s, g);
  } else {
    if (buf) hipLaunchKernelGGL((spmm_kernel<FORM, NB, RW, false, true>), grid, blk, lds, s, g);
    else hipLaunchKernelGGL((spmm_kernel<FORM, NB, RW, false, false>), grid, blk, lds, s, g);
  }
}

template <int FORM, int RW>
accspmm_status launch_nb(uint32_t nb, const SpmmArgs& g, uint32_t blocks,
                         uint32_t wpb, bool packed, hipStream_t s) {
  switch (nb) {
    case 1: launch_t<FORM, 1, RW>(g, blocks, wpb, packed, s); break;
    case 2: launch_t<FORM, 2, RW>(g, blocks, wpb, packed, s); break;
    case 4: launch_t<FORM, 4, RW>(g, blocks, wpb, packed, s); break;
    case 8: launch_t<FORM, 8, RW>(g, blocks, wpb, packed, s); break;
    default: return ACCSPMM_ERR_UNSUPPORTED;
  }
  return ACCSPMM_OK;
}

accspmm_status launch_form(int form, uint32_t rw, uint32_t nb,
                           const SpmmArgs& g, uint32_t blocks, uint32_t wpb,
                           bool packed, hipStream_t s) {
  if (rw == 8) {
    if (form == kFP32) return launch_nb<kFP32, 8>(nb, g, blocks, wpb, packed, s);
    if (form == kSPLIT) return launch_nb<kSPLIT, 8>(nb, g, blocks, wpb, packed, s);
    if (form == kSPLITF) return launch_nb<kSPLITF, 8>(nb, g, blocks, wpb, packed, s);
    return launch_nb<kBF16, 8>(nb, g, blocks, wpb, packed, s);
  }
  if (form == kFP32) return launch_nb<kFP32, 16>(nb, g, blocks, wpb, packed, s);
  if (form == kSPLIT) return launch_nb<kSPLIT, 16>(nb, g, blocks, wpb, packed, s);
  if (form == kSPLITF) return launch_nb<kSPLITF, 16>(nb, g, blocks, wpb, packed, s);
  return launch_nb<kBF16, 16>(nb, g, blocks, wpb, packed, s);
}

}  // namespace

// Default tile, from sweeps on gfx950 (232,965-node power-law graph and
// 27-point stencil, K = 128):
//   * NB: the widest feature chunk up to K (BF16 form: 8, the others 4; NB=8
//     spills registers in the FP32 and split forms);
//   * one wave per block (smaller blocks balance better; no shape measured
//     faster);
//   * long windows (mean >= 64 tiles, e.g. power-law graphs): column panels
//     when B exceeds 8 MiB (2 panels; 4 above 32 MiB), round-robin XCDs
//     (split form: feature chunks pinned to XCDs and twice the panels);
//   * short windows (banded matrices): one contiguous range of windows per
//     XCD, so neighbouring windows share B rows in that XCD's L2;
//   * windows much longer than the mean are split; short ones are packed
//     on large graphs with 8-row windows;
//   * rows reordered by length (long windows first): fewer panels and a
//     split threshold near the mean window (split-BF16: 4 panels, feature
//     chunks pinned to XCDs; FP32: no panels, pieces of 2x the mean; BF16:
//     2 panels) — measured on the power-law graph at RW 8 and 16.
void default_tile(const accspmm_metcf* h, uint32_t K, int form,
                  accspmm_tile* t) {
  uint32_t nb = K >= 128 && form == kBF16 ? 8 : K >= 64 ? 4 : K >= 32 ? 2 : 1;
  t->feat_per_lane = nb;
  t->waves_per_block = 1;
  const uint32_t mean = h->W ? h->T / h->W : 0;
  const uint32_t s = std::max<uint32_t>(64, 4 * std::max<uint32_t>(mean, 16));
  t->split_tiles = h->max_window_tiles > s ? s : 0;
  t->pack_tiles = 64;
  const uint64_t kp = ((uint64_t)K + 16 * nb - 1) / (16 * nb) * (16 * nb);
  const uint64_t b_bytes = ((uint64_t)h->cols + 1) * kp * (form == kBF16 ? 2 : 4);
  if (mean >= 64 && h->row_perm.p) {
    if (K >= 128) {
      // one 128-feature chunk per wave (one wave per SIMD, all registers):
      // A and the tile metadata are read once per launch instead of once
      // per 64-feature chunk (232,965-node graph, K = 128: split-BF16 2.95
      // against 3.07 ms, FP32 3.54 against 3.64 ms)
      nb = 8;
      t->feat_per_lane = nb;
    }
    const uint64_t kp8 = ((uint64_t)K + 16 * nb - 1) / (16 * nb) * (16 * nb);
    const uint64_t bb = ((uint64_t)h->cols + 1) * kp8 * (form == kBF16 ? 2 : 4);
    const uint32_t n_kc = (uint32_t)(kp8 / (16 * nb));
    const uint32_t panels = bb > (32ull << 20) ? 4 : bb > (8ull << 20) ? 2 : 0;
    t->xcd_slice = 0;
    t->split_tiles = h->max_window_tiles > mean ? std::max<uint32_t>(64, mean) : 0;
    if ((form == kSPLIT || form == kBF16) && nb == 8 && n_kc >= 4) {
      // four or more 128-feature chunks in flight: 4 column panels and
      // pieces of 1.5x the mean window (K = 512: split-BF16 10.44 ms
      // against 10.87 with pieces of the mean and 11.96 with 2 panels,
      // BF16 6.15 against 6.23 and 6.90; at K = 256, two chunks, the
      // one-chunk choices stay best: 5.35 against 5.46 ms with 4 panels)
      t->col_panels = bb > (32ull << 20) ? 4 : 0;
      const uint32_t piece = std::max<uint32_t>(64, mean + mean / 2);
      t->split_tiles = h->max_window_tiles > piece ? piece : 0;
    } else if (form == kSPLIT && nb == 8) {
      // split-BF16, one chunk: 2 column panels (0 / 2 / 4 panels: 3.28 /
      // 2.95 / 3.12 ms at K = 128)
      t->col_panels = bb > (32ull << 20) ? 2 : 0;
    } else if (form == kSPLIT) {
      // split-BF16 (B 4 bytes a feature): 4 panels above 64 MiB of B, 2
      // above 32 MiB (K = 32 / 64 on the 232,965-node graph: 0 / 2 panels,
      // 1.16 / 1.69 ms, against 1.27 / 1.79 ms with the thresholds of the
      // other forms)
      t->col_panels = bb > (64ull << 20) ? 4 : bb > (32ull << 20) ? 2 : 0;
      if (n_kc >= 2 && 8 % n_kc == 0) t->xcd_slice = 1;
    } else if (form == kFP32) {
      t->col_panels = 0;
      t->split_tiles = h->max_window_tiles > 2 * mean ? std::max<uint32_t>(64, 2 * mean) : 0;
    } else {
      t->col_panels = std::min<uint32_t>(panels, 2);
    }
  } else if (mean >= 64) {
    t->xcd_slice = 0;
    t->col_panels = b_bytes > (32ull << 20) ? 4 : b_bytes > (8ull << 20) ? 2 : 0;
    // split form: pin feature chunks to XCDs (each L2 then holds half of a
    // B row per chunk) and halve the panels' width (measured 2% faster)
    const uint32_t n_kc = (uint32_t)(kp / (16 * nb));
    if (form == kSPLIT && t->col_panels && n_kc >= 2 && 8 % n_kc == 0) {
      t->xcd_slice = 1;
      t->col_panels *= 2;
    }
  } else {
    // short windows: a banded graph runs one contiguous range of work units
    // per XCD, so neighbouring windows share their B rows in its L2 (27-point
    // stencil 100^3, split-BF16 K = 128: RW 16 0.558 against 0.614 ms
    // round-robin, RW 8 0.652 against 0.751); otherwise round-robin
    // (2M-node power-law graph, RW 8: 4.121 against 4.271 ms)
    t->xcd_slice = h->banded ? 2 : 0;
    t->col_panels = 0;
    // packing short windows into units of >= 64 tiles shortens the per-unit
    // pipeline start-up, which pays only while enough units remain to fill
    // the GPU and not with 16-row windows (split-BF16, K=128: 27-point
    // stencil 100^3 RW 8 0.674 vs 0.729 ms unpacked, RW 16 0.659 vs 0.591;
    // 20,000-node graph 0.054 vs 0.047; 3,000-node graph 0.037 vs 0.020)
    t->pack_tiles = h->rw == 8 && (uint64_t)h->T >= 64ull * 8192 ? 64 : 0;
  }
}

// Split-BF16 B path when the tile leaves it open.  The conversion pass costs
// one read and one write of B; splitting in registers costs four VALU
// operations per pair of gathered elements instead, i.e. per tile column.
// With few gathers per B row (banded matrices: ~14 on the 27-point stencil)
// the in-register split wins; with many (~380 on the 232,965-node power-law
// graph) the conversion pass does.  Threshold: 64 gathered tile columns per
// B row.
uint32_t default_b_split(const accspmm_metcf* h, uint32_t K) {
  (void)K;  // both costs scale with K
  return (uint64_t)h->T * kColWindow < 64ull * h->cols ? 2u : 1u;
}

// The single-graph SpMM kernel over g's work units: blocks of wpb waves,
// mapped to XCDs as g.xcd_slice says (multiples of 8 blocks for modes 1, 2).
accspmm_status launch_single(const SpmmArgs& g, int form, uint32_t rw, uint32_t nb,
                             uint32_t wpb, bool packed, uint32_t num_units, uint32_t n_kc,
                             hipStream_t s) {
  uint64_t blocks;
  if (g.xcd_slice == 1) {
    const uint64_t ub = (num_units + wpb - 1) / wpb;
    blocks = (ub * n_kc + 7) / 8 * 8;
  } else if (g.xcd_slice == 2) {
    blocks = ((uint64_t)num_units * n_kc + wpb - 1) / wpb;
    blocks = (blocks + 7) / 8 * 8;
  } else if (g.xcd_slice == 3) {
    blocks = 8 * (((uint64_t)g.pin_len * n_kc + wpb - 1) / wpb);
  } else {
    blocks = ((uint64_t)num_units * n_kc + wpb - 1) / wpb;
  }
  if (blocks == 0) return ACCSPMM_OK;
  if (blocks > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  ACC_TRY(launch_form(form, rw, nb, g, (uint32_t)blocks, wpb, packed, s));
  ACC_HIP(hipGetLastError());
  return ACCSPMM_OK;
}

// sweep.hip: the resident-window sweep over the same pieces
accspmm_status sweep_launch(const accspmm_metcf* h, const dev::SpmmArgs& g,
                            const accspmm_tile& tile, uint32_t nb, int form,
                            const Schedule* sch, hipStream_t s);

// Workspace bytes of one call: the split-window partials (pieces x RW x
// KC-padded K floats) and, for the BF16 form, the converted B ((cols + 1)
// rows of KC-padded K bf16, the last all zero), each rounded up to 256.
void spmm_workspace_need(const accspmm_metcf* h, const Schedule* sch, uint32_t K, int form,
                         const accspmm_tile& tile, uint64_t* partial_bytes,
                         uint64_t* bconv_bytes) {
  const uint32_t KC = 16 * tile.feat_per_lane;
  const uint64_t kcp = (uint64_t)((K + KC - 1) / KC) * KC;
  const int kform = kernel_form(form, tile);
  const uint64_t pb = tile.sweep ? 0 : (uint64_t)sch->num_pieces * h->rw * kcp * 4;
  // the FP32-gathering forms read the caller's B, or a permuted copy when
  // the graph has a column order
  const uint64_t cb = raw_b(kform) && !h->col_perm.p
                          ? 0
                          : ((uint64_t)h->cols + 1) * kcp * (kform == kBF16 ? 2 : 4);
  *partial_bytes = (pb + 255) / 256 * 256;
  *bconv_bytes = (cb + 255) / 256 * 256;
}

// The handle's workspace for calls on stream s (allocated on first use,
// grown as needed; caller holds h->mu).  Calls on one stream are ordered,
// so they may share it; calls on different streams never do.
accspmm_status stream_workspace(const accspmm_metcf* h, hipStream_t s, uint64_t partial_bytes,
                                uint64_t bconv_bytes, SpmmWorkspace* out) {
  std::unique_ptr<StreamWorkspace>& w = h->stream_ws[s];
  if (!w) w = std::make_unique<StreamWorkspace>();
  if (partial_bytes) ACC_TRY(w->partials.reserve(partial_bytes));
  if (bconv_bytes) ACC_TRY(w->bconv.reserve(bconv_bytes));
  out->partials = w->partials.as<float>();
  out->bconv = w->bconv.p;
  return ACCSPMM_OK;
}

// One SpMM call on a given schedule: d_units / d_sinfo are its work units
// and split-window records in device memory, ws the call's workspace (at
// least spmm_workspace_need's bytes).  phases: bit 0 = convert B (BF16
// form), bit 1 = SpMM kernel.  Caller holds h->mu.  Allocates nothing except
// through the sweep (tile.sweep), whose plan the handle caches.
accspmm_status spmm_run(const accspmm_metcf* h, const Schedule* sch, const WorkUnit* d_units,
                        const uint32_t* d_sinfo, const float* B, uint32_t ldb, float* C,
                        uint32_t ldc, uint32_t K, int form, const accspmm_tile& tile,
                        hipStream_t s, int phases, const SpmmWorkspace& w) {
  const uint32_t nb = tile.feat_per_lane;
  const int kform = kernel_form(form, tile);
  uint32_t wpb = tile.waves_per_block ? tile.waves_per_block : 4;
  // NB = 8 kernels run one wave per SIMD (blocks of at most 4 waves), which
  // leaves them the whole register file, accumulation registers included
  if (nb == 8) wpb = std::min<uint32_t>(wpb, 4);
  const uint32_t KC = 16 * nb;
  const uint32_t n_kc = (K + KC - 1) / KC;
  uint64_t pbytes = 0, cbytes = 0;
  spmm_workspace_need(h, sch, K, form, tile, &pbytes, &cbytes);

  SpmmArgs g{};
  g.rwo = h->rwo.as<uint32_t>();
  g.tco = h->tco.as<uint32_t>();
  g.bits = h->bits.as<unsigned long long>();
  g.a2b = h->a2b.as<uint32_t>();
  g.data = h->data.as<float>();
  g.units = d_units;
  g.split_info = d_sinfo;
  g.C = C;
  g.row_map = h->row_perm.p ? h->row_perm.as<uint32_t>() : nullptr;
  g.ldc = ldc;
  g.rows = h->rows;
  g.K = K;
  g.num_units = sch->num_units;
  g.n_kc = n_kc;
  g.kc_padded = n_kc * KC;
  g.xcd_slice = sch->pin_len ? 3u
                : tile.xcd_slice == 2 ? 2u
                : (tile.xcd_slice == 1 && (8 % n_kc) == 0) ? 1u : 0u;
  g.pin_len = sch->pin_len;
  g.zero_row = h->cols;
  g.b_rows = raw_b(kform) ? h->cols : h->cols + 1;
  g.pad_tile = h->T;
  g.wide = h->rw == 16 && h->max_group_values > dev::kSpanValues ? 1u : 0u;
  g.partials = pbytes ? w.partials : nullptr;

  const uint32_t* perm = h->col_perm.p ? h->col_perm.as<uint32_t>() : nullptr;
  if (raw_b(kform) && !perm) {
    g.B = B;
    g.ldb = ldb;
  } else {
    const uint32_t ldx = n_kc * KC;
    const uint64_t chunks = ((uint64_t)h->cols + 1) * (ldx / 4);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((chunks + 255) / 256, 8192);
    const uint32_t vec = (ldb % 4 == 0 && ((uintptr_t)B % 16) == 0) ? 1u : 0u;
    if (phases & 1) {
      if (raw_b(kform))
        hipLaunchKernelGGL(convert_b<kFP32>, dim3(blocks), dim3(256), 0, s, B, h->cols, K, ldb,
                           w.bconv, ldx, vec, perm);
      else if (kform == kSPLIT)
        hipLaunchKernelGGL(convert_b<kSPLIT>, dim3(blocks), dim3(256), 0, s, B, h->cols, K, ldb,
                           w.bconv, ldx, vec, perm);
      else
        hipLaunchKernelGGL(convert_b<kBF16>, dim3(blocks), dim3(256), 0, s, B, h->cols, K, ldb,
                           w.bconv, ldx, vec, perm);
      ACC_HIP(hipGetLastError());
    }  // else converted by an earlier prepare call on this stream
    g.B = w.bconv;
    g.ldb = ldx;
  }
  if (!(phases & 2)) return ACCSPMM_OK;
  if (tile.sweep) return sweep_launch(h, g, tile, nb, kform, sch, s);

  ACC_TRY(launch_single(g, kform, h->rw, nb, wpb, sch->packed, sch->num_units, n_kc, s));
  if (sch->num_split_windows) {
    if (h->rw == 8)
      hipLaunchKernelGGL(k_reduce_split<8>, dim3(sch->num_split_windows), dim3(256), 0, s,
                         d_sinfo, g.partials, g.kc_padded, K, C, ldc, h->rows, g.row_map);
    else
      hipLaunchKernelGGL(k_reduce_split<16>, dim3(sch->num_split_windows), dim3(256), 0, s,
                         d_sinfo, g.partials, g.kc_padded, K, C, ldc, h->rows, g.row_map);
    ACC_HIP(hipGetLastError());
  }
  return ACCSPMM_OK;
}

// Operand checks shared by the entry points: NB-wide vector loads of the
// caller's FP32 B and stores of C need their alignment.
accspmm_status spmm_check_operands(const float* B, uint32_t ldb, const float* C, uint32_t ldc,
                                   int form, const accspmm_tile& tile, int phases) {
  const uint32_t nb = tile.feat_per_lane;
  if (nb != 1 && nb != 2 && nb != 4 && nb != 8) return ACCSPMM_ERR_UNSUPPORTED;
  if (tile.waves_per_block > 8) return ACCSPMM_ERR_UNSUPPORTED;
  if (raw_b(kernel_form(form, tile))) {
    if ((ldb % nb) != 0 || ((uintptr_t)B % (4u * nb)) != 0) return ACCSPMM_ERR_UNSUPPORTED;
  }
  if ((phases & 2) && ((ldc % nb) != 0 || ((uintptr_t)C % (4u * nb)) != 0))
    return ACCSPMM_ERR_UNSUPPORTED;
  return ACCSPMM_OK;
}

// accspmm_spmm's body: the handle's cached schedule and, unless ws is
// given, the handle's workspace of stream s.
accspmm_status spmm_launch(const accspmm_metcf* h, const float* B, uint32_t ldb,
                           float* C, uint32_t ldc, uint32_t K, int form,
                           const accspmm_tile& tile, hipStream_t s,
                           int phases, const SpmmWorkspace* ws) {
  ACC_TRY(spmm_check_operands(B, ldb, C, ldc, form, tile, phases));
  std::lock_guard<std::mutex> lock(h->mu);
  const Schedule* sch = nullptr;
  ACC_TRY(get_schedule(h, tile, s, &sch));
  uint64_t pbytes = 0, cbytes = 0;
  spmm_workspace_need(h, sch, K, form, tile, &pbytes, &cbytes);
  SpmmWorkspace w{};
  if (ws) {
    w = *ws;
  } else {
    ACC_TRY(stream_workspace(h, s, (phases & 2) ? pbytes : 0, cbytes, &w));
  }
  return spmm_run(h, sch, sch->units.as<WorkUnit>(), sch->split_info.as<uint32_t>(), B, ldb,
                  C, ldc, K, form, tile, s, phases, w);
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
  if (form != kSPLIT) t.b_split = 0;
  else if (t.b_split == 0 || t.b_split > 2) t.b_split = default_b_split(h, K);
  if (t.feat_per_lane == 8 && t.waves_per_block > 4) t.waves_per_block = 4;
  while (t.feat_per_lane > 1 &&
         ((form == kFP32 &&
           ((ldb % t.feat_per_lane) || ((uintptr_t)B % (4u * t.feat_per_lane)))) ||
          (C && ((ldc % t.feat_per_lane) || ((uintptr_t)C % (4u * t.feat_per_lane))))))
    t.feat_per_lane /= 2;
  // the in-register split reads B like the FP32 form; otherwise convert
  if (t.b_split == 2 && ((ldb % t.feat_per_lane) || ((uintptr_t)B % (4u * t.feat_per_lane))))
    t.b_split = 1;
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
                     phases, nullptr);
}

extern "C" accspmm_status accspmm_spmm(const accspmm_metcf* h, const float* B,
                                       uint32_t ldb, float* C, uint32_t ldc,
                                       uint32_t K, accspmm_form form,
                                       const accspmm_tile* tile,
                                       void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    return spmm_entry(h, B, ldb, C, ldc, K, form, tile, stream, 3);
  });
}

extern "C" accspmm_status accspmm_spmm_prepare_b(const accspmm_metcf* h,
                                                 const float* B, uint32_t ldb,
                                                 uint32_t K, accspmm_form form,
                                                 const accspmm_tile* tile,
                                                 void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    return spmm_entry(h, B, ldb, nullptr, K, K, form, tile, stream, 1);
  });
}

extern "C" accspmm_status accspmm_spmm_prepared(const accspmm_metcf* h,
                                                const float* B, uint32_t ldb,
                                                float* C, uint32_t ldc,
                                                uint32_t K, accspmm_form form,
                                                const accspmm_tile* tile,
                                                void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    return spmm_entry(h, B, ldb, C, ldc, K, form, tile, stream, 2);
  });
}
