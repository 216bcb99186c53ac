// sweep.hip — resident-window sweep of the METCF SpMM (gfx950).
//
// The work-unit kernel (spmm.hip) runs the column-panel pieces of a window
// as independent units: their partial sums go to HBM and a second launch
// adds them.  Here one block per CU stays resident; wave v of round r owns
// up to M windows and keeps each window's running sum in an LDS slot.  It
// runs its pieces in (panel, window, slot) order, so all waves gather from
// the same column panel of B at about the same time and re-read that panel
// from L2, and only C goes to HBM.
//
// The pieces, their 4-tile groups and the summation order are those of the
// work-unit schedule with the same (split, pack, panels): a piece
// accumulates from zero (the split form's two A-term rows combined at its
// end), a window of one piece stores that sum, and a split window's sum is
// 0 + s_0 + s_1 + ... in slot order, as k_reduce_split computes it.  Both
// paths therefore give the same C bits, and the batched kernel (which runs
// work units) stays bit-equal to a standalone sweep.
#include <algorithm>
#include <numeric>
#include <vector>

#include "internal.hpp"

#include "spmm_device.hpp"

namespace accspmm {
namespace {
using namespace dev;

struct SweepArgs {
  SpmmArgs g;
  const uint32_t* __restrict__ pieces;   // per piece: tile_begin, tile_end, slot | split << 8
  const uint32_t* __restrict__ poff;     // [R*NW + 1]
  const uint32_t* __restrict__ ngroups;  // [R*NW]
  const uint32_t* __restrict__ wlist;    // [R*NW*M]
  uint32_t M, NW, R;
};

// Lanes that hold a window's rows after the piece's final combine: RW=8
// keeps rows 0-7 in quarters 0-1 (lanes 0-31); RW=16 uses all 64 lanes.
template <int RW>
constexpr uint32_t slot_lanes() { return RW == 8 ? 32u : 64u; }

// Ends a piece: combines the split form's two A-term rows and adds the piece sum
// into the window's LDS slot (first piece: stored, or 0 + s for a split
// window, matching the work-unit path's direct store and reduction).
template <int FORM, int NB, int RW>
__device__ inline void finish_piece(f4* acc, f4* slot, uint32_t lane, bool first,
                                    bool split) {
  if (is_split(FORM) && RW == 8) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[b][i] += __shfl_xor(acc[b][i], 32, 64);
  }
  constexpr uint32_t NL = slot_lanes<RW>();
  if (lane < NL) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      f4* p = slot + b * NL + lane;
      f4 v = acc[b];
      if (!first) v = *p + v;
      else if (split) v = f4{0.f, 0.f, 0.f, 0.f} + v;
      *p = v;
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
}

// Writes a window's rows from its LDS slot to C.
template <int FORM, int NB, int RW>
__device__ inline void store_slot(const f4* slot, uint32_t lane, uint32_t q,
                                  uint32_t w, uint32_t feat, uint32_t K,
                                  float* C, uint32_t ldc, uint32_t rows,
                                  const uint32_t* rmap) {
  constexpr uint32_t NL = slot_lanes<RW>();
  if (lane >= NL) return;
  f4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = slot[b * NL + lane];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t r = w * RW + 4 * q + i;
    if (r < rows) {
      float v[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) v[b] = acc[b][i];
      store_lane<FORM, NB>(C + out_row(rmap, r) * ldc, feat, K, v);
    }
  }
}

// Next group of the wave's piece list (only called while groups remain;
// the list holds no empty piece).  GroupPos::w is the piece index.
__device__ inline void sweep_next(GroupPos& p, const uint32_t* __restrict__ pieces) {
  p.t += 4;
  if (p.t >= p.tend) {
    ++p.w;
    p.t = __builtin_amdgcn_readfirstlane(pieces[3ull * p.w]);
    p.tend = __builtin_amdgcn_readfirstlane(pieces[3ull * p.w + 1]);
  }
}

// The work-unit kernel's three-stage pipeline (metadata g+2, gathers g+1,
// MFMAs g) over the wave's piece list; a group of a new piece first ends
// the previous piece into its window's LDS slot.
template <int FORM, int NB, int RW, bool FULL, bool MASK>
__device__ void sweep_loop(const SweepArgs& a, f4* acc, f4* slots, GroupPos pos,
                           uint32_t ngroups, uint32_t lane, uint32_t q,
                           uint32_t m, bool a_rows, uint32_t feat,
                           uint32_t& started, float* lv) {
  const SpmmArgs& g = a.g;
  constexpr uint32_t NL = slot_lanes<RW>();
  // the sweep gathers B with 64-bit addresses (the same C bits as the
  // buffer loads of the work-unit kernel: padded columns contribute zeros
  // either way)
  const BSrc src = make_bsrc<FORM, NB, false>(g.B, g.ldb, g.b_rows, g.zero_row);
  uint32_t cur = pos.w;
  auto end_piece = [&](uint32_t pi) {
    const uint32_t info = __builtin_amdgcn_readfirstlane(a.pieces[3ull * pi + 2]);
    const uint32_t s = info & 0xFFu;
    finish_piece<FORM, NB, RW>(acc, slots + (size_t)s * NB * NL, lane,
                               ((started >> s) & 1u) == 0u, (info >> 8) & 1u);
    started |= 1u << s;
  };
  Meta mA, mB;
  Stage<FORM, NB> dA, dB;
  const uint32_t* vsrc = reinterpret_cast<const uint32_t*>(g.data);
  const MetaSrc ms = meta_src<RW>(g, lane);
  auto consume = [&](Stage<FORM, NB>& d) {
    if (d.win != cur) {
      end_piece(cur);
      cur = d.win;
    }
    compute_stage<FORM, NB, RW, MASK>(d, acc, m, lv, lane, vsrc, g.wide != 0u);
  };
  load_meta<RW>(mA, g, pos, true, q, ms);
  if (ngroups > 1) sweep_next(pos, a.pieces);
  load_meta<RW>(mB, g, pos, ngroups > 1, q, ms);
  load_data<FORM, NB, FULL, RW, false>(dA, mA, g, m, a_rows, feat, src, lane);
  uint32_t gi = 0;
  for (; gi + 3 < ngroups; gi += 2) {
    sweep_next(pos, a.pieces);  // group gi + 2
    load_meta<RW>(mA, g, pos, true, q, ms);
    load_data<FORM, NB, FULL, RW, false>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    __builtin_amdgcn_sched_barrier(0);
    const bool v3 = gi + 3 < ngroups;
    if (v3) sweep_next(pos, a.pieces);  // group gi + 3
    load_meta<RW>(mB, g, pos, v3, q, ms);
    load_data<FORM, NB, FULL, RW, false>(dA, mA, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dB);
    __builtin_amdgcn_sched_barrier(0);
  }
  const uint32_t left = ngroups - gi;  // 1, 2 or 3
  if (left == 3) {
    sweep_next(pos, a.pieces);
    load_meta<RW>(mA, g, pos, true, q, ms);
    load_data<FORM, NB, FULL, RW, false>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    __builtin_amdgcn_sched_barrier(0);
    load_data<FORM, NB, FULL, RW, false>(dA, mA, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dB);
    consume(dA);
  } else if (left == 2) {
    load_data<FORM, NB, FULL, RW, false>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    consume(dB);
  } else {
    consume(dA);
  }
  end_piece(cur);
}

// One block per CU, wpb waves; dynamic LDS holds wpb * M slots of
// NB * slot_lanes f4 each.
template <int FORM, int NB, int RW>
__global__ __launch_bounds__(512) void sweep_kernel(SweepArgs a) {
  extern __shared__ f4 lds_slots[];
  __shared__ __align__(16) float lds_vals[8 * wave_lds<RW>()];
  const SpmmArgs& g = a.g;
  constexpr uint32_t KC = 16u * NB;
  constexpr uint32_t NL = slot_lanes<RW>();
  const uint32_t wpb = blockDim.x >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t v = blockIdx.x * wpb + wv;
  if (v >= a.NW) return;
  const uint32_t lane = threadIdx.x & 63u, q = lane >> 4, n = lane & 15u;
  const uint32_t m = n;
  const bool a_rows = RW == 16 || m < 8 || is_split(FORM);
  f4* slots = lds_slots + (size_t)wv * a.M * NB * NL;
  float* lv = lds_vals + wv * wave_lds<RW>();
  if (lane == 0) lv[group_values<RW>()] = 0.f;
  for (uint32_t r = 0; r < a.R; ++r) {
    const uint32_t idx = r * a.NW + v;
    const uint32_t pb = __builtin_amdgcn_readfirstlane(a.poff[idx]);
    const uint32_t ng = __builtin_amdgcn_readfirstlane(a.ngroups[idx]);
    const uint32_t* wl = a.wlist + (size_t)idx * a.M;
    for (uint32_t kc = 0; kc < g.n_kc; ++kc) {
      const uint32_t k0 = kc * KC, feat = lane_feat<FORM, NB>(k0, n);
      // B loads are full-width unless the chunk straddles K (FP32 B)
      const bool full = !raw_b(FORM) || k0 + KC <= g.K;
      uint32_t started = 0;
      if (ng) {
        f4 acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
        GroupPos pos{pb, (uint32_t)__builtin_amdgcn_readfirstlane(a.pieces[3ull * pb]),
                     (uint32_t)__builtin_amdgcn_readfirstlane(a.pieces[3ull * pb + 1])};
        const bool mask = row0_needs_mask<FORM, NB>(g.B, feat, g.K);
        if (full && mask)
          sweep_loop<FORM, NB, RW, true, true>(a, acc, slots, pos, ng, lane, q, m, a_rows, feat,
                                               started, lv);
        else if (full)
          sweep_loop<FORM, NB, RW, true, false>(a, acc, slots, pos, ng, lane, q, m, a_rows,
                                                feat, started, lv);
        else if (mask)
          sweep_loop<FORM, NB, RW, false, true>(a, acc, slots, pos, ng, lane, q, m, a_rows,
                                                feat, started, lv);
        else
          sweep_loop<FORM, NB, RW, false, false>(a, acc, slots, pos, ng, lane, q, m, a_rows,
                                                 feat, started, lv);
      }
      for (uint32_t s = 0; s < a.M; ++s) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(wl[s]);
        if (w == kPad) continue;
        if ((started >> s) & 1u)
          store_slot<FORM, NB, RW>(slots + (size_t)s * NB * NL, lane, q, w, feat, g.K, g.C,
                             g.ldc, g.rows, g.row_map);
        else
          zero_window<FORM, NB, RW>(w, q, feat, g.K, g.C, g.ldc, g.rows, g.row_map);
      }
    }
  }
}

// running-sum slots; with the per-wave value buffers (16.5 KiB for 16-row
// windows) the block stays within the CU's 160 KiB of LDS
constexpr uint32_t kSweepLdsBytes = 136u << 10;

template <int FORM, int NB, int RW>
accspmm_status launch_sweep_t(const SweepArgs& a, uint32_t blocks, uint32_t wpb,
                              uint32_t lds, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    ACC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&sweep_kernel<FORM, NB, RW>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kSweepLdsBytes));
    attr_set = true;
  }
  hipLaunchKernelGGL((sweep_kernel<FORM, NB, RW>), dim3(blocks), dim3(64 * wpb), lds, s, a);
  ACC_HIP(hipGetLastError());
  return ACCSPMM_OK;
}

template <int FORM, int RW>
accspmm_status launch_sweep_nb(uint32_t nb, const SweepArgs& a, uint32_t blocks,
                               uint32_t wpb, uint32_t lds, hipStream_t s) {
  switch (nb) {
    case 1: return launch_sweep_t<FORM, 1, RW>(a, blocks, wpb, lds, s);
    case 2: return launch_sweep_t<FORM, 2, RW>(a, blocks, wpb, lds, s);
    case 4: return launch_sweep_t<FORM, 4, RW>(a, blocks, wpb, lds, s);
    case 8: return launch_sweep_t<FORM, 8, RW>(a, blocks, wpb, lds, s);
  }
  return ACCSPMM_ERR_UNSUPPORTED;
}

}  // namespace

// Regroups the schedule's pieces per wave: windows sorted by tile count
// (heaviest first) are dealt to the NW waves in rounds of M per wave, in
// boustrophedon order so every round is balanced; a wave's pieces of one
// round are listed by (panel, window slot, piece slot).  Windows that the
// schedule does not cut are single pieces (panel 0).
accspmm_status get_sweep_plan(const accspmm_metcf* h, const Schedule* sch,
                              const accspmm_tile& tile, uint32_t NW, uint32_t M,
                              hipStream_t s, const SweepPlan** out) {
  const auto key = std::make_tuple(sch->split_tiles, sch->pack_tiles, sch->col_panels, NW, M);
  auto it = h->sweep_plans.find(key);
  if (it != h->sweep_plans.end()) {
    *out = it->second.get();
    return ACCSPMM_OK;
  }
  (void)tile;
  const uint32_t W = h->W;
  const std::vector<uint32_t>& rwo = h->h_rwo;
  struct Piece { uint32_t tb, te, panel; int32_t slot; };
  std::vector<std::vector<Piece>> per_w(W);
  std::vector<uint8_t> split(W, 0);
  for (const WorkUnit& u : sch->h_units) {
    if (u.slot >= 0) {
      split[u.window] = 1;
      if (u.tile_end > u.tile_begin)
        per_w[u.window].push_back(Piece{u.tile_begin, u.tile_end, u.panel, u.slot});
    } else {
      for (uint32_t w = u.window; w < u.window + u.nwin; ++w)
        if (rwo[w + 1] > rwo[w]) per_w[w].push_back(Piece{rwo[w], rwo[w + 1], 0, -1});
    }
  }
  for (auto& v : per_w)
    std::sort(v.begin(), v.end(), [](const Piece& x, const Piece& y) { return x.slot < y.slot; });
  std::vector<uint32_t> order(W);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return rwo[a + 1] - rwo[a] > rwo[b + 1] - rwo[b];
  });
  const uint64_t per_round = (uint64_t)NW * M;
  const uint32_t R = (uint32_t)std::max<uint64_t>(1, (W + per_round - 1) / per_round);
  std::vector<uint32_t> wlist((size_t)R * NW * M, kPad);
  for (uint32_t i = 0; i < W; ++i) {
    const uint32_t r = (uint32_t)(i / per_round), j = (uint32_t)(i % per_round);
    const uint32_t row = j / NW, col = j % NW;
    const uint32_t v = (row & 1u) ? NW - 1 - col : col;
    wlist[((size_t)r * NW + v) * M + row] = order[i];
  }
  struct Key { uint32_t panel, s, slot_rank, tb, te; };
  std::vector<uint32_t> pieces, poff((size_t)R * NW + 1, 0), ngroups((size_t)R * NW, 0);
  std::vector<Key> keys;
  for (uint32_t idx = 0; idx < R * NW; ++idx) {
    keys.clear();
    for (uint32_t sl = 0; sl < M; ++sl) {
      const uint32_t w = wlist[(size_t)idx * M + sl];
      if (w == kPad) continue;
      uint32_t rank = 0;
      for (const Piece& p : per_w[w]) keys.push_back(Key{p.panel, sl, rank++, p.tb, p.te});
    }
    std::sort(keys.begin(), keys.end(), [](const Key& x, const Key& y) {
      return x.panel != y.panel ? x.panel < y.panel
                                : x.s != y.s ? x.s < y.s : x.slot_rank < y.slot_rank;
    });
    poff[idx] = (uint32_t)(pieces.size() / 3);
    uint32_t ng = 0;
    for (const Key& k : keys) {
      const uint32_t w = wlist[(size_t)idx * M + k.s];
      pieces.push_back(k.tb);
      pieces.push_back(k.te);
      pieces.push_back(k.s | ((uint32_t)split[w] << 8));
      ng += (k.te - k.tb + 3) / 4;
    }
    ngroups[idx] = ng;
  }
  poff[(size_t)R * NW] = (uint32_t)(pieces.size() / 3);
  if (pieces.size() / 3 > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  auto plan = std::make_unique<SweepPlan>();
  plan->P = sch->col_panels;
  plan->M = M;
  plan->NW = NW;
  plan->R = R;
  auto up = [&](DevBuf& d, const std::vector<uint32_t>& v) -> accspmm_status {
    ACC_TRY(d.alloc(std::max<size_t>(v.size(), 4) * 4));
    if (!v.empty())
      ACC_HIP(hipMemcpyAsync(d.p, v.data(), v.size() * 4, hipMemcpyHostToDevice, s));
    return ACCSPMM_OK;
  };
  ACC_TRY(up(plan->pieces, pieces));
  ACC_TRY(up(plan->poff, poff));
  ACC_TRY(up(plan->ngroups, ngroups));
  ACC_TRY(up(plan->wlist, wlist));
  ACC_HIP(hipStreamSynchronize(s));
  *out = plan.get();
  h->sweep_plans[key] = std::move(plan);
  return ACCSPMM_OK;
}

// Launches the sweep for arguments prepared by spmm_launch (B converted,
// C, K, chunks): one block of wpb waves per CU.
accspmm_status sweep_launch(const accspmm_metcf* h, const dev::SpmmArgs& g,
                            const accspmm_tile& tile, uint32_t nb, int form,
                            const Schedule* sch, hipStream_t s) {
  uint32_t wpb = tile.waves_per_block;
  if (wpb != 4 && wpb != 8) wpb = 8;
  int dv = 0, cus = 0;
  ACC_HIP(hipGetDevice(&dv));
  ACC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dv));
  if (cus <= 0) return ACCSPMM_ERR_HIP;
  const uint32_t NL = h->rw == 8 ? 32u : 64u;
  const uint32_t slot_bytes = nb * NL * 16u;
  const uint32_t M = std::min<uint32_t>(32, kSweepLdsBytes / (wpb * slot_bytes));
  if (M == 0) return ACCSPMM_ERR_UNSUPPORTED;
  const uint32_t NW = (uint32_t)cus * wpb;
  const SweepPlan* plan = nullptr;
  ACC_TRY(get_sweep_plan(h, sch, tile, NW, M, s, &plan));
  SweepArgs a{};
  a.g = g;
  a.pieces = plan->pieces.as<uint32_t>();
  a.poff = plan->poff.as<uint32_t>();
  a.ngroups = plan->ngroups.as<uint32_t>();
  a.wlist = plan->wlist.as<uint32_t>();
  a.M = M;
  a.NW = NW;
  a.R = plan->R;
  const uint32_t lds = wpb * M * slot_bytes;
  if (h->rw == 8) {
    if (form == kFP32) return launch_sweep_nb<kFP32, 8>(nb, a, cus, wpb, lds, s);
    if (form == kSPLIT) return launch_sweep_nb<kSPLIT, 8>(nb, a, cus, wpb, lds, s);
    if (form == kSPLITF) return launch_sweep_nb<kSPLITF, 8>(nb, a, cus, wpb, lds, s);
    return launch_sweep_nb<kBF16, 8>(nb, a, cus, wpb, lds, s);
  }
  if (form == kFP32) return launch_sweep_nb<kFP32, 16>(nb, a, cus, wpb, lds, s);
  if (form == kSPLIT) return launch_sweep_nb<kSPLIT, 16>(nb, a, cus, wpb, lds, s);
  if (form == kSPLITF) return launch_sweep_nb<kSPLITF, 16>(nb, a, cus, wpb, lds, s);
  return launch_sweep_nb<kBF16, 16>(nb, a, cus, wpb, lds, s);
}

}  // namespace accspmm
