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

namespace accspmm {
namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef __bf16 bf16x2 __attribute__((__vector_size__(2 * sizeof(__bf16))));
typedef float f2 __attribute__((__vector_size__(2 * sizeof(float))));
typedef float f4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef uint32_t u4 __attribute__((__vector_size__(4 * sizeof(uint32_t))));

enum : int { kFP32 = 0, kSPLIT = 1, kBF16 = 2 };

struct SpmmArgs {
  const uint32_t* __restrict__ tco;
  const unsigned long long* __restrict__ bits;
  const uint32_t* __restrict__ a2b;
  const float* __restrict__ data;
  const WorkUnit* __restrict__ units;
  const uint32_t* __restrict__ piece_info;  // per piece: (first slot, count, counter id)
  uint32_t* __restrict__ counters;
  float* __restrict__ partials;
  const void* __restrict__ B;  // FP32: float; SPLIT: u32 (hi|lo<<16); BF16: u16
  float* __restrict__ C;
  uint32_t ldb, ldc;
  uint32_t rows, K;
  uint32_t num_units;
  uint32_t n_kc;  // feature chunks
  uint32_t kc_padded;  // KC * n_kc (partials row stride)
  uint32_t xcd_slice;
};

__device__ inline uint32_t popc64(unsigned long long x) { return __popcll(x); }

// bf16 round-to-nearest-even pair
__device__ inline uint32_t pack_bf16(float a, float b) {
  f2 v = {a, b};
  bf16x2 r = __builtin_convertvector(v, bf16x2);
  return __builtin_bit_cast(uint32_t, r);
}
__device__ inline float bf16_lo_to_f32(uint32_t p) {
  return __builtin_bit_cast(float, p << 16);
}
__device__ inline float bf16_hi_to_f32(uint32_t p) {
  return __builtin_bit_cast(float, p & 0xFFFF0000u);
}

// Elements each lane gathers per B row: NB features, as 32-bit words.
template <int FORM, int NB>
struct BRow {
  static constexpr int kWords = FORM == kBF16 ? (NB + 1) / 2 : NB;
  uint32_t w[kWords];
};

template <int FORM, int NB>
__device__ inline void load_brow(BRow<FORM, NB>& r, const void* B, uint32_t ldb,
                                 uint32_t col, uint32_t feat, uint32_t K,
                                 bool full) {
  constexpr int kW = BRow<FORM, NB>::kWords;
  if (col == kPad) {
#pragma unroll
    for (int i = 0; i < kW; ++i) r.w[i] = 0;
    return;
  }
  if (FORM == kBF16) {
    const uint16_t* p = static_cast<const uint16_t*>(B) + (uint64_t)col * ldb + feat;
    if (NB == 8) {
      u4 v = *reinterpret_cast<const u4*>(p);
      r.w[0] = v[0]; r.w[1] = v[1]; r.w[2] = v[2]; r.w[3] = v[3];
    } else if (NB == 4) {
      uint2 v = *reinterpret_cast<const uint2*>(p);
      r.w[0] = v.x; r.w[1] = v.y;
    } else if (NB == 2) {
      r.w[0] = *reinterpret_cast<const uint32_t*>(p);
    } else {
      r.w[0] = *p;
    }
    return;
  }
  const uint32_t* p = static_cast<const uint32_t*>(B) + (uint64_t)col * ldb + feat;
  if (full) {
    if (NB >= 4) {
#pragma unroll
      for (int i = 0; i < NB / 4; ++i) {
        u4 v = *reinterpret_cast<const u4*>(p + 4 * i);
        r.w[4 * i] = v[0]; r.w[4 * i + 1] = v[1];
        r.w[4 * i + 2] = v[2]; r.w[4 * i + 3] = v[3];
      }
    } else if (NB == 2) {
      uint2 v = *reinterpret_cast<const uint2*>(p);
      r.w[0] = v.x; r.w[1] = v.y;
    } else {
      r.w[0] = *p;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NB; ++i) r.w[i] = feat + i < K ? p[i] : 0u;
  }
}

// One pipeline stage: everything a lane needs for one 4-tile group.
template <int FORM, int NB>
struct Stage {
  BRow<FORM, NB> b[8];
  float a[8];
};

template <int FORM, int NB, int RW>
__device__ inline void load_stage(Stage<FORM, NB>& st, const SpmmArgs& g,
                                  uint32_t t, bool tile_ok, uint32_t m,
                                  bool a_rows, uint32_t feat, bool full,
                                  const void* B, uint32_t ldb) {
  uint32_t col[8];
  unsigned long long bw = 0;
  uint32_t off = 0;
  if (tile_ok) {
    const u4* cp = reinterpret_cast<const u4*>(g.a2b + (uint64_t)t * 8);
    u4 c0 = cp[0], c1 = cp[1];
    col[0] = c0[0]; col[1] = c0[1]; col[2] = c0[2]; col[3] = c0[3];
    col[4] = c1[0]; col[5] = c1[1]; col[6] = c1[2]; col[7] = c1[3];
    off = g.tco[t];
    if (RW == 8) {
      bw = g.bits[t];
    } else {
      unsigned long long w0 = g.bits[2ull * t];
      bw = g.bits[2ull * t + 1];
      if (m < 8) bw = w0; else off += popc64(w0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) col[j] = kPad;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    load_brow<FORM, NB>(st.b[j], B, ldb, col[j], feat, g.K, full);
  const uint32_t sh = 8u * (m & 7u);
  const uint32_t byte = a_rows ? (uint32_t)(bw >> sh) & 0xFFu : 0u;
  uint32_t idx = off + (sh ? popc64(bw & ((1ull << sh) - 1ull)) : 0u);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float v = 0.0f;
    if (byte & (1u << j)) {
      v = g.data[idx];
      ++idx;
    }
    st.a[j] = v;
  }
}

template <int FORM, int NB, int RW>
__device__ inline void compute_stage(const Stage<FORM, NB>& st, f4* acc,
                                     uint32_t m) {
  if (FORM == kFP32) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            st.a[j], __builtin_bit_cast(float, st.b[j].w[b]), acc[b], 0, 0, 0);
    return;
  }
  // A operand: 8 bf16 along the tile's condensed columns
  u4 aop, alo;
  {
    uint32_t hi[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) hi[d] = pack_bf16(st.a[2 * d], st.a[2 * d + 1]);
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      float l0 = st.a[2 * d] - bf16_lo_to_f32(hi[d]);
      float l1 = st.a[2 * d + 1] - bf16_hi_to_f32(hi[d]);
      alo[d] = FORM == kSPLIT ? pack_bf16(l0, l1) : 0u;
    }
    if (FORM == kSPLIT && RW == 8 && m >= 8) {
      aop = alo;  // rows 8-15 carry the low halves
    } else {
      aop = u4{hi[0], hi[1], hi[2], hi[3]};
    }
  }
  const bf16x8 A = __builtin_bit_cast(bf16x8, aop);
  const bf16x8 AL = __builtin_bit_cast(bf16x8, alo);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (FORM == kSPLIT) {
      u4 bh, bl;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        uint32_t x = st.b[2 * d].w[b], y = st.b[2 * d + 1].w[b];
        bh[d] = __builtin_amdgcn_perm(y, x, 0x05040100u);
        bl[d] = __builtin_amdgcn_perm(y, x, 0x07060302u);
      }
      const bf16x8 BH = __builtin_bit_cast(bf16x8, bh);
      const bf16x8 BL = __builtin_bit_cast(bf16x8, bl);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, BH, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, BL, acc[b], 0, 0, 0);
      if (RW == 16) {
        // all 16 rows are tile rows: the low halves of A need their own MFMAs
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL, BH, acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL, BL, acc[b], 0, 0, 0);
      }
    } else {
      u4 bv;
      const uint32_t sel = (b & 1) ? 0x07060302u : 0x05040100u;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        uint32_t x = st.b[2 * d].w[b / 2], y = st.b[2 * d + 1].w[b / 2];
        bv[d] = __builtin_amdgcn_perm(y, x, sel);
      }
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          A, __builtin_bit_cast(bf16x8, bv), acc[b], 0, 0, 0);
    }
  }
}

template <int NB>
__device__ inline void store_vec(float* p, const float* v, uint32_t nvalid) {
  if (nvalid >= (uint32_t)NB) {
    if (NB >= 4) {
#pragma unroll
      for (int i = 0; i < NB / 4; ++i)
        *reinterpret_cast<f4*>(p + 4 * i) = f4{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
    } else if (NB == 2) {
      *reinterpret_cast<f2*>(p) = f2{v[0], v[1]};
    } else {
      p[0] = v[0];
    }
  } else {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if ((uint32_t)i < nvalid) p[i] = v[i];
  }
}

template <int FORM, int NB, int RW>
__device__ void run_unit(const SpmmArgs& g, const WorkUnit& u, uint32_t kc,
                         const void* B, uint32_t ldb, float* C, uint32_t ldc,
                         uint32_t K, uint32_t rows) {
  constexpr uint32_t KC = 16u * NB;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t q = lane >> 4, n = lane & 15u;
  const uint32_t k0 = kc * KC;
  const uint32_t feat = k0 + NB * n;
  // B loads are full-width unless this chunk straddles K (FP32 form; the
  // converted BF16-form buffers are zero-padded to a multiple of KC)
  const bool full = FORM != kFP32 || k0 + KC <= K;
  // A rows: lanes m = n carry tile row m (RW=16) or m%8 (RW=8; rows 8-15
  // only in the split form, as the low halves)
  const uint32_t m = n;
  const bool a_rows = RW == 16 || m < 8 || FORM == kSPLIT;

  f4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};

  const uint32_t tb = u.tile_begin, te = u.tile_end;
  const uint32_t ngroups = (te - tb + 3) / 4;
  if (ngroups > 0) {
    // ping-pong register stages: loads of group g+1 are in flight while the
    // MFMAs of group g run
    Stage<FORM, NB> s0, s1;
    load_stage<FORM, NB, RW>(s0, g, tb + q, tb + q < te, m, a_rows, feat, full, B, ldb);
    for (uint32_t gi = 0; gi < ngroups; gi += 2) {
      if (gi + 1 < ngroups) {
        const uint32_t t = tb + 4 * (gi + 1) + q;
        load_stage<FORM, NB, RW>(s1, g, t, t < te, m, a_rows, feat, full, B, ldb);
      }
      compute_stage<FORM, NB, RW>(s0, acc, m);
      if (gi + 1 >= ngroups) break;
      if (gi + 2 < ngroups) {
        const uint32_t t = tb + 4 * (gi + 2) + q;
        load_stage<FORM, NB, RW>(s0, g, t, t < te, m, a_rows, feat, full, B, ldb);
      }
      compute_stage<FORM, NB, RW>(s1, acc, m);
    }
  }

  // combine rows r and r+8 (split form, RW=8): lanes l and l^32
  if (FORM == kSPLIT && RW == 8) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[b][i] += __shfl_xor(acc[b][i], 32, 64);
  }
  // rows held by this lane: 4q + i
  const bool store_lane = RW == 16 || q < 2;
  const uint32_t nvalid = feat < K ? min((uint32_t)NB, K - feat) : 0u;
  if (u.slot < 0) {
    if (store_lane) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        uint32_t r = u.window * RW + 4 * q + i;
        if (r < rows && nvalid) {
          float v[NB];
#pragma unroll
          for (int b = 0; b < NB; ++b) v[b] = acc[b][i];
          store_vec<NB>(C + (uint64_t)r * ldc + feat, v, nvalid);
        }
      }
    }
    return;
  }
  // split piece: write partial (RW rows x kc_padded floats per slot)
  float* part = g.partials + (uint64_t)u.slot * RW * g.kc_padded;
  if (store_lane) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) v[b] = acc[b][i];
      store_vec<NB>(part + (uint64_t)(4 * q + i) * g.kc_padded + NB * n + k0, v, NB);
    }
  }
  const uint32_t* info = g.piece_info + 3ull * (uint32_t)u.slot;
  const uint32_t first = info[0], count = info[1], cid = info[2] * g.n_kc + kc;
  __threadfence();
  uint32_t prev = 0;
  if (lane == 0)
    prev = __hip_atomic_fetch_add(g.counters + cid, 1u, __ATOMIC_ACQ_REL,
                                  __HIP_MEMORY_SCOPE_AGENT);
  prev = __shfl(prev, 0, 64);
  if (prev != count - 1) return;
  __threadfence();
  if (lane == 0)
    __hip_atomic_store(g.counters + cid, 0u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // last piece: sum partials of pieces first..first+count-1 in order
  if (!store_lane) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) v[b] = 0.f;
    for (uint32_t p = 0; p < count; ++p) {
      const float* src = g.partials + (uint64_t)(first + p) * RW * g.kc_padded +
                         (uint64_t)(4 * q + i) * g.kc_padded + NB * n + k0;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        v[b] += __hip_atomic_load(src + b, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    }
    uint32_t r = u.window * RW + 4 * q + i;
    if (r < rows && nvalid) store_vec<NB>(C + (uint64_t)r * ldc + feat, v, nvalid);
  }
}

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
  if (tile && tile->feat_per_lane) {
    t = *tile;
  } else {
    ACC_TRY(accspmm_tile_select(h, K, form, &t));
  }
  // fall back to narrower chunks when the caller's layout is misaligned
  while (t.feat_per_lane > 1 &&
         (((int)form == kFP32 &&
           ((ldb % t.feat_per_lane) || ((uintptr_t)B % (4u * t.feat_per_lane)))) ||
          ((phases & 2) &&
           ((ldc % t.feat_per_lane) || ((uintptr_t)C % (4u * t.feat_per_lane))))))
    t.feat_per_lane /= 2;
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
