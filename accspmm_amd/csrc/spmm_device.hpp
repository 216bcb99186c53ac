// spmm_device.hpp — device code shared by the single-graph SpMM kernel
// (spmm.hip) and the persistent batched kernel (batch.hip): the per-work-unit
// routine run_unit.  Both kernels execute the same arithmetic
// in the same order for a given (graph, K, NB, split threshold), so a problem
// computed alone and inside a batch gives bit-identical C.
#pragma once
#include "internal.hpp"

namespace accspmm {
namespace dev {


typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef __bf16 bf16x2 __attribute__((__vector_size__(2 * sizeof(__bf16))));
typedef float f2 __attribute__((__vector_size__(2 * sizeof(float))));
typedef float f4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef uint32_t u4 __attribute__((__vector_size__(4 * sizeof(uint32_t))));

// Forms.  kFP32 and kSPLITF gather the caller's FP32 B (padded tile columns
// read past the end of a buffer, or row 0 masked when that row is not
// finite); kSPLIT and kBF16 gather a converted copy of B (split-BF16: bf16
// hi | lo << 16 per element; BF16: bf16) with an all-zero last row.  kSPLIT
// and kSPLITF are the split-BF16 form with B split by the conversion pass or
// in registers, bit for bit the same split.  A's FP32 values are split or
// rounded in registers.
enum : int { kFP32 = 0, kSPLIT = 1, kBF16 = 2, kSPLITF = 3 };
constexpr bool is_split(int f) { return f == kSPLIT || f == kSPLITF; }
constexpr bool raw_b(int f) { return f == kFP32 || f == kSPLITF; }

struct SpmmArgs {
  const uint32_t* __restrict__ rwo;
  const uint32_t* __restrict__ tco;
  const unsigned long long* __restrict__ bits;
  const uint32_t* __restrict__ a2b;
  const float* __restrict__ data;
  const WorkUnit* __restrict__ units;
  const uint32_t* __restrict__ split_info;  // per split window: (window, first slot, count)
  float* __restrict__ partials;
  const void* __restrict__ B;  // FP32, SPLITF: float; SPLIT: u32 (hi | lo << 16); BF16: u16
  float* __restrict__ C;
  // C row of format row r: row_map[r] when the graph was built with its
  // rows reordered, r when row_map is null
  const uint32_t* __restrict__ row_map;
  // what padded columns read on the 64-bit gather path: converted B: row
  // zero_row (its last row, all zero); the caller's FP32 B: row 0, masked
  // after the load when that row holds an infinity or NaN (0 x inf would
  // otherwise leak into C; reading a separate zero buffer instead was
  // measured slower on gfx950)
  uint32_t zero_row;
  uint32_t b_rows;    // rows of the gathered B (buffer loads past them read zeros)
  uint32_t pad_tile;  // the all-padding tile (index T) past the real tiles
  uint32_t ldb, ldc;
  uint32_t rows, K;
  uint32_t num_units;
  uint32_t n_kc;  // feature chunks
  uint32_t kc_padded;  // KC * n_kc (partials row stride)
  uint32_t xcd_slice;
  uint32_t pin_len;    // xcd_slice 3: units per XCD (Schedule::pin_len)
  uint32_t rw, nb;     // used by the batched kernel's dispatch
  // 1: 16-row windows with groups of more than 256 values (the graph's
  // max_group_values), which load their values in two spans
  uint32_t wide;
};

__device__ inline uint32_t popc64(unsigned long long x) { return __popcll(x); }

// bf16 round-to-nearest-even pair (a in the low half)
__device__ inline uint32_t pack_bf16(float a, float b) {
  f2 v = {a, b};
  bf16x2 r = __builtin_convertvector(v, bf16x2);
  return __builtin_bit_cast(uint32_t, r);
}

// x - p.lo (lo = 1) or y - p.hi (lo = 0) for a bf16 pair p, in one
// v_dot2_f32_bf16: acc + p.lo * c.lo + p.hi * c.hi with c = (-1, 0) or
// (0, -1).  The products are exact and so is the difference (the split
// below only subtracts a rounding of the operand itself), so the result
// does not depend on how the instruction rounds.  The constant goes through
// a register: the compiler encodes the pair (-1, 0) as the inline constant
// -1.0, which gfx950 reads as the fp32 bits 0xBF800000, i.e. the pair
// (0, -1) (measured: 1,048,454 of 2^20 results off; none through a VGPR).
__device__ inline uint32_t in_vgpr(uint32_t c) {
  asm("" : "+v"(c));
  return c;
}
template <bool LO>
__device__ inline float sub_bf16(uint32_t p, float acc) {
  const uint32_t c = in_vgpr(LO ? 0x0000BF80u : 0xBF800000u);
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, p),
                                         __builtin_bit_cast(bf16x2, c), acc, false);
}

// Two-term bf16 split of two floats x, y: hi = rne(x), lo = rne(x - hi),
// each a bf16 pair (x in the low half).  x - hi is exact, and the remainder
// x - hi - lo is at most 2^-17 |x| (2^e <= |x|: |x - hi| <= 2^(e-8), and
// below that lo keeps 8 significant bits).  Four VALU operations per pair.
__device__ inline void split2(uint32_t x, uint32_t y, uint32_t& hi, uint32_t& lo) {
  const float fx = __builtin_bit_cast(float, x), fy = __builtin_bit_cast(float, y);
  hi = pack_bf16(fx, fy);
  lo = pack_bf16(sub_bf16<true>(hi, fx), sub_bf16<false>(hi, fy));
}

// Three-term bf16 split of two floats x, y (the three-way split of Henry,
// Tang and Heinecke, "Leveraging the bfloat16 Artificial Intelligence
// Datatype For Higher-Precision Computations", ARITH 2019): t0 = rne(x),
// t1 = rne(x - t0), t2 = x - t0 - t1, each a bf16 pair (x in the low half).
// x - t0 carries at most 16 significant bits and x - t0 - t1 at most 8, so
// t2 is exact and x = t0 + t1 + t2 with no error for every finite x whose
// last significant bit lies in bf16's range (|x| >= 2^-110 or so).  Seven
// VALU operations per pair.
__device__ inline void split3(uint32_t x, uint32_t y, uint32_t& t0, uint32_t& t1,
                              uint32_t& t2) {
  const float fx = __builtin_bit_cast(float, x), fy = __builtin_bit_cast(float, y);
  t0 = pack_bf16(fx, fy);
  const float rx = sub_bf16<true>(t0, fx), ry = sub_bf16<false>(t0, fy);
  t1 = pack_bf16(rx, ry);
  const float sx = sub_bf16<true>(t1, rx), sy = sub_bf16<false>(t1, ry);
  t2 = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, sy),
                             __builtin_bit_cast(uint32_t, sx), 0x07060302u);
}

// B conversion, one chunk of four consecutive features: chunk c of the
// (N + 1) x ldx / 4 chunks of the converted B (ldx a multiple of 16, zero
// past K and in the all-zero row N that padded tile columns read).  Row r
// takes B's row perm[r] (a graph built with a symmetric ordering), or row r
// when perm is null.  A 16-byte load when the segment lies inside K and B
// is 16-byte aligned with ldb % 4 == 0 (vec); one 16-byte (split: hi | lo
// << 16 per feature, split2's rounding; FP32: the values, the permuted copy
// the FP32-gathering forms read) or 8-byte (BF16) store.
template <int FORM>
__device__ inline void convert_chunk(const float* __restrict__ B, uint32_t N, uint32_t K,
                                     uint32_t ldb, void* __restrict__ out, uint32_t ldx,
                                     uint32_t vec, uint64_t c,
                                     const uint32_t* __restrict__ perm) {
  const uint32_t cpr = ldx / 4;  // chunks per row
  const uint32_t r = (uint32_t)(c / cpr), k = (uint32_t)(c % cpr) * 4;
  float x[4] = {0.f, 0.f, 0.f, 0.f};
  if (r < N) {
    const float* src = B + (uint64_t)(perm ? perm[r] : r) * ldb + k;
    if (vec && k + 4 <= K) {
      const f4 v = *reinterpret_cast<const f4*>(src);
      x[0] = v[0]; x[1] = v[1]; x[2] = v[2]; x[3] = v[3];
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = k + i < K ? src[i] : 0.f;
    }
  }
  if (FORM == kFP32) {
    reinterpret_cast<f4*>(out)[c] = f4{x[0], x[1], x[2], x[3]};
  } else if (FORM == kSPLIT) {
    u4 o;
#pragma unroll
    for (int i = 0; i < 4; i += 2) {
      uint32_t hi, lo;
      split2(__builtin_bit_cast(uint32_t, x[i]), __builtin_bit_cast(uint32_t, x[i + 1]), hi, lo);
      o[i] = __builtin_amdgcn_perm(lo, hi, 0x05040100u);      // x[i]: hi | lo << 16
      o[i + 1] = __builtin_amdgcn_perm(lo, hi, 0x07060302u);  // x[i+1]
    }
    reinterpret_cast<u4*>(out)[c] = o;
  } else {
    uint2 o;
    o.x = pack_bf16(x[0], x[1]);
    o.y = pack_bf16(x[2], x[3]);
    reinterpret_cast<uint2*>(out)[c] = o;
  }
}

// Elements each lane gathers per B row: NB features, as 32-bit words.
template <int FORM, int NB>
struct BRow {
  static constexpr int kWords = FORM == kBF16 ? (NB + 1) / 2 : NB;
  uint32_t w[kWords];
};

// Features of a lane's B segment.  Lane n of a quarter covers NB features
// of the chunk at k0: k0 + NB n + i for word i, except NB = 8 with 4-byte
// elements (FP32 B, split-BF16): words 0-3 are features k0 + 4n + i and
// words 4-7 features k0 + 64 + 4n + i - 4, so each of the lane's two
// 16-byte loads covers 256 contiguous bytes across the quarter's 16 lanes
// (consecutive 32-byte runs per lane would use every cache line of the row
// half in each load).  The output follows the same map.
template <int FORM, int NB>
constexpr bool split_halves() { return NB == 8 && FORM != kBF16; }
// Chunks of 128-byte rows, 8 bytes a lane (NB = 2 with 4-byte elements, NB
// = 4 with bf16): lane n of a quarter covers the P = 16 / (2 x element
// size) features from k0 + 2P (n & 7) + P (n >> 3), so lanes n and n ^ 8
// hold the 2P features of one 16-byte segment (the buffer-load gather loads
// two rows per instruction and exchanges halves, pair_gather).
template <int FORM, int NB>
constexpr bool pair_rows() {
  return FORM == kBF16 ? NB == 4 : NB == 2;
}
template <int FORM, int NB>
constexpr uint32_t word_feat(int i) {
  return split_halves<FORM, NB>() && i >= 4 ? 60u + (uint32_t)i : (uint32_t)i;
}
template <int FORM, int NB>
__device__ inline uint32_t lane_feat(uint32_t k0, uint32_t n) {
  if constexpr (pair_rows<FORM, NB>()) {
    constexpr uint32_t P = FORM == kBF16 ? 4u : 2u;
    return k0 + 2u * P * (n & 7u) + P * (n >> 3);
  }
  return k0 + (split_halves<FORM, NB>() ? 4u : (uint32_t)NB) * n;
}

// Where a lane's B row segments come from.  BUF: buffer loads through rsrc
// (num_records = b_rows * row_bytes) at 32-bit offsets c * row_bytes +
// feature bytes (one v_mad_u32_u24 of the column's low 24 bits); a padded
// column (kPad, low 24 bits 0xFFFFFF) lands past the end of the buffer and
// the hardware returns zeros for it, so no clamp, mask or zero row is
// needed.  Used when every offset fits 32 bits and that padded product
// does fall past the end (buffer_ok).  Otherwise 64-bit addresses from B with stride ldb elements:
// padded columns read the converted B's zero row (BF16 form) or row 0 of
// the caller's B (masked when consumed if that row is not finite).
struct BSrc {
  __amdgpu_buffer_rsrc_t rsrc;
  const void* B;
  uint32_t ldb;        // elements
  uint32_t row_bytes;  // ldb * element size
  uint32_t b_rows;     // rows of the gathered B
  uint32_t zero_row;   // 64-bit path, BF16 form: the all-zero row
};

// Whether every B offset of the BUF path fits its 32-bit arithmetic (rows
// and row bytes within v_mad_u32_u24's 24-bit operands, every row's segment
// below 2^32) and a padded column's offset, the low 32 bits of 0xFFFFFF *
// row_bytes, lies past the buffer's b_rows * row_bytes bytes with its whole
// row segment below 2^32 (so its loads return zeros without a clamp).
__host__ __device__ inline bool buffer_ok(uint64_t b_rows, uint64_t row_bytes) {
  if (!(b_rows < (1u << 24) && row_bytes < (1u << 24) &&
        (b_rows + 1) * row_bytes <= 0xFFFFFF00ull))
    return false;
  const uint64_t pad = (0xFFFFFFull * row_bytes) & 0xFFFFFFFFull;
  return pad >= b_rows * row_bytes && pad + row_bytes <= 0x100000000ull;
}

template <int FORM, int NB, bool BUF>
__device__ inline BSrc make_bsrc(const void* B, uint32_t ldb, uint32_t b_rows,
                                 uint32_t zero_row) {
  BSrc s;
  constexpr uint32_t es = FORM == kBF16 ? 2u : 4u;
  s.B = B;
  s.ldb = ldb;
  s.row_bytes = ldb * es;
  s.b_rows = b_rows;
  s.zero_row = zero_row;
  if (BUF)
    s.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(B), (short)0,
                                               (int)(b_rows * s.row_bytes), 0x00020000);
  return s;
}

// Branch-free gather of one B row segment.  Every lane issues the same
// vector loads and no result is touched until its stage is consumed, which
// keeps the count of outstanding loads static: the compiler can then wait for
// one pipeline stage without draining the younger ones.
template <int FORM, int NB, bool FULL, bool BUF>
__device__ inline void load_brow(BRow<FORM, NB>& r, const BSrc& src, uint32_t col,
                                 uint32_t feat, uint32_t K) {
  constexpr uint32_t es = FORM == kBF16 ? 2u : 4u;
  if constexpr (BUF) {
    const uint32_t off = __umul24(col, src.row_bytes) + feat * es;
    if constexpr (FORM == kBF16) {
      if constexpr (NB == 8) {
        const u4 v = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, off, 0, 0));
        r.w[0] = v[0]; r.w[1] = v[1]; r.w[2] = v[2]; r.w[3] = v[3];
      } else if constexpr (NB == 4) {
        const uint2 v = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(src.rsrc, off, 0, 0));
        r.w[0] = v.x; r.w[1] = v.y;
      } else if constexpr (NB == 2) {
        r.w[0] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, off, 0, 0);
      } else {
        r.w[0] = __builtin_amdgcn_raw_buffer_load_b16(src.rsrc, off, 0, 0);
      }
    } else if constexpr (FORM == kSPLIT || FULL) {
      if constexpr (NB >= 4) {
        // the second 16 bytes of NB = 8 (64 features on) through the
        // scalar offset: no per-row address add
#pragma unroll
        for (int i = 0; i < NB / 4; ++i) {
          const u4 v = __builtin_bit_cast(
              u4, __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, off,
                                                        4u * word_feat<FORM, NB>(4 * i), 0));
          r.w[4 * i] = v[0]; r.w[4 * i + 1] = v[1];
          r.w[4 * i + 2] = v[2]; r.w[4 * i + 3] = v[3];
        }
      } else if constexpr (NB == 2) {
        const uint2 v = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(src.rsrc, off, 0, 0));
        r.w[0] = v.x; r.w[1] = v.y;
      } else {
        r.w[0] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, off, 0, 0);
      }
    } else {
      // feature tail (FP32 B): features past K read past the end (zeros)
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const uint32_t wf = word_feat<FORM, NB>(i);
        r.w[i] = __builtin_amdgcn_raw_buffer_load_b32(
            src.rsrc, feat + wf < K ? off + 4u * wf : 0xFFFFFFFFu, 0, 0);
      }
    }
    return;
  }
  const void* B = src.B;
  const uint32_t ldb = src.ldb;
  if constexpr (FORM == kBF16) {
    const uint32_t c = col != kPad ? col : src.zero_row;
    const uint16_t* p = static_cast<const uint16_t*>(B) + (uint64_t)c * ldb + feat;
    if constexpr (NB == 8) {
      u4 v = *reinterpret_cast<const u4*>(p);
      r.w[0] = v[0]; r.w[1] = v[1]; r.w[2] = v[2]; r.w[3] = v[3];
    } else if constexpr (NB == 4) {
      uint2 v = *reinterpret_cast<const uint2*>(p);
      r.w[0] = v.x; r.w[1] = v.y;
    } else if constexpr (NB == 2) {
      r.w[0] = *reinterpret_cast<const uint32_t*>(p);
    } else {
      r.w[0] = *p;
    }
  } else if constexpr (FORM == kSPLIT || FULL) {
    // the converted split B: padded columns read its zero row; the caller's
    // FP32 B: they read row 0 (masked when consumed if that row is not
    // finite)
    const uint32_t row = FORM == kSPLIT ? (col != kPad ? col : src.zero_row)
                                        : (col != kPad ? col : 0u);
    const uint32_t* p = static_cast<const uint32_t*>(B) + (uint64_t)row * ldb + feat;
    if constexpr (NB >= 4) {
#pragma unroll
      for (int i = 0; i < NB / 4; ++i) {
        u4 v = *reinterpret_cast<const u4*>(p + word_feat<FORM, NB>(4 * i));
        r.w[4 * i] = v[0]; r.w[4 * i + 1] = v[1];
        r.w[4 * i + 2] = v[2]; r.w[4 * i + 3] = v[3];
      }
    } else if constexpr (NB == 2) {
      uint2 v = *reinterpret_cast<const uint2*>(p);
      r.w[0] = v.x; r.w[1] = v.y;
    } else {
      r.w[0] = *p;
    }
  } else {
    // feature tail (FP32 B): clamp inside the row, zero past K
    const uint32_t* row = static_cast<const uint32_t*>(B) + (uint64_t)(col != kPad ? col : 0u) * ldb;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const uint32_t f = feat + word_feat<FORM, NB>(i);
      const uint32_t x = row[min(f, K - 1)];
      r.w[i] = f < K ? x : 0u;
    }
  }
}

// A 4-tile group's A values are consecutive in the METCF values array (its
// tiles are consecutive): at most 4 x 64 of them with 8-row windows, 4 x 128
// with 16-row windows.  One 16-byte load per lane brings a 256-float span
// starting at the group's first value, which goes through a per-wave LDS
// buffer (group_values floats, then one zero read by empty bitmap slots).
// One vector-memory instruction instead of one per tile column: the
// texture-address unit, which handles every lane address of every load at
// four lanes a clock, bounds this kernel.  A 16-row group of more than 256
// values (tiles over half full) loads the rest when it is consumed.
constexpr uint32_t kSpanValues = 256;
template <int RW>
constexpr uint32_t group_values() { return RW == 16 ? 512u : 256u; }
template <int RW>
constexpr uint32_t wave_lds() { return group_values<RW>() + 4u; }  // 16-byte multiple

// One pipeline stage: everything a lane needs for one 4-tile group.  Padded
// columns of the BF16 forms read the converted B's zero row; bit j of bmask
// marks a real column j (the FP32 form masks the others).
template <int FORM, int NB>
struct Stage {
  BRow<FORM, NB> b[8];
  u4 vals;        // this lane's 4 floats of the group's value span
  uint32_t byte;  // the lane's tile row of the bitmap (0: no A row)
  uint32_t base;  // index of the row's first value within the span
  uint32_t bmask;
  uint32_t win;   // tag of the group (wave-uniform)
  uint32_t vbase; // RW=16: index of the group's first value (wave-uniform)
};

// Tile metadata as loaded: one 16-byte load per lane, the quarter's tile
// spread over each group of four lanes (r = lane & 3): r = 0 its columns
// 0-3, r = 1 columns 4-7 (sparseA2B), r = 2 its bitmap words (RW=8: the
// tile's word and the next tile's; RW=16: both words), r = 3
// tcOffset[t..t+3].  One vector-memory instruction per group for all of
// it: the texture-address unit, which bounds this kernel, costs the same per
// instruction whatever each lane reads.
struct Meta {
  u4 m;
  uint32_t win;   // tag of the group (wave-uniform)
};

// Position of a 4-tile group (wave-uniform): its window w, first tile t and
// the end tile of its range.  Groups never span two windows.
struct GroupPos {
  uint32_t w, t, tend;
};

// Advances to the next group of the unit; only called while groups remain.
// A unit of one window (or one piece) never leaves its range; packed units
// contain no empty windows, so the next window has tiles.
template <bool PACKED>
__device__ inline void next_group(GroupPos& p, const uint32_t* __restrict__ rwo) {
  p.t += 4;
  if (PACKED && p.t >= p.tend) {
    ++p.w;
    p.t = rwo[p.w];
    p.tend = rwo[p.w + 1];
  }
}

typedef uint32_t u4a __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u2a __attribute__((ext_vector_type(2), aligned(4)));

// Where lane r = lane & 3 of a quad loads its piece of a tile's metadata:
// tile tt's bytes base + tt * stride.  The builder pads tcOffset with
// three spare entries and the bitmaps with a spare tile so every such
// 16-byte read stays inside its array.
struct MetaSrc {
  const char* base;
  uint32_t stride;
};

template <int RW>
__device__ inline MetaSrc meta_src(const SpmmArgs& g, uint32_t lane) {
  const uint32_t r = lane & 3u;
  MetaSrc s;
  if (r < 2) {
    s.base = reinterpret_cast<const char*>(g.a2b) + 16u * r;
    s.stride = 32;
  } else if (r == 2) {
    s.base = reinterpret_cast<const char*>(g.bits);
    s.stride = RW == 16 ? 16 : 8;
  } else {
    s.base = reinterpret_cast<const char*>(g.tco);
    s.stride = 4;
  }
  return s;
}

template <int RW>
__device__ inline void load_meta(Meta& mt, const SpmmArgs& g, const GroupPos& p,
                                 bool valid, uint32_t q, const MetaSrc& src) {
  // a tile past the group's range (or an invalid group) reads the
  // all-padding tile: kPad columns, empty bitmap, offset nnz
  const uint32_t t = p.t + q;
  const bool tile_ok = valid && t < p.tend;
  const uint32_t tt = tile_ok ? t : g.pad_tile;
  const u4a v = *reinterpret_cast<const u4a*>(src.base + (uint64_t)tt * src.stride);
  mt.m = u4{v.x, v.y, v.z, v.w};
  mt.win = p.w;
}

// The value lane k (0-3) of this lane's quad holds (DPP quad broadcast).
template <int K>
__device__ inline uint32_t from_quad(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, K * 0x55, 0xF, 0xF, false);
}

// Gathers the B rows of the tile's 8 columns and the lane's share of the
// group's value span; records the lane's A row (bitmap byte, first value).
// Two-row gathers (pair_rows chunks, buffer loads, full chunks): lane n of
// a quarter loads the 16-byte segment holding its features of row col[j]
// (n < 8) or col[j + 1] (n >= 8), j = 0, 2, 4, 6: four loads for the tile's
// eight rows instead of eight 8-byte ones (232,965-node power-law graph:
// split-BF16 at K = 32 1.107 -> 1.044 ms, BF16 at K = 64 1.137 -> 1.052 ms,
// FP32 at K = 32 unchanged at 1.12 ms).  The stage keeps the raw words (row j's slot: the first two, row j + 1's: the last
// two) until exchange_pairs sorts them when it is consumed.
template <int FORM, int NB, bool FULL, bool BUF>
constexpr bool pair_gather() {
  return pair_rows<FORM, NB>() && BUF && (FORM == kSPLIT || FULL);
}

// Sorts a two-row gather's words: each lane keeps the two features it
// covers of the row it loaded and takes its features of the other row from
// lane n ^ 8 (a DPP rotate by 8 within each 16-lane row; lanes 0-7 of a
// row are its banks 0 and 1, written by bank mask 0x3).
template <int FORM, int NB>
__device__ inline void exchange_pairs(Stage<FORM, NB>& st) {
#pragma unroll
  for (int j = 0; j < 8; j += 2)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int x = (int)st.b[j].w[i], y = (int)st.b[j + 1].w[i];
      // lanes 8-15: row j from lane n - 8's last two words; lanes 0-7: row
      // j + 1 from lane n + 8's first two
      st.b[j].w[i] = (uint32_t)__builtin_amdgcn_update_dpp(x, y, 0x128, 0xF, 0xC, false);
      st.b[j + 1].w[i] = (uint32_t)__builtin_amdgcn_update_dpp(y, x, 0x128, 0xF, 0x3, false);
    }
}

template <int FORM, int NB, bool FULL, int RW, bool BUF>
__device__ inline void load_data(Stage<FORM, NB>& st, const Meta& mt,
                                 const SpmmArgs& g, uint32_t n, bool a_rows,
                                 uint32_t feat, const BSrc& src, uint32_t lane) {
  // columns 0-3 from quad lane 0, 4-7 from lane 1, the bitmap from lane 2,
  // tcOffset[t] from lane 3
  uint32_t col[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    col[j] = from_quad<0>(mt.m[j]);
    col[4 + j] = from_quad<1>(mt.m[j]);
  }
  unsigned long long bw;
  uint32_t off;
  if (RW == 8) {
    bw = (unsigned long long)from_quad<2>(mt.m[1]) << 32 | from_quad<2>(mt.m[0]);
    off = from_quad<3>(mt.m[0]);
  } else {
    // rows 0-7 (lanes n < 8) take word 0 and start at tcOffset[t]; rows
    // 8-15 word 1, after word 0's values.  Every lane of a quad has the
    // same n < 8, so the bitmap lane picks the word before the broadcast.
    const bool low = n < 8;
    const uint32_t lo = low ? mt.m[0] : mt.m[2], hi = low ? mt.m[1] : mt.m[3];
    const uint32_t skip = low ? 0u : (uint32_t)__popc(mt.m[0]) + (uint32_t)__popc(mt.m[1]);
    bw = (unsigned long long)from_quad<2>(hi) << 32 | from_quad<2>(lo);
    off = from_quad<3>(mt.m[0]) + from_quad<2>(skip);
  }
  if constexpr (pair_gather<FORM, NB, FULL, BUF>()) {
    constexpr uint32_t es = FORM == kBF16 ? 2u : 4u, seg_feats = 16u / es;
    const uint32_t seg = (feat & ~(seg_feats - 1u)) * es;
    // all ones in lanes 8-15 of each quarter: a mask select, which the
    // compiler would otherwise turn into a lane-indexed read of col[]
    // (scratch memory)
    const uint32_t upper = 0u - ((lane >> 3) & 1u);
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      const uint32_t c = col[j] ^ ((col[j] ^ col[j + 1]) & upper);
      const u4 v = __builtin_bit_cast(
          u4, __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, __umul24(c, src.row_bytes) + seg,
                                                    0, 0));
      st.b[j].w[0] = v[0]; st.b[j].w[1] = v[1];
      st.b[j + 1].w[0] = v[2]; st.b[j + 1].w[1] = v[3];
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) load_brow<FORM, NB, FULL, BUF>(st.b[j], src, col[j], feat, g.K);
  }
  uint32_t bm = 0;
  if (raw_b(FORM)) {
#pragma unroll
    for (int j = 0; j < 8; ++j) bm |= (col[j] != kPad ? 1u : 0u) << j;
  }
  st.bmask = bm;
  // the group's first value: tcOffset of quarter 0's tile (lane 0)
  const uint32_t vbase = __builtin_amdgcn_readfirstlane(off);
  const uint32_t* vsrc = reinterpret_cast<const uint32_t*>(g.data);
  // RW=8: lanes past the group's values (it ends at tcOffset[t3 + 1],
  // which lane 51 of quarter 3, its tcOffset lane, holds; a padding tile
  // there gives nnz, clamped to the span) re-read the span's first 16
  // bytes, so only the cache lines holding values are requested.  Every
  // lane still issues the load: the wait counts of the pipeline stay
  // static.
  uint32_t nv = kSpanValues;
  if (RW == 8) nv = min(__builtin_amdgcn_readlane(mt.m[1], 51) - vbase, kSpanValues);
  const uint32_t vi = 4u * lane < nv ? 4u * lane : 0u;
  const u4a v = *reinterpret_cast<const u4a*>(vsrc + vbase + vi);
  st.vals = u4{v.x, v.y, v.z, v.w};
  const uint32_t sh = 8u * (n & 7u);
  st.byte = a_rows ? (uint32_t)(bw >> sh) & 0xFFu : 0u;
  st.base = off - vbase + (sh ? popc64(bw & ((1ull << sh) - 1ull)) : 0u);
  st.win = mt.win;
  if (RW == 16) st.vbase = vbase;
}

// MFMAs of one 4-tile group.  MASK: the 64-bit gather path of the forms
// that gather the caller's FP32 B zeroes what padded columns loaded from
// B's row 0, needed only when that row holds an infinity or NaN (the A value
// there is zero, and 0 x inf is NaN).
template <int FORM, int NB, int RW, bool MASK, bool PAIR = false>
__device__ inline void compute_stage(Stage<FORM, NB>& st, f4* acc,
                                     uint32_t m, float* lv, uint32_t lane,
                                     const uint32_t* __restrict__ vsrc, bool wide) {
  if constexpr (PAIR) exchange_pairs(st);
  // the group's value span through the wave's LDS buffer: a set bitmap bit
  // j reads the row's next value, a clear one the zero after the values
  *reinterpret_cast<u4*>(lv + 4u * lane) = st.vals;
  if (RW == 16 && wide) {
    // the graph has 16-row groups of more than 256 values: count this
    // group's (every lane holds one tile row's bitmap byte) and load the
    // rest of a full one here.  The load waits for every load in flight;
    // graphs whose groups all fit the span (wide = false) never take it.
    uint32_t c = __popc(st.byte);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) c += __shfl_xor(c, o, 64);
    const uint32_t nv = __builtin_amdgcn_readfirstlane(c);
    if (nv > kSpanValues) {
      const uint32_t i = kSpanValues + 4u * lane;
      const u4a v = *reinterpret_cast<const u4a*>(vsrc + st.vbase + (i < nv ? i : 0u));
      *reinterpret_cast<u4*>(lv + kSpanValues + 4u * lane) = u4{v.x, v.y, v.z, v.w};
    }
  }
  // the row's values are consecutive from st.base: column j takes the next
  // one when its bitmap bit is set (then steps on) and zero otherwise.  The
  // sign-extended bit masks the value and steps the read position; a read
  // past the row's last value stays inside the buffer (at most the zero
  // after the group's values) and is masked.
  uint32_t a[8];
  const char* vp = reinterpret_cast<const char*>(lv + st.base);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int32_t m = -(int32_t)((st.byte >> j) & 1u);
    a[j] = *reinterpret_cast<const uint32_t*>(vp) & (uint32_t)m;
    vp += __mul24(m, -4);
  }
  if (raw_b(FORM) && MASK) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const bool bv = (st.bmask >> j) & 1u;
#pragma unroll
      for (int i = 0; i < NB; ++i) st.b[j].w[i] = bv ? st.b[j].w[i] : 0u;
    }
  }
  if (FORM == kFP32) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(
            __builtin_bit_cast(float, a[j]), __builtin_bit_cast(float, st.b[j].w[b]),
            acc[b], 0, 0, 0);
    return;
  }
  if (FORM == kBF16) {
    // A and B rounded to bf16 (B by the conversion pass): one MFMA a block
    u4 aop;
#pragma unroll
    for (int d = 0; d < 4; ++d)
      aop[d] = pack_bf16(__builtin_bit_cast(float, a[2 * d]),
                         __builtin_bit_cast(float, a[2 * d + 1]));
    const bf16x8 A = __builtin_bit_cast(bf16x8, aop);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
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
    return;
  }
  // split-BF16: A = A0 + A1 + A2 exactly (split3) and B = B0 + B1 + r with
  // |r| <= 2^-17 |b| (split2, by the conversion pass or in registers); the
  // five products A_i B_j with i + j <= 2 (j <= 1) are accumulated, so each
  // product a b is off by |a r| + |A2 B1| <= (2^-17 + 2^-25) |a b| before
  // FP32 accumulation.  The A operand's rows are tile rows: RW=16 all 16
  // (five MFMAs a block); RW=8 rows 0-7, so rows 8-15 carry a second A term:
  // [A0; A1] against B0 and B1 gives A0 B0, A1 B0, A0 B1, A1 B1, and
  // [A2; 0] against B0 the last product (three MFMAs a block); rows r and
  // r+8 are added when the window is stored.  Smallest products first.
  uint32_t t0[4], t1[4], t2[4];
#pragma unroll
  for (int d = 0; d < 4; ++d) split3(a[2 * d], a[2 * d + 1], t0[d], t1[d], t2[d]);
  u4 x0 = {t0[0], t0[1], t0[2], t0[3]}, x1 = {t1[0], t1[1], t1[2], t1[3]};
  const u4 x2 = {t2[0], t2[1], t2[2], t2[3]};
  if (RW == 8) {
    const bool upper = m >= 8;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const uint32_t a1 = x1[d];
      x1[d] = upper ? 0u : x2[d];
      x0[d] = upper ? a1 : x0[d];
    }
  }
  const bf16x8 X0 = __builtin_bit_cast(bf16x8, x0);  // RW=8: [A0; A1]; RW=16: A0
  const bf16x8 X1 = __builtin_bit_cast(bf16x8, x1);  // RW=8: [A2; 0];  RW=16: A1
  const bf16x8 X2 = __builtin_bit_cast(bf16x8, x2);  // RW=16: A2
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    u4 bh, bl;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const uint32_t x = st.b[2 * d].w[b], y = st.b[2 * d + 1].w[b];
      if (FORM == kSPLITF) {
        uint32_t h, l;
        split2(x, y, h, l);
        bh[d] = h;
        bl[d] = l;
      } else {
        // converted words hold hi | lo << 16 per feature
        bh[d] = __builtin_amdgcn_perm(y, x, 0x05040100u);
        bl[d] = __builtin_amdgcn_perm(y, x, 0x07060302u);
      }
    }
    const bf16x8 BH = __builtin_bit_cast(bf16x8, bh);
    const bf16x8 BL = __builtin_bit_cast(bf16x8, bl);
    if (RW == 8) {
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X1, BH, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X0, BL, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X0, BH, acc[b], 0, 0, 0);
    } else {
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X2, BH, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X1, BL, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X0, BL, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X1, BH, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X0, BH, acc[b], 0, 0, 0);
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

__device__ inline uint64_t out_row(const uint32_t* __restrict__ rmap, uint32_t r) {
  return rmap ? rmap[r] : r;
}

// Stores a lane's NB values of one output row (row points at its feature 0)
// at features feat + word_feat(i), those below lim.
template <int FORM, int NB>
__device__ inline void store_lane(float* row, uint32_t feat, uint32_t lim, const float* v) {
  if constexpr (split_halves<FORM, NB>()) {
    store_vec<4>(row + feat, v, feat < lim ? min(4u, lim - feat) : 0u);
    store_vec<4>(row + feat + 64, v + 4, feat + 64 < lim ? min(4u, lim - feat - 64) : 0u);
  } else {
    store_vec<NB>(row + feat, v, feat < lim ? min((uint32_t)NB, lim - feat) : 0u);
  }
}

// Writes the accumulated rows of window w to C (split form, RW=8: rows r
// and r+8 of the MFMA tile hold products of different A terms of tile row
// r and are added in place; the caller resets the accumulators afterwards).
template <int FORM, int NB, int RW>
__device__ inline void store_window(f4* acc, uint32_t w, uint32_t q,
                                    uint32_t feat, uint32_t K, float* C,
                                    uint32_t ldc, uint32_t rows,
                                    const uint32_t* rmap) {
  if (is_split(FORM) && RW == 8) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[b][i] += __shfl_xor(acc[b][i], 32, 64);
  }
  if (RW == 8 && q >= 2) return;
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

template <int FORM, int NB, int RW>
__device__ inline void zero_window(uint32_t w, uint32_t q, uint32_t feat,
                                   uint32_t K, float* C, uint32_t ldc,
                                   uint32_t rows, const uint32_t* rmap) {
  if (RW == 8 && q >= 2) return;
  float v[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) v[b] = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t r = w * RW + 4 * q + i;
    if (r < rows) store_lane<FORM, NB>(C + out_row(rmap, r) * ldc, feat, K, v);
  }
}

// Per-unit output state: the window being accumulated and its accumulators.
// A group of a later window first flushes the current one (and writes zeros
// for any empty windows in between).
template <int FORM, int NB, int RW>
struct Accum {
  f4 acc[NB];
  uint32_t cur_w;
  bool direct;  // packed unit: flush to C; split piece: keep for the partial
  uint32_t q, feat, K, ldc, rows;
  float* C;
  const uint32_t* rmap;
  __device__ inline void reset() {
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
  }
  // flush the current window and start window w (packed units hold no
  // empty windows except as a whole unit, so no zero-fill is needed here)
  __device__ inline void move_to(uint32_t w) {
    if (w == cur_w) return;
    store_window<FORM, NB, RW>(acc, cur_w, q, feat, K, C, ldc, rows, rmap);
    reset();
    cur_w = w;
  }
};

// Three-stage software pipeline over a unit's 4-tile groups: the metadata
// of group g+2 and the gathers of group g+1 are in flight while the MFMAs of
// group g run.  The steady-state loop issues the same loads on every path
// (order m(g+2), d(g+1), compute(g)), so each wait on the in-order
// vector-memory counter drains only the stage it needs; the last one to
// three groups run as straight-line tail code.  sched_barrier keeps each
// stage's loads issued before the previous stage's consumer code (the
// scheduler would otherwise hoist the cheap masking and force early waits).
template <int FORM, int NB, int RW, bool FULL, bool PACKED, bool BUF, bool MASK>
__device__ inline void unit_loop(const SpmmArgs& g, Accum<FORM, NB, RW>& out,
                                 GroupPos pos, uint32_t ngroups, uint32_t q,
                                 uint32_t m, bool a_rows, uint32_t feat,
                                 const BSrc& src, float* lv, uint32_t lane) {
  Meta mA, mB;
  Stage<FORM, NB> dA, dB;
  const uint32_t* vsrc = reinterpret_cast<const uint32_t*>(g.data);
  const MetaSrc ms = meta_src<RW>(g, lane);
  auto consume = [&](Stage<FORM, NB>& d) {
    if (PACKED && out.direct) out.move_to(d.win);
    compute_stage<FORM, NB, RW, MASK, pair_gather<FORM, NB, FULL, BUF>()>(d, out.acc, m, lv,
                                                                          lane, vsrc,
                                                                          g.wide != 0u);
  };
  // pos: position of the most recently loaded metadata
  load_meta<RW>(mA, g, pos, true, q, ms);
  if (ngroups > 1) next_group<PACKED>(pos, g.rwo);
  load_meta<RW>(mB, g, pos, ngroups > 1, q, ms);
  load_data<FORM, NB, FULL, RW, BUF>(dA, mA, g, m, a_rows, feat, src, lane);
  uint32_t gi = 0;
  for (; gi + 3 < ngroups; gi += 2) {
    next_group<PACKED>(pos, g.rwo);  // group gi + 2
    load_meta<RW>(mA, g, pos, true, q, ms);
    load_data<FORM, NB, FULL, RW, BUF>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    __builtin_amdgcn_sched_barrier(0);
    const bool v3 = gi + 3 < ngroups;
    if (v3) next_group<PACKED>(pos, g.rwo);  // group gi + 3
    load_meta<RW>(mB, g, pos, v3, q, ms);
    load_data<FORM, NB, FULL, RW, BUF>(dA, mA, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dB);
    __builtin_amdgcn_sched_barrier(0);
  }
  // dA holds group gi, mB the metadata of group gi+1
  const uint32_t left = ngroups - gi;  // 1, 2 or 3
  if (left == 3) {
    next_group<PACKED>(pos, g.rwo);  // group gi + 2
    load_meta<RW>(mA, g, pos, true, q, ms);
    load_data<FORM, NB, FULL, RW, BUF>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    __builtin_amdgcn_sched_barrier(0);
    load_data<FORM, NB, FULL, RW, BUF>(dA, mA, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dB);
    consume(dA);
  } else if (left == 2) {
    load_data<FORM, NB, FULL, RW, BUF>(dB, mB, g, m, a_rows, feat, src, lane);
    __builtin_amdgcn_sched_barrier(0);
    consume(dA);
    consume(dB);
  } else {
    consume(dA);
  }
}

// Wave-uniform: whether row 0 of the caller's FP32 B, which padded tile
// columns read, holds an infinity or NaN in this lane group's features (the
// loaded words must then be masked).  Only called for a unit with tiles,
// so B has at least one row.
template <int FORM, int NB>
__device__ inline bool row0_needs_mask(const void* B, uint32_t feat, uint32_t K) {
  if (!raw_b(FORM)) return false;
  const float* r0 = static_cast<const float*>(B);
  bool bad = false;
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (feat + word_feat<FORM, NB>(i) < K)
      bad |= !__builtin_isfinite(r0[feat + word_feat<FORM, NB>(i)]);
  return __builtin_amdgcn_ballot_w64(bad) != 0;
}

// Writes a split piece's sums to its partial slot (RW rows x kc_padded
// floats; the split form with RW=8 first adds rows r and r+8).
template <int FORM, int NB, int RW>
__device__ inline void store_partial(f4* acc, uint32_t slot, uint32_t q, uint32_t feat,
                                     const SpmmArgs& g) {
  if (is_split(FORM) && RW == 8) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[b][i] += __shfl_xor(acc[b][i], 32, 64);
  }
  const bool holds_rows = RW == 16 || q < 2;
  float* part = g.partials + (uint64_t)slot * RW * g.kc_padded;
  if (holds_rows) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) v[b] = acc[b][i];
      store_lane<FORM, NB>(part + (uint64_t)(4 * q + i) * g.kc_padded, feat, g.kc_padded, v);
    }
  }
}

// PACKED: the schedule may hold multi-window units (the loop then flushes C
// at window changes); without it every direct unit is one window and the
// loop carries no flush code (fewer registers).  BUF: B is gathered by
// buffer loads (g.b_rows rows; buffer_ok holds).
template <int FORM, int NB, int RW, bool PACKED, bool BUF>
__device__ void run_unit(const SpmmArgs& g, const WorkUnit& u, uint32_t kc,
                         const void* B, uint32_t ldb, float* C, uint32_t ldc,
                         uint32_t K, uint32_t rows, float* lv) {
  constexpr uint32_t KC = 16u * NB;
  if (u.nwin == 0) return;  // padding of a pinned schedule's unit lists
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t q = lane >> 4, n = lane & 15u;
  const uint32_t k0 = kc * KC;
  const uint32_t feat = lane_feat<FORM, NB>(k0, n);
  // B loads are full-width unless this chunk straddles K (FP32 B; the
  // converted BF16-form buffer is zero-padded to a multiple of KC)
  const bool full = !raw_b(FORM) || k0 + KC <= K;
  // A rows: lanes m = n carry tile row m (RW=16) or m%8 (RW=8; rows 8-15
  // only in the split form, as a second A term)
  const uint32_t m = n;
  const bool a_rows = RW == 16 || m < 8 || is_split(FORM);

  Accum<FORM, NB, RW> out;
  out.reset();
  out.cur_w = u.window;
  out.direct = u.slot < 0;
  out.q = q;
  out.feat = feat;
  out.K = K;
  out.C = C;
  out.ldc = ldc;
  out.rows = rows;
  out.rmap = g.row_map;

  const uint32_t ngroups = u.ngroups;
  if (ngroups > 0) {
    // the unit is wave-uniform: keep its position in scalar registers
    GroupPos pos{(uint32_t)__builtin_amdgcn_readfirstlane(u.window),
                 (uint32_t)__builtin_amdgcn_readfirstlane(u.tile_begin),
                 (uint32_t)__builtin_amdgcn_readfirstlane(u.tile_end)};
    // packed unit: group ranges are whole windows (its first window is
    // not empty when it has groups)
    if (u.slot < 0) pos.tend = g.rwo[pos.w + 1];
    const BSrc src = make_bsrc<FORM, NB, BUF>(B, ldb, g.b_rows, g.zero_row);
    if (BUF) {
      // padded columns read zeros past the end of the buffer: no mask
      if (full)
        unit_loop<FORM, NB, RW, true, PACKED, BUF, false>(g, out, pos, ngroups, q, m, a_rows,
                                                          feat, src, lv, lane);
      else
        unit_loop<FORM, NB, RW, false, PACKED, BUF, false>(g, out, pos, ngroups, q, m, a_rows,
                                                           feat, src, lv, lane);
    } else if (row0_needs_mask<FORM, NB>(B, feat, K)) {
      if (full)
        unit_loop<FORM, NB, RW, true, PACKED, BUF, true>(g, out, pos, ngroups, q, m, a_rows,
                                                         feat, src, lv, lane);
      else
        unit_loop<FORM, NB, RW, false, PACKED, BUF, true>(g, out, pos, ngroups, q, m, a_rows,
                                                          feat, src, lv, lane);
    } else {
      if (full)
        unit_loop<FORM, NB, RW, true, PACKED, BUF, false>(g, out, pos, ngroups, q, m, a_rows,
                                                          feat, src, lv, lane);
      else
        unit_loop<FORM, NB, RW, false, PACKED, BUF, false>(g, out, pos, ngroups, q, m, a_rows,
                                                           feat, src, lv, lane);
    }
  }

  if (u.slot < 0) {
    // last window of the unit; a unit of empty windows writes zero rows
    store_window<FORM, NB, RW>(out.acc, out.cur_w, q, feat, K, C, ldc, rows, g.row_map);
    for (uint32_t e = out.cur_w + 1; e < u.window + u.nwin; ++e)
      zero_window<FORM, NB, RW>(e, q, feat, K, C, ldc, rows, g.row_map);
    return;
  }
  // split piece: write partial (RW rows x kc_padded floats per slot); the
  // partials are combined by reduce_split_window in a second launch
  store_partial<FORM, NB, RW>(out.acc, (uint32_t)u.slot, q, feat, g);
}

// Sums the piece partials of split window `sw` in slot order and writes the
// window's C rows; the block's threads stride over its RW x K outputs.
// Runs in a launch after the SpMM kernel: the kernel boundary orders the
// partial writes before these reads, so the SpMM kernel needs no
// device-scope fence (on gfx950 one writes back and invalidates the XCD's
// whole L2).
template <int RW>
__device__ inline void reduce_split_window(const uint32_t* __restrict__ sinfo,
                                           uint32_t sw,
                                           const float* __restrict__ partials,
                                           uint32_t kcp, uint32_t K,
                                           float* __restrict__ C, uint32_t ldc,
                                           uint32_t rows,
                                           const uint32_t* __restrict__ rmap) {
  const uint32_t w = sinfo[3ull * sw], first = sinfo[3ull * sw + 1],
                 count = sinfo[3ull * sw + 2];
  // four features per thread when C rows allow 16-byte stores (the
  // partials are: kcp is a multiple of 16); the same per-element order
  if ((K & 3u) == 0 && (ldc & 3u) == 0 && (reinterpret_cast<uintptr_t>(C) & 15u) == 0) {
    const uint32_t K4 = K >> 2;
    for (uint32_t i = threadIdx.x; i < RW * K4; i += blockDim.x) {
      const uint32_t r = i / K4, k = (i - r * K4) * 4u;
      const uint64_t row = (uint64_t)w * RW + r;
      if (row >= rows) continue;
      const float* src = partials + ((uint64_t)first * RW + r) * kcp + k;
      f4 v = f4{0.f, 0.f, 0.f, 0.f};
      for (uint32_t p = 0; p < count; ++p)
        v += *reinterpret_cast<const f4*>(src + (uint64_t)p * RW * kcp);
      *reinterpret_cast<f4*>(C + out_row(rmap, (uint32_t)row) * ldc + k) = v;
    }
    return;
  }
  for (uint32_t i = threadIdx.x; i < RW * K; i += blockDim.x) {
    const uint32_t r = i / K, k = i - r * K;
    const uint64_t row = (uint64_t)w * RW + r;
    if (row >= rows) continue;
    const float* src = partials + ((uint64_t)first * RW + r) * kcp + k;
    float v = 0.f;
    for (uint32_t p = 0; p < count; ++p) v += src[(uint64_t)p * RW * kcp];
    C[out_row(rmap, (uint32_t)row) * ldc + k] = v;
  }
}

}  // namespace dev
}  // namespace accspmm
