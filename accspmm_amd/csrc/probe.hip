// probe.hip — the gather ceiling: how fast MI355X serves the SpMM kernels'
// B access pattern with nothing else in the way.  Each wave reads rows
// idx[...] of a region of B-like rows (row_bytes each, a multiple of 256)
// the way the kernels gather them: quarter-wave q takes one row of a group
// of four, lane n of the quarter 16 bytes at n * 16 + 256 * h for every
// 256-byte half h of the row, through buffer loads.  The rows' words are
// folded into one word per lane so no load is dead.  A measurement tool for
// the roofline report; no reference counterpart.
#include "internal.hpp"

namespace {

typedef uint32_t u4 __attribute__((ext_vector_type(4)));

// rows_per_wave is a multiple of 64; waves * rows_per_wave indices; every
// index < the region's row count (checked on the host)
template <int HALVES>
__global__ __launch_bounds__(256) void k_gather_probe(const void* __restrict__ region,
                                                      uint32_t region_bytes,
                                                      const uint32_t* __restrict__ idx,
                                                      uint32_t rows_per_wave,
                                                      uint32_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t* my = idx + wave * rows_per_wave;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void*>(region), (short)0, (int)region_bytes, 0x00020000);
  const uint32_t q = lane >> 4, n = lane & 15u;
  uint32_t acc = 0;
  for (uint32_t base = 0; base < rows_per_wave; base += 64) {
    const uint32_t row_of_lane = my[base + lane];
#pragma unroll
    for (uint32_t g0 = 0; g0 < 16; g0 += 8) {
      u4 v[8][HALVES];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t row = __shfl(row_of_lane, 4 * (g0 + u) + q, 64);
        const uint32_t off = row * (256u * HALVES) + 16u * n;
#pragma unroll
        for (int h = 0; h < HALVES; ++h)
          v[u][h] = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 256u * h, 0));
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
        for (int h = 0; h < HALVES; ++h) acc ^= v[u][h].x ^ v[u][h].y ^ v[u][h].z ^ v[u][h].w;
    }
  }
  out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

extern "C" accspmm_status accspmm_gather_probe(const void* region, uint32_t region_rows,
                                               uint32_t row_bytes, const uint32_t* idx,
                                               uint32_t waves, uint32_t rows_per_wave,
                                               uint32_t* out, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!region || !idx || !out || region_rows == 0 || waves == 0 || waves % 4 != 0 ||
        rows_per_wave == 0 || rows_per_wave % 64 != 0)
      return ACCSPMM_ERR_INVALID_ARG;
    if (row_bytes != 256 && row_bytes != 512) return ACCSPMM_ERR_UNSUPPORTED;
    if ((uint64_t)region_rows * row_bytes > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
    const uint32_t bytes = region_rows * row_bytes;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(waves / 4), blk(256);
    if (row_bytes == 512)
      hipLaunchKernelGGL(k_gather_probe<2>, grid, blk, 0, s, region, bytes, idx, rows_per_wave, out);
    else
      hipLaunchKernelGGL(k_gather_probe<1>, grid, blk, 0, s, region, bytes, idx, rows_per_wave, out);
    ACC_HIP(hipGetLastError());
    return ACCSPMM_OK;
  });
}
