// scan.hip — device-wide prefix sums over u32 (reduce-then-scan, recursive
// over block totals).  Used by the METCF builder for window tile counts,
// tile value offsets and column-condensing ranks.
#include "internal.hpp"

namespace accspmm {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = (uint64_t)kScanThreads * kScanItems;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Scans one tile of kScanTile items; writes the tile total to sums[bid].
template <bool kInclusive>
__global__ __launch_bounds__(kScanThreads) void scan_tiles(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n,
    uint32_t* __restrict__ sums) {
  __shared__ uint32_t wave_tot[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  const int tid = threadIdx.x;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + (uint64_t)tid * kScanItems + j;
    v[j] = i < n ? in[i] : 0u;
    acc += v[j];
  }
  uint32_t incl = wave_incl_scan(acc);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  uint32_t wave_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    uint32_t t = wave_tot[w];
    if (w < wave) wave_off += t;
    total += t;
  }
  uint32_t run = wave_off + incl - acc;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + (uint64_t)tid * kScanItems + j;
    uint32_t x = v[j];
    if (kInclusive) run += x;
    if (i < n) out[i] = run;
    if (!kInclusive) run += x;
  }
  if (tid == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void add_tile_offsets(
    uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ offs) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  const uint32_t o = offs[blockIdx.x];
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
    if (i < n) out[i] += o;
  }
}

template <bool kInclusive>
accspmm_status scan_rec(const uint32_t* in, uint32_t* out, uint64_t n,
                        hipStream_t s) {
  if (n == 0) return ACCSPMM_OK;
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  DevBuf sums;
  ACC_TRY(sums.alloc(nb * sizeof(uint32_t)));
  hipLaunchKernelGGL(scan_tiles<kInclusive>, dim3((uint32_t)nb),
                     dim3(kScanThreads), 0, s, in, out, n,
                     sums.as<uint32_t>());
  ACC_HIP(hipGetLastError());
  if (nb > 1) {
    ACC_TRY(scan_rec<false>(sums.as<uint32_t>(), sums.as<uint32_t>(), nb, s));
    hipLaunchKernelGGL(add_tile_offsets, dim3((uint32_t)nb),
                       dim3(kScanThreads), 0, s, out, n,
                       sums.as<uint32_t>());
    ACC_HIP(hipGetLastError());
  }
  // the block-sum buffer is freed on return; make sure its users finished
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}

}  // namespace

accspmm_status exclusive_scan_u32(const uint32_t* in, uint32_t* out,
                                  uint64_t n, hipStream_t s) {
  return scan_rec<false>(in, out, n, s);
}
accspmm_status inclusive_scan_u32(const uint32_t* in, uint32_t* out,
                                  uint64_t n, hipStream_t s) {
  return scan_rec<true>(in, out, n, s);
}

}  // namespace accspmm
