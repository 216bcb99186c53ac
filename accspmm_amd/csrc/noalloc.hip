// noalloc.hip — accspmm_spmm_noalloc: accspmm_spmm with every per-call
// device buffer in one file-scope device array, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  No counterpart in the reference.
//
// The array `scratch` (ACCSPMM_NOALLOC_SCRATCH_BYTES, 256-byte aligned)
// exists once per GPU the library runs on, for the life of the process.  Its
// device address is looked up per device with hipGetSymbolAddress on the
// first call there (with that device current) and only that address is ever
// handed to kernels or copies: host code never uses the array's name as
// memory.  One call lays out in it:
//   [ work units | split-window records ]  only when the handle has no
//                                           device schedule for the tile
//   [ split-window partial sums ]           pieces x RW x KC-padded K floats
//   [ converted B ]                         BF16 form, (cols + 1) rows
// Every region is written before it is read within the call (units and
// records by the upload, partials by the pieces, the converted B by the
// conversion pass), so nothing has to start at zero; the previous call's
// contents stay behind.  Calls that overlap on one GPU share the array and
// race: they must be ordered (one stream, or events between streams).
#include <algorithm>
#include <mutex>

#include "internal.hpp"

namespace accspmm {
accspmm_status spmm_fit_tile(const accspmm_metcf* h, const float* B, uint32_t ldb, float* C,
                             uint32_t ldc, uint32_t K, int form, const accspmm_tile* tile,
                             accspmm_tile* out);
accspmm_status spmm_check_operands(const float* B, uint32_t ldb, const float* C, uint32_t ldc,
                                   int form, const accspmm_tile& tile, int phases);
uint64_t panel_bound_words(const accspmm_metcf* h, const accspmm_tile& tile);
accspmm_status plan_schedule(const accspmm_metcf* h, const accspmm_tile& tile, hipStream_t s,
                             uint32_t* bounds_dev, Schedule* sch);
uint64_t schedule_unit_bytes(const Schedule& sch);
uint64_t schedule_sinfo_bytes(const Schedule& sch);
accspmm_status upload_schedule(const Schedule& sch, void* units, void* sinfo, hipStream_t s);
const Schedule* find_schedule(const accspmm_metcf* h, const accspmm_tile& tile);
accspmm_status spmm_run(const accspmm_metcf* h, const Schedule* sch, const WorkUnit* d_units,
                        const uint32_t* d_sinfo, const float* B, uint32_t ldb, float* C,
                        uint32_t ldc, uint32_t K, int form, const accspmm_tile& tile,
                        hipStream_t s, int phases, const SpmmWorkspace& w);
}  // namespace accspmm

using namespace accspmm;

namespace {

constexpr uint64_t kScratchBytes = ACCSPMM_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
accspmm_status scratch_address(void** out) {
  int dev = 0;
  ACC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ACCSPMM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    ACC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return ACCSPMM_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ACCSPMM_OK;
}

inline uint64_t align256(uint64_t x) { return (x + 255) / 256 * 256; }

}  // namespace

extern "C" accspmm_status accspmm_spmm_noalloc(const accspmm_metcf* h, const float* B,
                                               uint32_t ldb, float* C, uint32_t ldc,
                                               uint32_t K, accspmm_form form,
                                               const accspmm_tile* tile, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !B || !C || K == 0) return ACCSPMM_ERR_INVALID_ARG;
    if (ldb < K || ldc < K) return ACCSPMM_ERR_SHAPE;
    if ((int)form < 0 || (int)form > 2) return ACCSPMM_ERR_INVALID_ARG;
    // the resident-window sweep keeps a device plan per handle: not offered
    if (tile && tile->feat_per_lane && tile->sweep) return ACCSPMM_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    accspmm_tile t{};
    ACC_TRY(spmm_fit_tile(h, B, ldb, C, ldc, K, (int)form, tile, &t));
    t.sweep = 0;
    ACC_TRY(spmm_check_operands(B, ldb, C, ldc, (int)form, t, 3));
    void* base = nullptr;
    ACC_TRY(scratch_address(&base));
    char* sp = static_cast<char*>(base);

    std::lock_guard<std::mutex> lock(h->mu);
    // the handle's device schedule when an earlier call built it; otherwise
    // one planned on the host (kept in the handle's host memory for the next
    // calls) and uploaded into the scratch array
    const Schedule* sch = find_schedule(h, t);
    const WorkUnit* d_units = nullptr;
    const uint32_t* d_sinfo = nullptr;
    uint64_t off = 0;
    if (sch) {
      d_units = sch->units.as<WorkUnit>();
      d_sinfo = sch->split_info.as<uint32_t>();
    } else {
      const auto key = schedule_key(t);
      auto it = h->host_schedules.find(key);
      if (it == h->host_schedules.end()) {
        // the panel cut search writes W * (P - 1) words, before the layout
        // below is in use
        if (panel_bound_words(h, t) * 4 > kScratchBytes) return ACCSPMM_ERR_SCRATCH;
        auto plan = std::make_unique<Schedule>();
        ACC_TRY(plan_schedule(h, t, s, reinterpret_cast<uint32_t*>(sp), plan.get()));
        it = h->host_schedules.emplace(key, std::move(plan)).first;
      }
      sch = it->second.get();
      const uint64_t ub = align256(schedule_unit_bytes(*sch));
      const uint64_t sb = align256(schedule_sinfo_bytes(*sch));
      if (ub + sb > kScratchBytes) return ACCSPMM_ERR_SCRATCH;
      d_units = reinterpret_cast<const WorkUnit*>(sp);
      d_sinfo = reinterpret_cast<const uint32_t*>(sp + ub);
      off = ub + sb;
    }
    uint64_t pbytes = 0, cbytes = 0;
    spmm_workspace_need(h, sch, K, (int)form, t, &pbytes, &cbytes);
    if (off + pbytes + cbytes > kScratchBytes) return ACCSPMM_ERR_SCRATCH;
    if (off)
      ACC_TRY(upload_schedule(*sch, const_cast<WorkUnit*>(d_units),
                              const_cast<uint32_t*>(d_sinfo), s));
    SpmmWorkspace w;
    w.partials = reinterpret_cast<float*>(sp + off);
    w.bconv = sp + off + pbytes;
    return spmm_run(h, sch, d_units, d_sinfo, B, ldb, C, ldc, K, (int)form, t, s, 3, w);
  });
}
