// internal.hpp — shared internals of libaccspmm (handle layout, HIP checks).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/accspmm.h"
#include "guard.hpp"

#define ACC_HIP(expr)                                        \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) {                                  \
      (void)hipGetLastError();                               \
      return _e == hipErrorOutOfMemory ? ACCSPMM_ERR_OUT_OF_MEMORY \
                                       : ACCSPMM_ERR_HIP;    \
    }                                                        \
  } while (0)

#define ACC_TRY(expr)                          \
  do {                                         \
    accspmm_status _s = (expr);                \
    if (_s != ACCSPMM_OK) return _s;           \
  } while (0)

namespace accspmm {

constexpr uint32_t kColWindow = 8;
constexpr uint32_t kPad = 0xFFFFFFFFu;
// zero floats kept past the METCF values array (a 4-tile group's values are
// loaded as one 256-float span)
constexpr uint32_t kValuePad = 264;
// value offsets are 32-bit in the kernels (a group's span may run up to 512
// values past its first): nonzeros up to 2^32 - 1024
constexpr uint32_t kMaxNnz = 0xFFFFFFFFu - 1023u;

// Device buffer owned by RAII.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  accspmm_status alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return ACCSPMM_ERR_OUT_OF_MEMORY;
    }
    bytes = n;
    return ACCSPMM_OK;
  }
  // grows (never shrinks) to at least n bytes; contents not preserved
  accspmm_status reserve(size_t n) {
    if (p && bytes >= n) return ACCSPMM_OK;
    return alloc(n);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
};

// Work unit of the SpMM kernels: either a run of consecutive whole row
// windows [window, window + nwin) with their tiles [tile_begin, tile_end)
// (slot < 0; C rows are written directly, window by window), or one piece
// of a split window (nwin == 1, slot >= 0: a partial, combined in slot order
// by a second launch).  ngroups counts the unit's 4-tile groups; a group
// never spans two windows.  panel is the column panel of a piece (0 for
// whole windows).
struct alignas(16) WorkUnit {
  uint32_t window;
  uint32_t nwin;
  uint32_t tile_begin;
  uint32_t tile_end;
  int32_t slot;
  uint32_t ngroups;
  uint32_t panel, pad1;
};
// A unit with nwin == 0 does nothing: it pads the per-XCD unit lists of a
// pinned schedule (Schedule::pin_len).

struct Schedule {
  uint32_t split_tiles = 0xFFFFFFFFu;
  uint32_t pack_tiles = 0;
  uint32_t col_panels = 1;
  uint32_t num_units = 0;
  uint32_t num_split_windows = 0;
  uint32_t num_pieces = 0;  // units with slot >= 0
  bool packed = false;      // some unit spans several non-empty windows
  // > 0: units pinned to XCDs (tile.xcd_slice 3): XCD x runs units
  // [x * pin_len, (x + 1) * pin_len), the column panels p with p % 8 == x
  // and a load-balanced share of the whole windows, padded with no-op units
  uint32_t pin_len = 0;
  DevBuf units;             // WorkUnit[num_units]
  DevBuf split_info;        // per split window: window, first slot, piece count
  std::vector<WorkUnit> h_units;  // host copy (the sweep plan is derived from it)
  std::vector<uint32_t> h_sinfo;  // host copy of split_info
};

// Workspace of one accspmm_spmm call: split-window partials and the BF16
// form's converted B.
struct StreamWorkspace {
  DevBuf partials, bconv;
};
struct SpmmWorkspace {
  float* partials = nullptr;
  void* bconv = nullptr;
};

// Plan of the resident-window sweep (sweep.hip): the schedule's pieces
// regrouped per wave.  Wave v of round r owns up to M windows (their running
// sums live in LDS) and runs its pieces in (panel, window, slot) order.
struct SweepPlan {
  uint32_t P = 1, M = 1, NW = 0, R = 0;
  DevBuf pieces;   // {tile_begin, tile_end, window slot | split flag << 8} per piece
  DevBuf poff;     // [R*NW + 1] first piece of each (round, wave)
  DevBuf ngroups;  // [R*NW] 4-tile groups of each (round, wave)
  DevBuf wlist;    // [R*NW*M] windows of each (round, wave); kPad: none
};

}  // namespace accspmm

struct accspmm_metcf {
  int device = 0;
  uint32_t rows = 0, cols = 0, nnz = 0;
  uint32_t rw = 8, words = 1;
  uint32_t W = 0, T = 0, G = 0;
  uint32_t ibd = 0, balanced = 0, max_window_tiles = 0;
  uint32_t max_group_values = 0;  // values of the fullest 4-tile group
  // most non-empty windows gather B rows within 1/8 of B (a banded
  // matrix: neighbouring windows share B rows)
  bool banded = false;
  accspmm_build_opts opts{};
  accspmm::DevBuf rwo, tco, bits, a2b, data, grp, rix;
  // rows reordered at build (opts.reorder_rows or opts.symmetric_order):
  // format row i is input row row_perm[i]; empty otherwise
  accspmm::DevBuf row_perm;
  // symmetric_order: format column j is input column (B row) col_perm[j];
  // empty otherwise
  accspmm::DevBuf col_perm;
  std::vector<uint32_t> h_rwo;  // host copy of rowWindowOffset
  std::string hist_key;         // degree-histogram part of the tuning key
  // lazily built state (guarded by mu; handles are otherwise immutable)
  mutable std::mutex mu;
  // keyed by (split_tiles, pack_tiles, col_panels, pinned to XCDs)
  mutable std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>,
                   std::unique_ptr<accspmm::Schedule>>
      schedules;
  // host-only schedules of accspmm_spmm_noalloc (no device buffers), same key
  mutable std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>,
                   std::unique_ptr<accspmm::Schedule>>
      host_schedules;
  // keyed by (split_tiles, pack_tiles, col_panels, waves, slots)
  mutable std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>,
                   std::unique_ptr<accspmm::SweepPlan>>
      sweep_plans;
  // per-stream workspaces of accspmm_spmm (split partials, converted B):
  // calls on one stream share one, calls on different streams never do
  mutable std::map<hipStream_t, std::unique_ptr<accspmm::StreamWorkspace>> stream_ws;
  uint64_t device_bytes() const {
    return rwo.bytes + tco.bytes + bits.bytes + a2b.bytes + data.bytes +
           grp.bytes + rix.bytes + row_perm.bytes + col_perm.bytes;
  }
};

namespace accspmm {
// canon.hip: validates the CSR's row_ptr and writes the CSR with its rows
// in order of decreasing length (stable) into rp/ci/va; perm[i] = input row
// of row i
accspmm_status reorder_rows_by_length(const accspmm_csr_view* in, DevBuf& rp, DevBuf& ci,
                                      DevBuf& va, DevBuf& perm, hipStream_t s);
// canon.hip: the canonical CSR of P A P^T into rp/ci/va, perm[i] = input
// row (and column) at position i; order 1 reverse Cuthill-McKee, 2 rows by
// decreasing length.  The input must be square and canonical.
accspmm_status symmetric_permute(const accspmm_csr_view* in, uint32_t order, DevBuf& rp,
                                 DevBuf& ci, DevBuf& va, DevBuf& perm, hipStream_t s);
// scan.hip
accspmm_status exclusive_scan_u32(const uint32_t* in, uint32_t* out,
                                  uint64_t n, hipStream_t s);
accspmm_status inclusive_scan_u32(const uint32_t* in, uint32_t* out,
                                  uint64_t n, hipStream_t s);
// spmm.hip
// cache key of a tile's work-unit schedule: (split_tiles, pack_tiles,
// col_panels, pinned to XCDs)
std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> schedule_key(const accspmm_tile& tile);
accspmm_status get_schedule(const accspmm_metcf* h, const accspmm_tile& tile,
                            hipStream_t s, const Schedule** out);
// one SpMM call (spmm.hip).  phases: bit 0 converts B (BF16 form), bit 1
// runs the kernels; ws: the call's workspace, or null for the handle's
// workspace of stream s
struct SpmmWorkspace;
accspmm_status spmm_launch(const accspmm_metcf* h, const float* B, uint32_t ldb,
                           float* C, uint32_t ldc, uint32_t K, int form,
                           const accspmm_tile& tile, hipStream_t s, int phases,
                           const SpmmWorkspace* ws);
// workspace bytes of one call with this tile and schedule (spmm.hip)
void spmm_workspace_need(const accspmm_metcf* h, const Schedule* sch, uint32_t K, int form,
                         const accspmm_tile& tile, uint64_t* partial_bytes,
                         uint64_t* bconv_bytes);
// sets h->max_group_values and h->banded (builder.hip); after a build or a
// load
accspmm_status measure_group_values(accspmm_metcf* h, hipStream_t s);
}  // namespace accspmm
