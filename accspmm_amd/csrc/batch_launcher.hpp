// batch_launcher.hpp — host side of the persistent batched SpMM launch,
// shared by the batch API (batch.hip) and the serving loop (server.hip).
#pragma once
#include "internal.hpp"
#include "spmm_device.hpp"

namespace accspmm {

// What one problem needs inside a batch: its tile (the one its standalone
// call would use), its work-unit schedule and its workspace sizes.
struct ProblemPlan {
  uint32_t nb = 1, n_kc = 1, kc_padded = 16;
  const Schedule* sched = nullptr;
  uint64_t tasks = 0;          // units x feature chunks
  uint64_t bconv_bytes = 0;    // converted B (split-BF16 and BF16 forms)
  // the single-graph call's choices: B split in registers (split form),
  // waves per block and XCD mapping
  bool b_inreg = false;
  bool permuted = false;       // the graph has a column order: B is permuted
                               // by the conversion pass in every form
  uint32_t wpb = 4, xcd = 0;
  uint64_t partial_bytes = 0;  // split-window partial sums
  uint64_t workspace_bytes() const;
};

// Batch-wide choices on top of each problem's own tile, neither of which
// changes C's bits: the widest NB the persistent kernel runs (nb_cap) and
// the packing of short windows (pack: 0 keeps the tile's packing, or 64
// tiles when it has none; otherwise units of at least `pack` tiles).
struct BatchPolicy {
  uint32_t nb_cap = 4;
  uint32_t pack = 0;
};

accspmm_status problem_plan(const accspmm_problem& p, int form, hipStream_t s,
                            ProblemPlan* out, const BatchPolicy& policy = BatchPolicy());

// Problems of one NB, consecutive in the batch's order: one persistent
// kernel each.
struct BatchClass {
  uint32_t nb, p0, p1;
  uint64_t tasks;
  uint32_t grid = 1;
  bool buf = true;  // every problem's B gathered by buffer loads
};
// NB 4, 2 and 1 (the persistent kernels carry NB <= 4)
constexpr uint32_t kMaxBatchClasses = 3;

// Device buffers and pinned staging for one stream's batched launches.
// Buffers only grow; prepare() uploads asynchronously on the stream and
// launch() enqueues the kernels, so steady-state use neither allocates nor
// synchronizes.  prepare() must not be called while a previous launch()
// from the same launcher may still be running.
struct BatchLauncher {
  int form = 0;
  uint32_t rw = 8;
  DevBuf arena, dev_meta, head;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  hipEvent_t upload_event = nullptr;
  bool upload_event_valid = false;
  dev::SpmmArgs* descs = nullptr;
  uint32_t* prefix = nullptr;
  void* conv = nullptr;
  uint32_t* conv_prefix = nullptr;
  uint32_t* split_prefix = nullptr;  // split windows before each problem
  uint64_t total_tasks = 0;
  uint32_t nprob = 0, num_conv = 0, conv_blocks = 0;
  BatchClass cls[kMaxBatchClasses] = {};
  uint32_t nclass = 0;
  // a batch of one split-form problem whose B is split in registers runs
  // the single-graph kernel
  bool single = false;
  dev::SpmmArgs single_g{};
  uint32_t single_nb = 1, single_wpb = 4, single_units = 0, single_nkc = 1;
  bool single_packed = false;
  uint32_t total_split = 0;          // split windows of all problems
  BatchLauncher() = default;
  BatchLauncher(const BatchLauncher&) = delete;
  BatchLauncher& operator=(const BatchLauncher&) = delete;
  ~BatchLauncher();
  accspmm_status ensure_pinned(size_t bytes);
  accspmm_status prepare(const accspmm_problem* P, uint32_t n,
                         const ProblemPlan* plans, hipStream_t s);
  accspmm_status launch(hipStream_t s) const;
};

}  // namespace accspmm
