// batch.hip — many (graph, B, C, K) SpMM problems in one persistent launch.
//
// No reference counterpart (the reference runs one matrix per process,
// src/mma_tf32.cu:185-243).  Every problem keeps the work units its
// standalone call would use (same tile, same split threshold), so its C is
// bit-identical alone and inside a batch.  The task space is implicit: task
// t belongs to problem p with prefix[p] <= t < prefix[p+1] and is unit
// (t - prefix[p]) / n_kc, feature chunk (t - prefix[p]) % n_kc of it.
// A run is a few kernels on one stream: reset of the queue heads, one
// conversion pass over every problem's B (BF16 form), one persistent kernel
// per NB class of problems whose waves pop tasks from the class's atomic
// head until it passes the class's last task (every wave reaches that test,
// so the grid always drains), and one reduction block per split window of
// any problem, which sums its piece partials in slot order (as the
// single-graph call does).  Problems are ordered by NB, widest first, then
// largest first.
//
// BatchLauncher is shared with the serving loop: planning is host
// arithmetic over cached per-graph state, uploads go through pinned staging
// with hipMemcpyAsync, and device buffers only grow, so steady-state
// launches neither allocate nor synchronize.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch_launcher.hpp"
#include "spmm_device.hpp"

namespace accspmm {
accspmm_status launch_single(const dev::SpmmArgs& g, int form, uint32_t rw, uint32_t nb,
                             uint32_t wpb, bool packed, uint32_t num_units, uint32_t n_kc,
                             hipStream_t s);
accspmm_status spmm_fit_tile(const accspmm_metcf* h, const float* B,
                             uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                             int form, const accspmm_tile* tile,
                             accspmm_tile* out);
namespace {
using namespace dev;

struct ConvDesc {
  const float* src;
  void* dst;
  const uint32_t* perm;  // the graph's column order (null: none)
  uint32_t ldb, K, rows, ldx;
  uint32_t vec;  // 16-byte loads allowed (ldb % 4 == 0, src 16-byte aligned)
};

// four-feature chunks of converted B per block (8 per thread)
constexpr uint32_t kConvChunks = 2048;

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// last index i in [0, n) with a[i] <= key (a[0] == 0 <= key)
__device__ inline uint32_t last_le(const uint32_t* __restrict__ a, uint32_t n,
                                   uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (a[mid] <= key) lo = mid; else hi = mid;
  }
  return lo;
}

// Every problem's B conversion in one launch (the converted forms): the
// blocks of problem p convert its chunks [(b - prefix[p]) * kConvChunks,
// ...) of the (rows + 1) * ldx / 4 (convert_chunk, the single-graph
// conversion's code, so the converted bits are the same).
template <int FORM>
__global__ __launch_bounds__(256) void k_batch_convert(
    const ConvDesc* __restrict__ d, uint32_t nconv,
    const uint32_t* __restrict__ prefix) {
  const uint32_t p = last_le(prefix, nconv, blockIdx.x);
  const ConvDesc c = d[p];
  const uint64_t total = (uint64_t)(c.rows + 1) * (c.ldx / 4);
  const uint64_t c0 = (uint64_t)(blockIdx.x - prefix[p]) * kConvChunks;
  const uint64_t c1 = min<uint64_t>(c0 + kConvChunks, total);
  for (uint64_t ch = c0 + threadIdx.x; ch < c1; ch += blockDim.x)
    convert_chunk<FORM>(c.src, c.rows, c.K, c.ldb, c.dst, c.ldx, c.vec, ch, c.perm);
}

// The persistent kernel of one (form, row window, NB) class: problems
// [p0, p1) of the batch (consecutive: the batch orders its problems by NB),
// whose tasks are prefix[p0] .. prefix[p1] - 1.  One NB per kernel keeps the
// register budget of the standalone kernel (three NB variants in one body
// spilled).
template <int FORM, int RW, int NB, bool BUF>
__global__ __launch_bounds__(256) void k_batch_spmm(
    const SpmmArgs* __restrict__ descs, const uint32_t* __restrict__ prefix,
    uint32_t p0, uint32_t p1, uint32_t* __restrict__ head) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t first = prefix[p0], total = prefix[p1] - first;
  __shared__ __align__(16) float lds_vals[4 * wave_lds<RW>()];
  float* lv = lds_vals + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * wave_lds<RW>();
  if (lane == 0) lv[group_values<RW>()] = 0.f;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(head, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (t >= total) return;
    t += first;
    const uint32_t p = p0 + last_le(prefix + p0, p1 - p0, t);
    const SpmmArgs g = descs[p];
    const uint32_t local = t - prefix[p];
    const WorkUnit u = g.units[local / g.n_kc];
    const uint32_t kc = local % g.n_kc;
    run_unit<FORM, NB, RW, true, BUF>(g, u, kc, g.B, g.ldb, g.C, g.ldc, g.K, g.rows, lv);
  }
}

template <int FORM, int RW, bool BUF>
void launch_batch_buf(uint32_t nb, dim3 grid, hipStream_t s, const SpmmArgs* descs,
                      const uint32_t* prefix, uint32_t p0, uint32_t p1, uint32_t* head) {
  const dim3 blk(256);
  if (nb == 4)
    hipLaunchKernelGGL((k_batch_spmm<FORM, RW, 4, BUF>), grid, blk, 0, s, descs, prefix, p0, p1,
                       head);
  else if (nb == 2)
    hipLaunchKernelGGL((k_batch_spmm<FORM, RW, 2, BUF>), grid, blk, 0, s, descs, prefix, p0, p1,
                       head);
  else
    hipLaunchKernelGGL((k_batch_spmm<FORM, RW, 1, BUF>), grid, blk, 0, s, descs, prefix, p0, p1,
                       head);
}

template <int FORM, int RW>
void launch_batch_class(const BatchClass& k, hipStream_t s, const SpmmArgs* descs,
                        const uint32_t* prefix, uint32_t* head) {
  if (k.buf)
    launch_batch_buf<FORM, RW, true>(k.nb, dim3(k.grid), s, descs, prefix, k.p0, k.p1, head);
  else
    launch_batch_buf<FORM, RW, false>(k.nb, dim3(k.grid), s, descs, prefix, k.p0, k.p1, head);
}

template <int RW>
__global__ __launch_bounds__(256) void k_batch_reduce(
    const SpmmArgs* __restrict__ descs, const uint32_t* __restrict__ sprefix,
    uint32_t nprob) {
  const uint32_t p = last_le(sprefix, nprob, blockIdx.x);
  const SpmmArgs& g = descs[p];
  reduce_split_window<RW>(g.split_info, blockIdx.x - sprefix[p], g.partials,
                          g.kc_padded, g.K, g.C, g.ldc, g.rows, g.row_map);
}

inline uint64_t align256(uint64_t x) { return (x + 255) / 256 * 256; }

}  // namespace

// Per-problem plan: tile, schedule and workspace sizes.
accspmm_status problem_plan(const accspmm_problem& p, int form, hipStream_t s,
                            ProblemPlan* out, const BatchPolicy& policy) {
  if (!p.graph || !p.B || !p.C || p.K == 0) return ACCSPMM_ERR_INVALID_ARG;
  if (p.ldb < p.K || p.ldc < p.K) return ACCSPMM_ERR_SHAPE;
  const accspmm_metcf* h = p.graph;
  accspmm_tile t;
  // the same tile (after alignment narrowing) as the single-graph call
  ACC_TRY(spmm_fit_tile(h, p.B, p.ldb, p.C, p.ldc, p.K, form, nullptr, &t));
  // the persistent kernel carries NB <= 4 (register budget), or the
  // policy's narrower cap; NB does not change C's bits
  const uint32_t cap = std::min<uint32_t>(4, std::max<uint32_t>(1, policy.nb_cap));
  while (t.feat_per_lane > cap) t.feat_per_lane /= 2;
  // it also packs short windows into units of >= 64 tiles (or the policy's
  // size) whatever the single-graph call does: its parallelism comes from
  // the many problems, and each task costs a queue pop (256-graph batch:
  // 2.65 ms packed, 5.33 unpacked).  Packing does not change C: a packed
  // unit sums each of its windows from zero, group by group, as a unit of
  // that window alone does.
  if (h->W && h->T / h->W < 64) {
    if (policy.pack) t.pack_tiles = policy.pack;
    else if (t.pack_tiles == 0) t.pack_tiles = 64;
  }
  const uint32_t KC = 16 * t.feat_per_lane;
  out->nb = t.feat_per_lane;
  out->n_kc = (p.K + KC - 1) / KC;
  out->kc_padded = out->n_kc * KC;
  {
    std::lock_guard<std::mutex> lock(h->mu);
    ACC_TRY(get_schedule(h, t, s, &out->sched));
  }
  out->tasks = (uint64_t)out->sched->num_units * out->n_kc;
  // a graph with a column order gathers a permuted copy of B in every form
  // (the split from the conversion pass is bit for bit the in-register one)
  out->permuted = h->col_perm.p != nullptr;
  out->b_inreg = form == kSPLIT && t.b_split == 2 && !out->permuted;
  out->wpb = t.waves_per_block ? t.waves_per_block : 4;
  out->xcd = t.xcd_slice;
  // the persistent kernel gathers a converted B in the split form (a batch
  // of one problem whose tile splits B in registers leaves it unused)
  out->bconv_bytes =
      form == kFP32 && !out->permuted
          ? 0
          : align256(((uint64_t)h->cols + 1) * out->kc_padded * (form == kBF16 ? 2 : 4));
  out->partial_bytes =
      align256((uint64_t)out->sched->num_pieces * h->rw * out->kc_padded * 4);
  return ACCSPMM_OK;
}

uint64_t ProblemPlan::workspace_bytes() const {
  return bconv_bytes + partial_bytes;
}

BatchLauncher::~BatchLauncher() {
  if (pinned) (void)hipHostFree(pinned);
}

accspmm_status BatchLauncher::ensure_pinned(size_t bytes) {
  if (pinned && pinned_bytes >= bytes) return ACCSPMM_OK;
  if (pinned) (void)hipHostFree(pinned);
  pinned = nullptr;
  size_t n = std::max<size_t>(bytes, 64 << 10);
  ACC_HIP(hipHostMalloc(&pinned, n, hipHostMallocDefault));
  pinned_bytes = n;
  return ACCSPMM_OK;
}

accspmm_status BatchLauncher::prepare(const accspmm_problem* P, uint32_t n,
                                      const ProblemPlan* plans, hipStream_t s) {
  nprob = 0;
  total_tasks = 0;
  single = false;
  if (n == 0) return ACCSPMM_OK;
  // one persistent kernel serves one row-window size
  const uint32_t rw0 = P[0].graph->rw;
  for (uint32_t i = 1; i < n; ++i)
    if (P[i].graph->rw != rw0) return ACCSPMM_ERR_UNSUPPORTED;
  rw = rw0;
  // one persistent kernel per NB class (widest first); largest problems
  // first within a class, so heavy work starts early
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (plans[a].nb != plans[b].nb) return plans[a].nb > plans[b].nb;
    return P[a].graph->T * (uint64_t)P[a].K > P[b].graph->T * (uint64_t)P[b].K;
  });
  // workspace layout: [partials + converted B per problem]
  uint64_t ws = 256, tot = 0, tot_split = 0;
  for (uint32_t i = 0; i < n; ++i) {
    ws += plans[i].bconv_bytes + plans[i].partial_bytes;
    tot += plans[i].tasks;
    tot_split += plans[i].sched->num_split_windows;
  }
  if (tot > 0xFFFFFFFFull || tot_split > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
  // host staging: descs | prefix (n+1) | conv descs | conv prefix (n+1) |
  // split prefix (n+1)
  const size_t d_bytes = sizeof(dev::SpmmArgs) * n;
  const size_t p_bytes = 4 * ((size_t)n + 1);
  const size_t c_bytes = sizeof(ConvDesc) * n;
  const size_t off_p = align256(d_bytes), off_c = off_p + align256(p_bytes),
               off_cp = off_c + align256(c_bytes), off_sp = off_cp + align256(p_bytes);
  const size_t stage = off_sp + align256(p_bytes);
  // the staging buffer is reused: wait until its previous upload consumed it
  if (upload_event_valid) ACC_HIP(hipEventSynchronize(upload_event));
  if (!upload_event)
    ACC_HIP(hipEventCreateWithFlags(&upload_event, hipEventDisableTiming));
  ACC_TRY(ensure_pinned(stage));
  ACC_TRY(dev_meta.reserve(stage));
  ACC_TRY(arena.reserve(ws));
  ACC_TRY(head.reserve(256));
  nclass = 0;
  char* hb = static_cast<char*>(pinned);
  auto* D = reinterpret_cast<dev::SpmmArgs*>(hb);
  auto* PF = reinterpret_cast<uint32_t*>(hb + off_p);
  auto* CV = reinterpret_cast<ConvDesc*>(hb + off_c);
  auto* CP = reinterpret_cast<uint32_t*>(hb + off_cp);
  auto* SP = reinterpret_cast<uint32_t*>(hb + off_sp);
  uint64_t bc = 0;
  uint32_t run = 0, cblocks = 0, nconv = 0, nsplit = 0;
  char* A = arena.as<char>();
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t i = order[j];
    const accspmm_problem& p = P[i];
    const ProblemPlan& q = plans[i];
    const accspmm_metcf* h = p.graph;
    dev::SpmmArgs& g = D[j];
    std::memset(&g, 0, sizeof(g));
    g.rwo = h->rwo.as<uint32_t>();
    g.tco = h->tco.as<uint32_t>();
    g.bits = h->bits.as<unsigned long long>();
    g.a2b = h->a2b.as<uint32_t>();
    g.data = h->data.as<float>();
    g.units = q.sched->units.as<WorkUnit>();
    g.split_info = q.sched->split_info.as<uint32_t>();
    g.C = p.C;
    g.row_map = h->row_perm.p ? h->row_perm.as<uint32_t>() : nullptr;
    g.ldc = p.ldc;
    g.rows = h->rows;
    g.K = p.K;
    g.num_units = q.sched->num_units;
    g.n_kc = q.n_kc;
    g.kc_padded = q.kc_padded;
    g.rw = h->rw;
    g.nb = q.nb;
      g.zero_row = h->cols;
    g.b_rows = form == kFP32 ? h->cols : h->cols + 1;
    g.pad_tile = h->T;
    g.wide = h->rw == 16 && h->max_group_values > dev::kSpanValues ? 1u : 0u;
    g.partials = reinterpret_cast<float*>(A + bc);
    bc += q.partial_bytes;
    // a batch of one split-BF16 problem whose tile splits B in registers
    // (the server's large requests) runs the single-graph kernel with the
    // batch's own workspace
    single = n == 1 && q.b_inreg;
    if (single) {
      // the single-graph kernel with the batch's own workspace
      g.xcd_slice = q.xcd == 2 ? 2u : (q.xcd == 1 && q.n_kc >= 2 && (8 % q.n_kc) == 0) ? 1u : 0u;
      single_g = g;
      single_g.B = p.B;
      single_g.ldb = p.ldb;
      single_nb = q.nb;
      single_wpb = q.wpb;
      single_units = q.sched->num_units;
      single_nkc = q.n_kc;
      single_packed = q.sched->packed;
    }
    if (single) g.b_rows = h->cols;
    if ((form == kFP32 && !q.permuted) || single) {
      g.B = p.B;
      g.ldb = p.ldb;
    } else {
      void* dst = A + bc;
      bc += q.bconv_bytes;
      g.B = dst;
      g.ldb = q.kc_padded;
      const uint32_t vec = (p.ldb % 4 == 0 && ((uintptr_t)p.B % 16) == 0) ? 1u : 0u;
      CV[nconv] = ConvDesc{static_cast<const float*>(p.B), dst,
                           h->col_perm.p ? h->col_perm.as<uint32_t>() : nullptr, p.ldb, p.K,
                           h->cols, q.kc_padded, vec};
      CP[nconv] = cblocks;
      const uint64_t chunks = ((uint64_t)h->cols + 1) * (q.kc_padded / 4);
      const uint64_t nblk = (chunks + kConvChunks - 1) / kConvChunks;
      if ((uint64_t)cblocks + nblk > 0x7FFFFFFFull) return ACCSPMM_ERR_UNSUPPORTED;
      cblocks += (uint32_t)nblk;
      ++nconv;
    }
    if (nclass == 0 || cls[nclass - 1].nb != q.nb) {
      cls[nclass] = BatchClass{q.nb, j, j + 1, 0};
      ++nclass;
    } else {
      cls[nclass - 1].p1 = j + 1;
    }
    cls[nclass - 1].tasks += q.tasks;
    // the class gathers by buffer loads when every problem's B allows it
    cls[nclass - 1].buf &= dev::buffer_ok(g.b_rows, (uint64_t)g.ldb * (form == kBF16 ? 2 : 4));
    if (single) single_g.b_rows = g.b_rows;
    PF[j] = run;
    run += (uint32_t)q.tasks;
    SP[j] = nsplit;
    nsplit += q.sched->num_split_windows;
  }
  PF[n] = run;
  CP[nconv] = cblocks;
  SP[n] = nsplit;
  ACC_HIP(hipMemcpyAsync(dev_meta.p, pinned, stage, hipMemcpyHostToDevice, s));
  ACC_HIP(hipEventRecord(upload_event, s));
  upload_event_valid = true;
  descs = reinterpret_cast<dev::SpmmArgs*>(dev_meta.as<char>());
  prefix = reinterpret_cast<uint32_t*>(dev_meta.as<char>() + off_p);
  conv = dev_meta.as<char>() + off_c;
  conv_prefix = reinterpret_cast<uint32_t*>(dev_meta.as<char>() + off_cp);
  split_prefix = reinterpret_cast<uint32_t*>(dev_meta.as<char>() + off_sp);
  total_split = nsplit;
  nprob = n;
  num_conv = nconv;
  conv_blocks = cblocks;
  total_tasks = run;
  int dv = 0, cus = 0;
  ACC_HIP(hipGetDevice(&dv));
  ACC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dv));
  for (uint32_t c = 0; c < nclass; ++c) {
    const uint64_t want = (cls[c].tasks + 3) / 4;
    cls[c].grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)cus * 4));
  }
  return ACCSPMM_OK;
}

accspmm_status BatchLauncher::launch(hipStream_t s) const {
  if (nprob == 0) return ACCSPMM_OK;
  if (single) {
    ACC_TRY(launch_single(single_g, kSPLITF, rw, single_nb, single_wpb, single_packed,
                          single_units, single_nkc, s));
    if (total_split) {
      const dim3 blk(256);
      if (rw == 8)
        hipLaunchKernelGGL(k_batch_reduce<8>, dim3(total_split), blk, 0, s, descs,
                           split_prefix, nprob);
      else
        hipLaunchKernelGGL(k_batch_reduce<16>, dim3(total_split), blk, 0, s, descs,
                           split_prefix, nprob);
      ACC_HIP(hipGetLastError());
    }
    return ACCSPMM_OK;
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, head.as<uint32_t>(),
                     (uint64_t)kMaxBatchClasses, 0u);
  ACC_HIP(hipGetLastError());
  if (conv_blocks) {
    if (form == kFP32)
      hipLaunchKernelGGL(k_batch_convert<kFP32>, dim3(conv_blocks), dim3(256), 0, s,
                         static_cast<const ConvDesc*>(conv), num_conv, conv_prefix);
    else if (form == kSPLIT)
      hipLaunchKernelGGL(k_batch_convert<kSPLIT>, dim3(conv_blocks), dim3(256), 0, s,
                         static_cast<const ConvDesc*>(conv), num_conv, conv_prefix);
    else
      hipLaunchKernelGGL(k_batch_convert<kBF16>, dim3(conv_blocks), dim3(256), 0, s,
                         static_cast<const ConvDesc*>(conv), num_conv, conv_prefix);
    ACC_HIP(hipGetLastError());
  }
  if (total_tasks == 0) return ACCSPMM_OK;
  auto* H = head.as<uint32_t>();
  const dim3 blk(256);
  for (uint32_t c = 0; c < nclass; ++c) {
    const BatchClass& k = cls[c];
    if (k.tasks == 0) continue;
    if (rw == 8) {
      if (form == kFP32) launch_batch_class<kFP32, 8>(k, s, descs, prefix, H + c);
      else if (form == kSPLIT) launch_batch_class<kSPLIT, 8>(k, s, descs, prefix, H + c);
      else launch_batch_class<kBF16, 8>(k, s, descs, prefix, H + c);
    } else {
      if (form == kFP32) launch_batch_class<kFP32, 16>(k, s, descs, prefix, H + c);
      else if (form == kSPLIT) launch_batch_class<kSPLIT, 16>(k, s, descs, prefix, H + c);
      else launch_batch_class<kBF16, 16>(k, s, descs, prefix, H + c);
    }
    ACC_HIP(hipGetLastError());
  }
  if (total_split) {
    if (rw == 8)
      hipLaunchKernelGGL(k_batch_reduce<8>, dim3(total_split), blk, 0, s, descs,
                         split_prefix, nprob);
    else
      hipLaunchKernelGGL(k_batch_reduce<16>, dim3(total_split), blk, 0, s, descs,
                         split_prefix, nprob);
    ACC_HIP(hipGetLastError());
  }
  return ACCSPMM_OK;
}

}  // namespace accspmm

using namespace accspmm;

struct accspmm_batch {
  // one persistent kernel serves one row-window size: the problems on
  // 8-row graphs go to L[0], those on 16-row graphs to L[1]
  BatchLauncher L[2];
  int form = 0;
  BatchPolicy policy;
  std::vector<accspmm_problem> probs;
  std::vector<ProblemPlan> plans;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~accspmm_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto& l : L)
      if (l.upload_event) (void)hipEventDestroy(l.upload_event);
  }
  // uploads each row-window group's descriptors (problems in input order)
  accspmm_status prepare_groups(hipStream_t s) {
    for (int gi = 0; gi < 2; ++gi) {
      std::vector<accspmm_problem> P;
      std::vector<ProblemPlan> Q;
      const uint32_t rw = gi ? 16u : 8u;
      for (size_t i = 0; i < probs.size(); ++i)
        if (probs[i].graph->rw == rw) {
          P.push_back(probs[i]);
          Q.push_back(plans[i]);
        }
      L[gi].form = form;
      ACC_TRY(L[gi].prepare(P.data(), (uint32_t)P.size(), Q.data(), s));
    }
    return ACCSPMM_OK;
  }
  accspmm_status launch(hipStream_t s) const {
    for (const auto& l : L) ACC_TRY(l.launch(s));
    return ACCSPMM_OK;
  }
};

extern "C" accspmm_status accspmm_batch_create(const accspmm_problem* problems,
                                               uint32_t count, accspmm_form form,
                                               accspmm_batch** out, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!out || (count && !problems)) return ACCSPMM_ERR_INVALID_ARG;
    if ((int)form < 0 || (int)form > 2) return ACCSPMM_ERR_INVALID_ARG;
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    auto* b = new (std::nothrow) accspmm_batch();
    if (!b) return ACCSPMM_ERR_OUT_OF_MEMORY;
    b->form = (int)form;
    b->probs.assign(problems, problems + count);
    b->plans.resize(count);
    accspmm_status st = ACCSPMM_OK;
    for (uint32_t i = 0; i < count && st == ACCSPMM_OK; ++i)
      st = problem_plan(problems[i], (int)form, s, &b->plans[i], b->policy);
    if (st == ACCSPMM_OK) st = b->prepare_groups(s);
    if (st == ACCSPMM_OK && hipStreamSynchronize(s) != hipSuccess) st = ACCSPMM_ERR_HIP;
    if (st != ACCSPMM_OK) {
      (void)hipStreamSynchronize(s);
      delete b;
      return st;
    }
    *out = b;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_batch_rebind(accspmm_batch* b,
                                               const accspmm_problem* problems,
                                               uint32_t count, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b || (count && !problems)) return ACCSPMM_ERR_INVALID_ARG;
    if (count != b->probs.size()) return ACCSPMM_ERR_SHAPE;
    hipStream_t s = (hipStream_t)stream;
    std::vector<ProblemPlan> plans(count);
    for (uint32_t i = 0; i < count; ++i) {
      if (problems[i].graph != b->probs[i].graph || problems[i].K != b->probs[i].K)
        return ACCSPMM_ERR_SHAPE;
      ACC_TRY(problem_plan(problems[i], b->form, s, &plans[i], b->policy));
      // the new operands must keep the planned tile (captured graphs replay it)
      if (plans[i].nb != b->plans[i].nb) return ACCSPMM_ERR_UNSUPPORTED;
    }
    ACC_HIP(hipStreamSynchronize(s));
    b->probs.assign(problems, problems + count);
    b->plans = plans;
    // same sizes => no buffer is reallocated, so a captured graph stays valid
    ACC_TRY(b->prepare_groups(s));
    ACC_HIP(hipStreamSynchronize(s));
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_batch_run(accspmm_batch* b, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b) return ACCSPMM_ERR_INVALID_ARG;
    return b->launch((hipStream_t)stream);
  });
}

extern "C" accspmm_status accspmm_batch_run_graph(accspmm_batch* b, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b) return ACCSPMM_ERR_INVALID_ARG;
    if (!b->exec) {
      if (!b->cap_stream)
        ACC_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
      ACC_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
      accspmm_status st = b->launch(b->cap_stream);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
      if (st != ACCSPMM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return ACCSPMM_ERR_HIP;
      b->graph = g;
      ACC_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    ACC_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return ACCSPMM_OK;
  });
}

namespace {
// Re-plans every problem of the batch under policy p and uploads the
// descriptors (no launch); synchronises the stream.
accspmm_status replan(accspmm_batch* b, const BatchPolicy& p, hipStream_t s) {
  std::vector<ProblemPlan> plans(b->probs.size());
  for (size_t i = 0; i < b->probs.size(); ++i)
    ACC_TRY(problem_plan(b->probs[i], b->form, s, &plans[i], p));
  b->policy = p;
  b->plans = plans;
  ACC_TRY(b->prepare_groups(s));
  ACC_HIP(hipStreamSynchronize(s));
  return ACCSPMM_OK;
}
}  // namespace

extern "C" accspmm_status accspmm_batch_autotune(accspmm_batch* b, void* stream,
                                                 float* best_ms) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b) return ACCSPMM_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    // a captured graph may hold buffers that re-planning replaces
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph) {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ACC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return ACCSPMM_ERR_HIP;
    }
    BatchPolicy best = b->policy;
    float bms = 1e30f;
    accspmm_status st = ACCSPMM_OK;
    for (uint32_t cap : {4u, 2u, 1u}) {
      for (uint32_t pack : {0u, 32u, 128u, 512u}) {
        BatchPolicy p;
        p.nb_cap = cap;
        p.pack = pack;
        st = replan(b, p, s);
        if (st != ACCSPMM_OK) break;
        for (int i = 0; i < 2 && st == ACCSPMM_OK; ++i) st = b->launch(s);
        if (st != ACCSPMM_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = ACCSPMM_ERR_HIP; break; }
        for (int i = 0; i < 5 && st == ACCSPMM_OK; ++i) st = b->launch(s);
        if (st != ACCSPMM_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          st = ACCSPMM_ERR_HIP;
          break;
        }
        if (ms / 5 < bms) {
          bms = ms / 5;
          best = p;
        }
      }
      if (st != ACCSPMM_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ACC_TRY(st);
    ACC_TRY(replan(b, best, s));
    if (best_ms) *best_ms = bms;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_batch_destroy(accspmm_batch* b) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b) return ACCSPMM_ERR_INVALID_ARG;
    (void)hipDeviceSynchronize();
    delete b;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_batch_num_tasks(const accspmm_batch* b,
                                                  uint64_t* tasks) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!b || !tasks) return ACCSPMM_ERR_INVALID_ARG;
    *tasks = b->L[0].total_tasks + b->L[1].total_tasks;
    return ACCSPMM_OK;
  });
}
