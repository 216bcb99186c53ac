// server.hip — continuous-batching SpMM serving loop.
//
// No reference counterpart.  Graphs are registered once and stay resident in
// METCF form; requests (graph id + caller-owned B and C) are queued and a
// dispatcher thread admits them into persistent batched launches:
//   * admission: a request's workspace (converted B, split partials) is
//     charged against the HBM budget together with the resident graphs;
//     requests are taken in arrival order while they fit;
//   * continuous batching: up to `max_inflight` launches run at once, each
//     on its own stream; whenever one completes, pending requests (including
//     those that arrived meanwhile) form the next launch, so new work
//     overlaps work already on the GPU;
//   * size classes: a request of nnz * K >= large_work runs in a launch of
//     its own, one large launch at a time, so small requests are never
//     queued behind (or batched with) a large one and do not wait for it to
//     finish; each class is served in arrival order and the older of the
//     two queue heads goes first;
//   * completion: the dispatcher polls each launch's event and stamps the
//     requests' latency (submit -> observed completion).
// In one_per_launch mode every request is its own standalone SpMM call and
// launches do not overlap (the baseline the batched mode is measured against).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_launcher.hpp"

namespace accspmm {
accspmm_status spmm_fit_tile(const accspmm_metcf* h, const float* B,
                             uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                             int form, const accspmm_tile* tile,
                             accspmm_tile* out);
}  // namespace accspmm

using namespace accspmm;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  uint32_t gid = 0;
  accspmm_problem prob{};
  ProblemPlan plan;
  uint64_t ws = 0;
  Clock::time_point t0;
  double latency = 0;
  int state = 0;  // 0 pending, 1 in flight, 2 done, 3 failed
  bool large = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  BatchLauncher L;
  std::vector<uint64_t> rids;
  uint64_t ws = 0;
  bool busy = false;
  bool large = false;  // holds one large request
};

}  // namespace

struct accspmm_server {
  accspmm_server_opts o{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_idle;
  struct Graph {
    accspmm_metcf* h = nullptr;
    uint64_t bytes = 0;
    uint32_t users = 0;  // pending + in-flight requests
  };
  std::map<uint32_t, Graph> graphs;
  uint32_t next_gid = 1;
  std::unordered_map<uint64_t, Req> reqs;
  uint64_t next_rid = 1;
  std::deque<uint64_t> pending, pending_large;
  std::vector<std::unique_ptr<Slot>> slots;
  hipStream_t setup = nullptr;
  uint64_t resident = 0, inflight = 0, peak_inflight = 0;
  uint64_t submitted = 0, completed = 0, launches = 0, rejected_budget = 0;
  double busy_s = 0;
  bool paused = false, stop = false;
  std::thread worker;

  void run();
  bool poll_completions();
  bool admit();
  void finish_slot(Slot& sl, bool ok);
};

void accspmm_server::finish_slot(Slot& sl, bool ok) {
  const auto now = Clock::now();
  for (uint64_t rid : sl.rids) {
    auto it = reqs.find(rid);
    if (it == reqs.end()) continue;
    Req& r = it->second;
    r.latency = std::chrono::duration<double>(now - r.t0).count();
    r.state = ok ? 2 : 3;
    auto g = graphs.find(r.gid);
    if (g != graphs.end() && g->second.users) g->second.users--;
    ++completed;
  }
  inflight -= sl.ws;
  sl.ws = 0;
  sl.rids.clear();
  sl.busy = false;
}

bool accspmm_server::poll_completions() {
  bool any = false;
  for (auto& sp : slots) {
    Slot& sl = *sp;
    if (!sl.busy) continue;
    hipError_t e = hipEventQuery(sl.done);
    if (e == hipErrorNotReady) continue;
    finish_slot(sl, e == hipSuccess);
    any = true;
  }
  if (any) cv_idle.notify_all();
  return any;
}

bool accspmm_server::admit() {
  bool any = false;
  const uint64_t budget = o.hbm_budget_bytes > resident ? o.hbm_budget_bytes - resident : 0;
  for (auto& sp : slots) {
    if (paused || (pending.empty() && pending_large.empty())) break;
    Slot& sl = *sp;
    if (sl.busy) continue;
    uint32_t large_busy = 0;
    for (auto& q : slots) large_busy += (q->busy && q->large) ? 1u : 0u;
    // a large request may take this slot if no other large one runs and
    // its workspace fits; the older queue head goes first.  One large
    // launch saturates the memory system by itself: two at once only
    // stretch both, and first-in first-out finishes the first sooner (the
    // serving pool's batched p99 6.2-6.4 ms against 10.7-11.3 ms with two,
    // at the same throughput)
    bool take_large = false;
    if (!pending_large.empty() && large_busy == 0) {
      const Req& rl = reqs[pending_large.front()];
      const bool fits = inflight + rl.ws <= budget;
      const bool older = pending.empty() || rl.t0 <= reqs[pending.front()].t0;
      take_large = fits && older;
    }
    std::vector<accspmm_problem> probs;
    std::vector<ProblemPlan> plans;
    uint64_t ws = 0;
    if (take_large) {
      Req& r = reqs[pending_large.front()];
      ws = r.ws;
      probs.push_back(r.prob);
      plans.push_back(r.plan);
      sl.rids.push_back(pending_large.front());
      r.state = 1;
      pending_large.pop_front();
    } else {
      const uint32_t cap = o.one_per_launch ? 1u : std::max(1u, o.max_batch);
      uint32_t rw = 0;
      while (!pending.empty() && probs.size() < cap) {
        Req& r = reqs[pending.front()];
        const uint32_t grw = r.prob.graph->rw;
        if (rw && grw != rw) break;  // one launch serves one row-window size
        if (inflight + ws + r.ws > budget) break;  // FIFO: wait for memory
        rw = grw;
        ws += r.ws;
        probs.push_back(r.prob);
        plans.push_back(r.plan);
        sl.rids.push_back(pending.front());
        r.state = 1;
        pending.pop_front();
      }
    }
    if (probs.empty()) continue;
    accspmm_status st;
    if (o.one_per_launch) {
      accspmm_tile t;
      const accspmm_problem& p = probs[0];
      st = spmm_fit_tile(p.graph, p.B, p.ldb, p.C, p.ldc, p.K, (int)o.form, nullptr, &t);
      if (st == ACCSPMM_OK)
        st = spmm_launch(p.graph, p.B, p.ldb, p.C, p.ldc, p.K, (int)o.form, t,
                         sl.stream, 3, nullptr);
    } else {
      // large requests too go through the slot's own launcher (a batch of
      // one): its workspace is the slot's, so two requests for one graph
      // on two streams never share a conversion buffer
      sl.L.form = (int)o.form;
      st = sl.L.prepare(probs.data(), (uint32_t)probs.size(), plans.data(), sl.stream);
      if (st == ACCSPMM_OK) st = sl.L.launch(sl.stream);
    }
    if (st == ACCSPMM_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess)
      st = ACCSPMM_ERR_HIP;
    sl.ws = ws;
    sl.large = take_large;
    inflight += ws;
    peak_inflight = std::max(peak_inflight, inflight);
    sl.busy = true;
    ++launches;
    if (st != ACCSPMM_OK) finish_slot(sl, false);
    any = true;
  }
  return any;
}

void accspmm_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  auto busy_since = Clock::now();
  bool was_busy = false;
  for (;;) {
    poll_completions();
    admit();
    bool busy_now = false;
    for (auto& sp : slots) busy_now |= sp->busy;
    if (busy_now && !was_busy) busy_since = Clock::now();
    if (!busy_now && was_busy)
      busy_s += std::chrono::duration<double>(Clock::now() - busy_since).count();
    was_busy = busy_now;
    if (stop && pending.empty() && pending_large.empty() && !busy_now) break;
    if (busy_now) {
      // short poll: completion latency is bounded by this interval
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    } else {
      cv_work.wait_for(lock, std::chrono::milliseconds(2));
    }
  }
}

extern "C" void accspmm_server_opts_default(accspmm_server_opts* o) {
  if (!o) return;
  o->hbm_budget_bytes = 64ull << 30;
  o->max_batch = 64;
  o->max_queue = 1u << 16;
  o->max_inflight = 2;
  o->form = ACCSPMM_FORM_SPLIT_BF16;
  o->one_per_launch = 0;
  o->large_work = 1ull << 30;
}

extern "C" accspmm_status accspmm_server_create(const accspmm_server_opts* opts,
                                                accspmm_server** out) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!out) return ACCSPMM_ERR_INVALID_ARG;
    *out = nullptr;
    accspmm_server_opts o;
    if (opts) o = *opts; else accspmm_server_opts_default(&o);
    if (o.form > 2 || o.max_inflight == 0 || o.max_inflight > 8 || o.max_queue == 0)
      return ACCSPMM_ERR_INVALID_ARG;
    if (o.one_per_launch) o.max_inflight = 1;
    auto s = std::make_unique<accspmm_server>();
    s->o = o;
    ACC_HIP(hipGetDevice(&s->device));
    ACC_HIP(hipStreamCreateWithFlags(&s->setup, hipStreamNonBlocking));
    for (uint32_t i = 0; i < o.max_inflight; ++i) {
      auto sl = std::make_unique<Slot>();
      ACC_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
      ACC_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
      s->slots.push_back(std::move(sl));
    }
    accspmm_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return ACCSPMM_OK;
  });
}

namespace {
// Takes ownership of h: prepares its default schedule (and the BF16 forms'
// split values) so the first request does not, then charges it against the
// budget.
accspmm_status register_handle(accspmm_server* s, accspmm_metcf* h, uint32_t* graph_id) {
  accspmm_tile t;
  accspmm_status st = accspmm_tile_select(h, 128, (accspmm_form)s->o.form, &t);
  if (st == ACCSPMM_OK) {
    std::lock_guard<std::mutex> hl(h->mu);
    const Schedule* sch = nullptr;
    st = get_schedule(h, t, s->setup, &sch);
  }
  if (st != ACCSPMM_OK) {
    accspmm_metcf_destroy(h);
    return st;
  }
  const uint64_t bytes = h->device_bytes();
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->resident + bytes > s->o.hbm_budget_bytes) {
    s->rejected_budget++;
    accspmm_metcf_destroy(h);
    return ACCSPMM_ERR_BUDGET;
  }
  s->resident += bytes;
  uint32_t id = s->next_gid++;
  s->graphs[id] = accspmm_server::Graph{h, bytes, 0};
  *graph_id = id;
  return ACCSPMM_OK;
}
}  // namespace

extern "C" accspmm_status accspmm_server_register_opts(accspmm_server* s,
                                                       const accspmm_csr_view* csr,
                                                       const accspmm_build_opts* opts,
                                                       uint32_t* graph_id) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s || !csr || !graph_id) return ACCSPMM_ERR_INVALID_ARG;
    accspmm_build_opts bo;
    if (opts) {
      bo = *opts;
    } else {
      // served graphs are never exported: the library's automatic layout
      accspmm_build_opts_default(&bo);
      bo.row_window = 0;
      bo.reorder_rows = 2;
    }
    bo.build_adpbme = 0;  // the served kernels use work units, not AdpBME groups
    accspmm_metcf* h = nullptr;
    ACC_TRY(accspmm_metcf_build(csr, &bo, &h, s->setup));
    return register_handle(s, h, graph_id);
  });
}

extern "C" accspmm_status accspmm_server_register(accspmm_server* s,
                                                  const accspmm_csr_view* csr,
                                                  uint32_t* graph_id) {
  return accspmm::guard([&]() -> accspmm_status {
    return accspmm_server_register_opts(s, csr, nullptr, graph_id);
  });
}

extern "C" accspmm_status accspmm_server_register_file(accspmm_server* s, const char* path,
                                                       uint32_t* graph_id) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s || !path || !graph_id) return ACCSPMM_ERR_INVALID_ARG;
    accspmm_metcf* h = nullptr;
    ACC_TRY(accspmm_metcf_load(path, &h, s->setup));
    return register_handle(s, h, graph_id);
  });
}

extern "C" accspmm_status accspmm_server_unregister(accspmm_server* s,
                                                    uint32_t graph_id) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s) return ACCSPMM_ERR_INVALID_ARG;
    accspmm_metcf* h = nullptr;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      auto it = s->graphs.find(graph_id);
      if (it == s->graphs.end()) return ACCSPMM_ERR_NOT_FOUND;
      if (it->second.users) return ACCSPMM_ERR_NOT_READY;
      h = it->second.h;
      s->resident -= it->second.bytes;
      s->graphs.erase(it);
    }
    return accspmm_metcf_destroy(h);
  });
}

extern "C" accspmm_status accspmm_server_submit(accspmm_server* s, uint32_t graph_id,
                                                const float* B, uint32_t ldb, float* C,
                                                uint32_t ldc, uint32_t K,
                                                uint64_t* request_id) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s || !request_id || !B || !C || K == 0) return ACCSPMM_ERR_INVALID_ARG;
    const auto t0 = Clock::now();
    accspmm_metcf* h = nullptr;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      auto it = s->graphs.find(graph_id);
      if (it == s->graphs.end()) return ACCSPMM_ERR_NOT_FOUND;
      if (s->pending.size() + s->pending_large.size() >= s->o.max_queue)
        return ACCSPMM_ERR_QUEUE_FULL;
      h = it->second.h;
      it->second.users++;  // keeps the graph registered until completion
    }
    Req r;
    r.gid = graph_id;
    r.prob = accspmm_problem{h, B, ldb, C, ldc, K};
    r.t0 = t0;
    accspmm_status st = problem_plan(r.prob, (int)s->o.form, s->setup, &r.plan);
    std::lock_guard<std::mutex> lock(s->mu);
    auto g = s->graphs.find(graph_id);
    if (st != ACCSPMM_OK) {
      if (g != s->graphs.end()) g->second.users--;
      return st;
    }
    r.ws = s->o.one_per_launch ? 0 : r.plan.workspace_bytes();
    if (s->o.one_per_launch) {
      // the standalone path converts into the graph's own workspace
      r.ws = r.plan.bconv_bytes + r.plan.partial_bytes;
    }
    if (s->resident + r.ws > s->o.hbm_budget_bytes) {
      g->second.users--;
      s->rejected_budget++;
      return ACCSPMM_ERR_BUDGET;  // can never fit
    }
    const uint64_t id = s->next_rid++;
    const uint64_t large_work = s->o.large_work ? s->o.large_work : (1ull << 30);
    r.large = !s->o.one_per_launch && (uint64_t)h->nnz * K >= large_work;
    s->reqs.emplace(id, r);
    (r.large ? s->pending_large : s->pending).push_back(id);
    s->submitted++;
    *request_id = id;
    s->cv_work.notify_one();
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_server_poll(accspmm_server* s, uint64_t request_id,
                                              double* latency_s) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s) return ACCSPMM_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return ACCSPMM_ERR_NOT_FOUND;
    if (it->second.state < 2) return ACCSPMM_ERR_NOT_READY;
    const bool ok = it->second.state == 2;
    if (latency_s) *latency_s = it->second.latency;
    s->reqs.erase(it);
    return ok ? ACCSPMM_OK : ACCSPMM_ERR_HIP;
  });
}

extern "C" accspmm_status accspmm_server_drain(accspmm_server* s) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s) return ACCSPMM_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv_work.notify_one();
    s->cv_idle.wait(lock, [s] {
      if ((!s->pending.empty() || !s->pending_large.empty()) && !s->paused) return false;
      for (auto& sp : s->slots)
        if (sp->busy) return false;
      return true;
    });
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_server_pause(accspmm_server* s, int paused) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s) return ACCSPMM_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(s->mu);
    s->paused = paused != 0;
    s->cv_work.notify_one();
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_server_get_stats(accspmm_server* s,
                                                   accspmm_server_stats* st) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s || !st) return ACCSPMM_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(s->mu);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->launches = s->launches;
    st->rejected_budget = s->rejected_budget;
    st->resident_bytes = s->resident;
    st->peak_inflight_bytes = s->peak_inflight;
    st->busy_s = s->busy_s;
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_server_destroy(accspmm_server* s) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!s) return ACCSPMM_ERR_INVALID_ARG;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      s->stop = true;
      s->paused = false;
      s->cv_work.notify_one();
    }
    if (s->worker.joinable()) s->worker.join();
    for (auto& sp : s->slots) {
      (void)hipStreamSynchronize(sp->stream);
      (void)hipEventDestroy(sp->done);
      if (sp->L.upload_event) (void)hipEventDestroy(sp->L.upload_event);
      (void)hipStreamDestroy(sp->stream);
    }
    s->slots.clear();
    for (auto& kv : s->graphs) accspmm_metcf_destroy(kv.second.h);
    if (s->setup) (void)hipStreamDestroy(s->setup);
    delete s;
    return ACCSPMM_OK;
  });
}
