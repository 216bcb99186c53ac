// tune.cpp — tile-shape autotuner for the METCF SpMM kernels.
//
// The reference hard-codes its launch shape: warps per block K/32 or K/64
// (include/tf32_comp.hpp:19-25), one block per row window or AdpBME group
// (:27-31).  Here the shape (features per lane NB, waves per block, window
// splitting threshold, XCD feature slicing, window packing, column panels)
// is chosen per (degree histogram, K, form): candidates are timed with HIP
// events on the caller's stream and the fastest is cached in memory and,
// optionally, in a JSON file (also read at first use from
// $ACCSPMM_TUNE_CACHE).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace accspmm {
void default_tile(const accspmm_metcf* h, uint32_t K, int form,
                  accspmm_tile* t);
accspmm_status spmm_fit_tile(const accspmm_metcf* h, const float* B,
                             uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                             int form, const accspmm_tile* tile,
                             accspmm_tile* out);
}  // namespace accspmm

using namespace accspmm;

namespace {

std::mutex g_mu;
std::map<std::string, accspmm_tile> g_cache;
bool g_env_loaded = false;

// Tuning key: form, K and the handle's degree-histogram key (log2 buckets of
// tiles per window in 10% steps, log2 of the window count); graphs with the
// same key get the same tile.
std::string make_key(const accspmm_metcf* h, uint32_t K, int form) {
  return "f" + std::to_string(form) + "-k" + std::to_string(K) + "-" + h->hist_key;
}

// An entry the kernels can launch (a cache file is outside input: others
// are skipped, and the heuristic default serves their key)
bool valid_tile(const accspmm_tile& t) {
  const uint32_t nb = t.feat_per_lane;
  return (nb == 1 || nb == 2 || nb == 4 || nb == 8) && t.waves_per_block <= 8 &&
         t.xcd_slice <= 3 && t.sweep <= 1 && t.b_split <= 2 && t.col_panels <= 1024;
}

void load_file(const char* path, std::map<std::string, accspmm_tile>& into) {
  std::ifstream in(path);
  if (!in) return;
  std::string line;
  // one entry per line: "key": [nb, wpb, split, xcd, pack, panels, sweep, b_split],
  while (std::getline(in, line)) {
    size_t a = line.find('"'), b = line.find('"', a + 1);
    size_t l = line.find('['), r = line.find(']');
    if (a == std::string::npos || b == std::string::npos ||
        l == std::string::npos || r == std::string::npos)
      continue;
    std::string key = line.substr(a + 1, b - a - 1);
    unsigned v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int got = std::sscanf(line.c_str() + l + 1, "%u , %u , %u , %u , %u , %u , %u , %u",
                                &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
    const accspmm_tile t{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    // older entries lack col_panels, sweep and b_split (all 0)
    if (got >= 5 && valid_tile(t)) into[key] = t;
  }
}

// Writes the file's entries merged with this process's (which win), so
// processes tuning different graphs into one file keep each other's picks.
void save_file(const char* path) {
  std::map<std::string, accspmm_tile> merged;
  load_file(path, merged);
  for (auto& kv : g_cache) merged[kv.first] = kv.second;
  std::ofstream out(path);
  if (!out) return;
  out << "{\n";
  size_t i = 0;
  for (auto& kv : merged) {
    out << "  \"" << kv.first << "\": [" << kv.second.feat_per_lane << ", "
        << kv.second.waves_per_block << ", " << kv.second.split_tiles << ", "
        << kv.second.xcd_slice << ", " << kv.second.pack_tiles << ", "
        << kv.second.col_panels << ", " << kv.second.sweep << ", " << kv.second.b_split
        << "]"
        << (++i < merged.size() ? "," : "")
        << "\n";
  }
  out << "}\n";
}

void ensure_env() {
  if (g_env_loaded) return;
  g_env_loaded = true;
  const char* p = std::getenv("ACCSPMM_TUNE_CACHE");
  if (p && *p) load_file(p, g_cache);
}

}  // namespace

extern "C" accspmm_status accspmm_tune_key(const accspmm_metcf* h, uint32_t K,
                                           accspmm_form form, char* buf,
                                           size_t cap) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !buf || cap == 0) return ACCSPMM_ERR_INVALID_ARG;
    std::string k = make_key(h, K, (int)form);
    std::snprintf(buf, cap, "%s", k.c_str());
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_tile_select(const accspmm_metcf* h,
                                              uint32_t K, accspmm_form form,
                                              accspmm_tile* out) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !out || K == 0) return ACCSPMM_ERR_INVALID_ARG;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      ensure_env();
      auto it = g_cache.find(make_key(h, K, (int)form));
      if (it != g_cache.end()) {
        *out = it->second;
        return ACCSPMM_OK;
      }
    }
    default_tile(h, K, (int)form, out);
    return ACCSPMM_OK;
  });
}

extern "C" accspmm_status accspmm_autotune(const accspmm_metcf* h,
                                           const float* B, uint32_t ldb,
                                           float* C, uint32_t ldc, uint32_t K,
                                           accspmm_form form,
                                           const char* cache_path,
                                           accspmm_tile* best,
                                           float* best_ms, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!h || !B || !C || K == 0) return ACCSPMM_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t mean = h->W ? std::max<uint32_t>(1, h->T / h->W) : 1;
    std::vector<uint32_t> splits = {0};
    for (uint32_t f : {1u, 2u, 4u, 8u}) {
      uint32_t sp = std::max<uint32_t>(32, mean * f);
      if (h->max_window_tiles > sp) splits.push_back(sp);
    }
    hipEvent_t e0, e1;
    ACC_HIP(hipEventCreate(&e0));
    ACC_HIP(hipEventCreate(&e1));
    // median of 5 timed launches after 2 warm-up launches; < 0 if rejected
    auto time_tile = [&](const accspmm_tile& c) -> float {
      if (spmm_launch(h, B, ldb, C, ldc, K, (int)form, c, s, 3, nullptr) != ACCSPMM_OK)
        return -1.f;  // e.g. alignment not met
      for (int i = 0; i < 2; ++i) spmm_launch(h, B, ldb, C, ldc, K, (int)form, c, s, 3, nullptr);
      std::vector<float> t;
      for (int i = 0; i < 5; ++i) {
        (void)hipEventRecord(e0, s);
        spmm_launch(h, B, ldb, C, ldc, K, (int)form, c, s, 3, nullptr);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.push_back(ms);
      }
      std::sort(t.begin(), t.end());
      return t[2];
    };
    float bms = 1e30f;
    accspmm_tile bt{};
    auto consider = [&](const accspmm_tile& c) {
      const float ms = time_tile(c);
      if (ms >= 0.f && ms < bms) {
        bms = ms;
        bt = c;
      }
    };
    // stage 1: lane width, waves per block, split, XCD slicing, packing
    for (uint32_t nb : {1u, 2u, 4u, 8u}) {
      if (nb > 1 && 16 * nb > 2 * ((K + 15) / 16 * 16)) continue;
      const uint32_t n_kc = (K + 16 * nb - 1) / (16 * nb);
      for (uint32_t wpb : {1u, 2u, 4u})
        for (uint32_t sp : splits)
          for (uint32_t xcd : {0u, 1u, 2u})
            for (uint32_t pack : {0u, 64u}) {
              if (xcd == 1 && (n_kc == 1 || (8 % n_kc) != 0)) continue;
              if (pack && mean >= 64) continue;  // windows already long
              consider(accspmm_tile{nb, wpb, sp, xcd, pack, 0, 0, 0});
            }
    }
    // stage 2: column panels around the best shape, when the B columns one
    // XCD gathers exceed its 4 MiB L2
    if (bms < 1e30f) {
      const uint32_t KC = 16 * bt.feat_per_lane;
      const uint32_t n_kc = (K + KC - 1) / KC;
      const uint64_t es = form == ACCSPMM_FORM_BF16 ? 2 : 4;
      const uint64_t row_bytes = (bt.xcd_slice == 1 ? KC : (uint64_t)n_kc * KC) * es;
      if ((uint64_t)h->cols * row_bytes > (4ull << 20)) {
        const accspmm_tile base = bt;
        for (uint32_t p : {2u, 4u, 8u, 16u, 32u})
          for (uint32_t sp : splits) {
            accspmm_tile c = base;
            c.col_panels = p;
            c.split_tiles = sp;
            consider(c);
          }
        // panels pinned to XCDs (each B row cached by one L2)
        for (uint32_t p : {8u, 16u}) {
          accspmm_tile c = base;
          c.col_panels = p;
          c.xcd_slice = 3;
          consider(c);
        }
      }
    }
    // stage 3 (split-BF16): both ways of splitting B, around the best shape
    if (bms < 1e30f && form == ACCSPMM_FORM_SPLIT_BF16) {
      const accspmm_tile base = bt;
      for (uint32_t bs : {1u, 2u}) {
        accspmm_tile c = base;
        c.b_split = bs;
        consider(c);
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bms >= 1e30f) return ACCSPMM_ERR_UNSUPPORTED;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      ensure_env();
      g_cache[make_key(h, K, (int)form)] = bt;
      if (cache_path && *cache_path) save_file(cache_path);
    }
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return ACCSPMM_OK;
  });
}

// Layout search: the row window (8, 16) and the row order (input, by
// decreasing length) are build options, so each candidate is a build.  Each
// layout runs its library tile (tuned pick or default) and the median of 5
// timed launches after 2 warm-up launches decides.
extern "C" accspmm_status accspmm_autotune_layout(const accspmm_csr_view* csr,
                                                  const float* B, uint32_t ldb,
                                                  float* C, uint32_t ldc, uint32_t K,
                                                  accspmm_form form,
                                                  accspmm_build_opts* best,
                                                  float* best_ms, void* stream) {
  return accspmm::guard([&]() -> accspmm_status {
    if (!csr || !B || !C || !best || K == 0) return ACCSPMM_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    accspmm_build_opts base;
    accspmm_build_opts_default(&base);
    hipEvent_t e0, e1;
    ACC_HIP(hipEventCreate(&e0));
    ACC_HIP(hipEventCreate(&e1));
    float bms = 1e30f;
    uint32_t brw = 8, bro = 0, bso = 0;
    accspmm_status err = ACCSPMM_OK;
    // row orders: input, rows by length, and for square matrices reverse
    // Cuthill-McKee (rows and columns; the order that undoes a scattered
    // labelling of a mesh).  Rows and columns by length was measured slower
    // than rows by length alone on every graph tried, so it is not a
    // candidate.
    const uint32_t norders = csr->rows == csr->cols ? 3u : 2u;
    for (uint32_t rw : {8u, 16u}) {
      for (uint32_t ord = 0; ord < norders; ++ord) {
        const uint32_t ro = ord == 1 ? 1u : 0u, so = ord == 2 ? 1u : 0u;
        accspmm_build_opts o = base;
        o.row_window = rw;
        o.reorder_rows = ro;
        o.symmetric_order = so;
        o.build_adpbme = 0;  // the SpMM path does not read the AdpBME groups
        accspmm_metcf* h = nullptr;
        err = accspmm_metcf_build(csr, &o, &h, stream);
        if (err != ACCSPMM_OK) break;
        accspmm_tile t{};
        err = spmm_fit_tile(h, B, ldb, C, ldc, K, (int)form, nullptr, &t);
        std::vector<float> ms;
        for (int i = 0; i < 7 && err == ACCSPMM_OK; ++i) {
          (void)hipEventRecord(e0, s);
          err = spmm_launch(h, B, ldb, C, ldc, K, (int)form, t, s, 3, nullptr);
          (void)hipEventRecord(e1, s);
          (void)hipEventSynchronize(e1);
          float x = 0.f;
          (void)hipEventElapsedTime(&x, e0, e1);
          if (i >= 2) ms.push_back(x);
        }
        (void)accspmm_metcf_destroy(h);
        if (err != ACCSPMM_OK) break;
        std::sort(ms.begin(), ms.end());
        if (ms[2] < bms) {
          bms = ms[2];
          brw = rw;
          bro = ro;
          bso = so;
        }
      }
      if (err != ACCSPMM_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != ACCSPMM_OK) return err;
    *best = base;
    best->row_window = brw;
    best->reorder_rows = bro;
    best->symmetric_order = bso;
    if (best_ms) *best_ms = bms;
    return ACCSPMM_OK;
  });
}

extern "C" const char* accspmm_status_string(accspmm_status s) {
  switch (s) {
    case ACCSPMM_OK: return "ok";
    case ACCSPMM_ERR_INVALID_ARG: return "invalid argument";
    case ACCSPMM_ERR_NOT_CANONICAL: return "CSR row not strictly increasing in column";
    case ACCSPMM_ERR_HIP: return "HIP runtime error";
    case ACCSPMM_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ACCSPMM_ERR_UNSUPPORTED: return "unsupported option";
    case ACCSPMM_ERR_BUDGET: return "HBM budget exhausted";
    case ACCSPMM_ERR_NOT_FOUND: return "not found";
    case ACCSPMM_ERR_QUEUE_FULL: return "request queue full";
    case ACCSPMM_ERR_MTX: return "Matrix Market parse error";
    case ACCSPMM_ERR_NOT_READY: return "not ready";
    case ACCSPMM_ERR_SHAPE: return "operand shapes do not agree";
    case ACCSPMM_ERR_FILE: return "file not writable, or METCF file unreadable, corrupt or of another format";
    case ACCSPMM_ERR_SCRATCH: return "allocation-free call needs more than the scratch array holds";
  }
  return "unknown status";
}

extern "C" const char* accspmm_version(void) {
  return "accspmm-mi355x 0.1 gfx950";
}
