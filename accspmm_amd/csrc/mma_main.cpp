// mma_main.cpp — the `./mma <matrix.mtx> <feature_dim>` driver, a drop-in for
// the reference's src/mma_tf32.cu:185-243.
//
// Same positional arguments, exit codes (4 on missing arguments, 3 on a read
// failure), stdout lines and result.csv row "name,K,ms,GFLOPs".  As in the
// reference: B is all ones unless --random-b is given, 10 warm-up launches
// are followed by 10 timed launches between two events, and ms is their mean
// (include/tf32_comp.hpp:33-46,79,112).  The imbalance test
// (include/utils.hpp:75-85) picks the balanced schedule (heavy row windows
// split into pieces) or the plain one.  Options:
//   --form fp32|split_bf16|bf16   precision form (default fp32)
//   --random-b SEED               B uniform in [-1,1) instead of ones
//   --warmup N / --exe N          launch counts (default 10 / 10)
//   --check                       random B (seed 1 unless --random-b) and a
//                                 host FP64 CSR product: prints the
//                                 componentwise max-rel error (against
//                                 |A||B|) and the Frobenius error, exit 6
//                                 above the form's tolerance (the reference
//                                 keeps B = 1 and never checks C)
//   --layout reference|auto       the graph the launches use: the reference's
//                                 METCF (default) or the library's automatic
//                                 layout (16-row windows and rows reordered by
//                                 length when rows average >= 128 nonzeros;
//                                 16-row windows alone for large banded
//                                 graphs); C is in input row order either way
// The CSR is canonicalised on load (rows sorted by column, duplicates
// summed), which the METCF tile layout requires.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "accspmm/mmio.hpp"

static std::string match_filename(const std::string& s) {
  size_t slash = s.rfind('/');
  return slash == std::string::npos ? s : s.substr(slash + 1);
}

#define HIP_OK(x)                                                         \
  do {                                                                    \
    hipError_t e_ = (x);                                                  \
    if (e_ != hipSuccess) {                                               \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                   \
      return 5;                                                           \
    }                                                                     \
  } while (0)

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("Usage: ./mma <matrix.mtx> <feature_dim> [--form fp32|split_bf16|bf16]"
                " [--random-b SEED] [--warmup N] [--exe N] [--check]"
                " [--layout reference|auto]\n");
    return ERROR_ARGS;
  }
  const char* filename = argv[1];
  const int feature_dim = std::atoi(argv[2]);
  if (feature_dim <= 0) {
    std::printf("feature_dim must be a positive integer\n");
    return ERROR_ARGS;
  }
  accspmm_form form = ACCSPMM_FORM_FP32;
  long long bseed = -1;
  int warmup = 10, exe = 10;
  bool check = false, auto_layout = false;
  for (int i = 3; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--form" && i + 1 < argc) {
      std::string f = argv[++i];
      if (f == "fp32") form = ACCSPMM_FORM_FP32;
      else if (f == "split_bf16") form = ACCSPMM_FORM_SPLIT_BF16;
      else if (f == "bf16") form = ACCSPMM_FORM_BF16;
      else { std::printf("unknown form %s\n", f.c_str()); return ERROR_ARGS; }
    } else if (a == "--random-b" && i + 1 < argc) {
      bseed = std::atoll(argv[++i]);
    } else if (a == "--warmup" && i + 1 < argc) {
      warmup = std::atoi(argv[++i]);
    } else if (a == "--exe" && i + 1 < argc) {
      exe = std::max(1, std::atoi(argv[++i]));
    } else if (a == "--check") {
      check = true;
    } else if (a == "--layout" && i + 1 < argc) {
      std::string l = argv[++i];
      if (l == "reference") auto_layout = false;
      else if (l == "auto") auto_layout = true;
      else { std::printf("unknown layout %s\n", l.c_str()); return ERROR_ARGS; }
    } else {
      std::printf("unknown option %s\n", a.c_str());
      return ERROR_ARGS;
    }
  }
  const std::string mtx_name = match_filename(filename);

  COO<MAT_VAL_TYPE> coo;
  if (read_from_mtx<MAT_VAL_TYPE>(filename, &coo, /*canonicalize=*/true) != SUCCESS)
    return ERROR_READ_MTX;
  CSR<MAT_VAL_TYPE> csr(&coo);

  METCFBit<MAT_VAL_TYPE> metcf_bit;
  try {
    metcf_bit.convertFromCSR(csr);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return ERROR_READ_MTX;
  }
  const vint numEdges = csr.nnz;
  const bool load_balance =
      balanceStrategy(metcf_bit, (vint)metcf_bit.tcOffset.size() - 1,
                      (vint)metcf_bit.rowWindowOffset.size() - 1);

  if (check && bseed < 0) bseed = 1;
  const size_t nB = (size_t)csr.cols * feature_dim, nC = (size_t)csr.rows * feature_dim;
  std::vector<float> hB(nB, 1.0f);
  if (bseed >= 0) accspmm_gen_dense((uint64_t)bseed, nB, hB.data());
  float *dB = nullptr, *dC = nullptr;
  HIP_OK(hipMalloc((void**)&dB, nB * sizeof(float)));
  HIP_OK(hipMalloc((void**)&dC, nC * sizeof(float)));
  HIP_OK(hipMemcpy(dB, hB.data(), nB * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(dC, 0, nC * sizeof(float)));

  // the launches run on the reference's METCF, or on a second build in the
  // library's automatic layout
  accspmm_metcf* graph = metcf_bit.handle;
  accspmm_metcf* own = nullptr;
  if (auto_layout) {
    accspmm_build_opts o;
    accspmm_build_opts_default(&o);
    o.row_window = 0;
    o.reorder_rows = 2;
    o.build_adpbme = 0;
    const accspmm_csr_view v = csr.view();
    const accspmm_status bst = accspmm_metcf_build_host(&v, &o, &own);
    if (bst != ACCSPMM_OK) {
      std::fprintf(stderr, "build: %s\n", accspmm_status_string(bst));
      return 5;
    }
    graph = own;
  }
  accspmm_tile tile;
  accspmm_tile_select(graph, feature_dim, form, &tile);
  // the reference's unbalanced path does not split heavy windows
  if (!load_balance && !auto_layout) tile.split_tiles = 0;
  std::cout << (load_balance ? "doing AdpBalance pipelining ..." : "doing pipelining mma ...")
            << std::endl;

  auto run = [&]() {
    return accspmm_spmm(graph, dB, feature_dim, dC, feature_dim, feature_dim,
                        form, &tile, nullptr);
  };
  for (int i = 0; i < warmup; ++i) {
    accspmm_status st = run();
    if (st != ACCSPMM_OK) {
      std::fprintf(stderr, "spmm: %s\n", accspmm_status_string(st));
      return 5;
    }
  }
  HIP_OK(hipDeviceSynchronize());
  hipEvent_t start, stop;
  HIP_OK(hipEventCreate(&start));
  HIP_OK(hipEventCreate(&stop));
  HIP_OK(hipEventRecord(start, nullptr));
  for (int i = 0; i < exe; ++i) run();
  HIP_OK(hipEventRecord(stop, nullptr));
  HIP_OK(hipEventSynchronize(stop));
  float total_ms = 0;
  HIP_OK(hipEventElapsedTime(&total_ms, start, stop));
  const double elapsed_time = total_ms / exe;
  const double spmm_throughput =
      ((double)numEdges * feature_dim * 2.0 * 1000.) / (elapsed_time * 1e9);

  std::ofstream outFile("result.csv", std::ios::app);
  if (!outFile) std::cerr << "Error Opening result.csv" << std::endl;
  outFile << mtx_name << "," << feature_dim << "," << elapsed_time << "," << spmm_throughput
          << "\n";
  outFile.close();
  int rc = 0;
  if (check) {
    std::vector<float> hC(nC);
    HIP_OK(hipMemcpy(hC.data(), dC, nC * sizeof(float), hipMemcpyDeviceToHost));
    // FP64 product and |A||B| row by row; componentwise error against
    // |A||B| (an entry with |A||B| = 0 must be exactly 0)
    std::vector<double> ref(feature_dim), mag(feature_dim);
    double max_rel = 0, err2 = 0, ref2 = 0;
    for (vint r = 0; r < csr.rows; ++r) {
      std::fill(ref.begin(), ref.end(), 0.0);
      std::fill(mag.begin(), mag.end(), 0.0);
      for (vint e = csr.row_ptr[r]; e < csr.row_ptr[r + 1]; ++e) {
        const double a = csr.data[e];
        const float* b = hB.data() + (size_t)csr.col_idx[e] * feature_dim;
        for (int k = 0; k < feature_dim; ++k) {
          ref[k] += a * b[k];
          mag[k] += std::fabs(a * b[k]);
        }
      }
      const float* c = hC.data() + (size_t)r * feature_dim;
      for (int k = 0; k < feature_dim; ++k) {
        const double d = std::fabs((double)c[k] - ref[k]);
        const double rel = mag[k] > 0 ? d / mag[k] : (d > 0 ? INFINITY : 0.0);
        max_rel = std::max(max_rel, rel);
        err2 += d * d;
        ref2 += ref[k] * ref[k];
      }
    }
    const double fro = ref2 > 0 ? std::sqrt(err2 / ref2) : std::sqrt(err2);
    const double tol = form == ACCSPMM_FORM_BF16 ? 1e-2 : 1e-5;
    std::printf("check: max_rel=%.3e fro_rel=%.3e tol=%.1e %s\n", max_rel, fro, tol,
                max_rel <= tol ? "ok" : "FAILED");
    if (!(max_rel <= tol)) rc = 6;
  }
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (own) (void)accspmm_metcf_destroy(own);
  return rc;
}
