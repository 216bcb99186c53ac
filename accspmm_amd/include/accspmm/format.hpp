// format.hpp — source-compatible drop-in for AccSpMM's format API
// (include/format.hpp and include/common.hpp of the reference): the same
// class, member and method names, with METCF / AdpBME construction running
// on the MI355X through libaccspmm's C ABI.
//
//   COO<T>, CSR<T>(const COO<T>*)           include/format.hpp:33-169
//   METCFBit<T>::convertFromCSR(csr)        include/format.hpp:202-274
//   AdpBME<T>::CSR2AdpBME(csr, 3, 1, 2)     include/format.hpp:327-429
//   balanceStrategy(metcf, numTc, numWin)   include/utils.hpp:75-85
//
// The public std::vector members hold the byte-exact export of the GPU build.
// Differences: DataType must be float (the library stores FP32 values);
// build errors throw std::runtime_error instead of printing and continuing;
// the CSR must be canonical (rows strictly increasing in column) because the
// tile values are placed by bit rank, which equals the reference's CSR-order
// placement only for such rows.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/accspmm.h"

typedef uint32_t vint;
#ifndef ROW_WINDOW
#define ROW_WINDOW 8
#endif
#define COL_WINDOW 8
#define TARGET_NUM 16
#define TCLOCAL_TYPE uint64_t
#define IMBALANCE_THRESHOLD 8
#define MAT_VAL_TYPE float
#define MAT_PTR_TYPE uint32_t

namespace accspmm_detail {
inline void check(accspmm_status s, const char* what) {
  if (s != ACCSPMM_OK)
    throw std::runtime_error(std::string(what) + ": " + accspmm_status_string(s));
}
}  // namespace accspmm_detail

template <typename DataType>
class SparseMatrixFormat {
 public:
  vint rows = 0, cols = 0, nnz = 0;
  virtual void show() const = 0;
  virtual long long printSize(bool print = true) const = 0;
  virtual ~SparseMatrixFormat() = default;
};

template <typename DataType>
class COO : public SparseMatrixFormat<DataType> {
 public:
  using SparseMatrixFormat<DataType>::rows;
  using SparseMatrixFormat<DataType>::cols;
  using SparseMatrixFormat<DataType>::nnz;
  vint* row = nullptr;
  vint* col = nullptr;
  DataType* data = nullptr;
  COO() = default;
  COO(const COO&) = delete;
  COO& operator=(const COO&) = delete;
  ~COO() override {
    std::free(row);
    std::free(col);
    std::free(data);
  }
  void show() const override {
    for (vint i = 0; i < nnz; ++i)
      std::cout << "row = " << row[i] << ", col = " << col[i]
                << ", data = " << data[i] << "\n";
  }
  long long printSize(bool print = true) const override {
    long long size = (long long)nnz * (2 * sizeof(vint) + sizeof(DataType));
    if (print)
      std::cout << "\n========== size of COO =========\nTotal size: " << size
                << " bytes" << std::endl;
    return size;
  }
};

template <typename DataType>
class CSR : public SparseMatrixFormat<DataType> {
 public:
  using SparseMatrixFormat<DataType>::rows;
  using SparseMatrixFormat<DataType>::cols;
  using SparseMatrixFormat<DataType>::nnz;
  std::vector<vint> row_ptr;
  std::vector<vint> col_idx;
  std::vector<DataType> data;
  CSR() = default;
  // stable counting sort by row (reference include/format.hpp:110-142)
  explicit CSR(const COO<DataType>* coo) {
    rows = coo->rows;
    cols = coo->cols;
    nnz = coo->nnz;
    row_ptr.assign(rows + 1, 0);
    col_idx.resize(nnz);
    data.resize(nnz);
    for (vint i = 0; i < nnz; ++i) row_ptr[coo->row[i] + 1]++;
    for (vint i = 0; i < rows; ++i) row_ptr[i + 1] += row_ptr[i];
    std::vector<vint> fill(row_ptr.begin(), row_ptr.end() - 1);
    for (vint i = 0; i < nnz; ++i) {
      vint d = fill[coo->row[i]]++;
      col_idx[d] = coo->col[i];
      data[d] = coo->data[i];
    }
  }
  void show() const override {
    std::cout << "Matrix in CSR format, Row: " << rows << ", Col: " << cols
              << ", NNZ: " << nnz << "\n";
    for (vint i = 0; i + 1 < row_ptr.size(); ++i) {
      std::cout << "\nrow = " << i << "\n";
      for (vint j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
        std::cout << "col = " << col_idx[j] << ": data = " << data[j] << "\n";
    }
  }
  long long printSize(bool print = true) const override {
    long long size = (long long)(row_ptr.size() + col_idx.size()) * sizeof(vint) +
                     (long long)data.size() * sizeof(DataType);
    if (print)
      std::cout << "\n========== size of CSR =========\nTotal size: " << size
                << " bytes" << std::endl;
    return size;
  }
  accspmm_csr_view view() const {
    return accspmm_csr_view{rows, cols, nnz, row_ptr.data(), col_idx.data(),
                            reinterpret_cast<const float*>(data.data())};
  }
};

template <typename DataType>
class TensorCoreFormat : public SparseMatrixFormat<DataType> {
  static_assert(std::is_same<DataType, float>::value,
                "the MI355X METCF build stores FP32 values");

 public:
  std::vector<vint> sparseA2B;
  std::vector<DataType> data;
  accspmm_metcf* handle = nullptr;  // device-resident graph for accspmm_spmm
  TensorCoreFormat() = default;
  TensorCoreFormat(const TensorCoreFormat&) = delete;
  TensorCoreFormat& operator=(const TensorCoreFormat&) = delete;
  ~TensorCoreFormat() override {
    if (handle) accspmm_metcf_destroy(handle);
  }

 protected:
  void build(const CSR<DataType>& csr, const accspmm_build_opts& o) {
    if (handle) accspmm_metcf_destroy(handle);
    handle = nullptr;
    accspmm_csr_view v = csr.view();
    accspmm_detail::check(accspmm_metcf_build_host(&v, &o, &handle), "METCF build");
    accspmm_detail::check(accspmm_metcf_get_info(handle, &info), "METCF info");
    this->rows = info.rows;
    this->cols = info.cols;
    this->nnz = info.nnz;
  }
  accspmm_metcf_info info{};
};

template <typename DataType>
class METCFBit : public TensorCoreFormat<DataType> {
 public:
  using TensorCoreFormat<DataType>::sparseA2B;
  using TensorCoreFormat<DataType>::data;
  std::vector<vint> rowWindowOffset;
  std::vector<vint> tcOffset;
  std::vector<TCLOCAL_TYPE> tcLocalBit;
  METCFBit() = default;

  void convertFromCSR(const CSR<DataType>& csr) {
    accspmm_build_opts o;
    accspmm_build_opts_default(&o);
    o.row_window = ROW_WINDOW;
    o.build_adpbme = 0;
    this->build(csr, o);
    const auto& in = this->info;
    rowWindowOffset.resize(in.num_windows + 1);
    tcOffset.resize(in.num_tiles + 1);
    tcLocalBit.resize((size_t)in.num_tiles * (in.row_window / 8));
    sparseA2B.resize((size_t)in.num_tiles * COL_WINDOW);
    data.resize(in.nnz);
    accspmm_detail::check(
        accspmm_metcf_export(this->handle, rowWindowOffset.data(), tcOffset.data(),
                             tcLocalBit.data(), sparseA2B.data(), data.data(),
                             nullptr, nullptr),
        "METCF export");
  }
  bool balanced() const { return this->info.balanced != 0; }
  vint ibd() const { return this->info.ibd; }

  void show() const override {
    std::cout << "\n========== METCFBit =========\nrowWindowOffset:\n";
    for (auto v : rowWindowOffset) std::cout << v << " ";
    std::cout << "\ntcLocalBit:\n";
    for (auto v : tcLocalBit) std::cout << v << "\n";
    std::cout << "\nsparseA2B:\n";
    for (auto v : sparseA2B) std::cout << v << " ";
    std::cout << "\ndata:\n";
    for (auto v : data) std::cout << v << " ";
    std::cout << std::endl;
  }
  long long printSize(bool print = true) const override {
    long long s1 = rowWindowOffset.size() * 4, s2 = tcOffset.size() * 4,
              s3 = tcLocalBit.size() * 8, s4 = sparseA2B.size() * 4;
    long long total = s1 + s2 + s3 + s4;
    if (print)
      std::cout << "\n========== size of ME-CTF-bit =========\ntotal size(without data): "
                << total << " bytes\nrowWindowOffset: " << s1 << " bytes\ntcOffset: "
                << s2 << " bytes\ntcLocalBit: " << s3 << " bytes\nsparseA2B: " << s4
                << " bytes\n"
                << std::endl;
    return total;
  }
};

template <typename DataType>
class AdpBME : public TensorCoreFormat<DataType> {
 public:
  using TensorCoreFormat<DataType>::sparseA2B;
  using TensorCoreFormat<DataType>::data;
  std::vector<vint> groupOffset;
  std::vector<vint> tcOffset;
  std::vector<vint> rowIndices;
  std::vector<TCLOCAL_TYPE> tcLocalBit;
  AdpBME() = default;

  void CSR2AdpBME(const CSR<DataType>& csr, vint t_load, vint t_comp, vint t_write,
                  vint target_num = TARGET_NUM) {
    accspmm_build_opts o;
    accspmm_build_opts_default(&o);
    o.row_window = ROW_WINDOW;
    o.t_load = t_load;
    o.t_comp = t_comp;
    o.t_write = t_write;
    o.target_num = target_num;
    o.build_adpbme = 1;
    this->build(csr, o);
    const auto& in = this->info;
    groupOffset.resize(in.num_groups + 1);
    tcOffset.resize(in.num_tiles + 1);
    rowIndices.resize(in.num_tiles);
    tcLocalBit.resize((size_t)in.num_tiles * (in.row_window / 8));
    sparseA2B.resize((size_t)in.num_tiles * COL_WINDOW);
    data.resize(in.nnz);
    accspmm_detail::check(
        accspmm_metcf_export(this->handle, nullptr, tcOffset.data(), tcLocalBit.data(),
                             sparseA2B.data(), data.data(), groupOffset.data(),
                             rowIndices.data()),
        "AdpBME export");
  }

  void show() const override {
    std::cout << "\n========== AdpBME =========:\ngroupOffset:\n";
    for (auto v : groupOffset) std::cout << v << " ";
    std::cout << "\ntcOffset:\n";
    for (auto v : tcOffset) std::cout << v << " ";
    std::cout << "\nrowIndices:\n";
    for (auto v : rowIndices) std::cout << v << " ";
    std::cout << std::endl;
  }
  long long printSize(bool print = true) const override {
    long long total = (long long)(groupOffset.size() + tcOffset.size() +
                                  rowIndices.size() + sparseA2B.size()) * 4 +
                      (long long)tcLocalBit.size() * 8;
    if (print)
      std::cout << "\n========== size of AdpBME (ours) =========\ntotal size(without data array): "
                << total << " bytes" << std::endl;
    return total;
  }
};

// include/utils.hpp:75-85, computed on the GPU when the METCF was built.
inline bool balanceStrategy(METCFBit<MAT_VAL_TYPE>& metcf, vint numTcBlocks,
                            vint numBlocks) {
  if (numTcBlocks != metcf.tcOffset.size() - 1 ||
      numBlocks != metcf.rowWindowOffset.size() - 1)
    throw std::runtime_error("balanceStrategy: arguments do not match the METCF");
  return metcf.balanced();
}
