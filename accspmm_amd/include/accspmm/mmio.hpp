// mmio.hpp — drop-in for the reference's read_from_mtx
// (include/mmio.hpp:171-210): Matrix Market -> COO<T>, through libaccspmm.
//
// Return codes follow the reference: SUCCESS (0), MM_LOADING_FILE_ERROR (-1)
// when the file cannot be opened, MM_READ_BANNER_ERROR (-2) for a bad banner,
// ERROR_READ_MTX (3) for a bad size line or entry.  The COO comes back in row
// order (the order of the CSR the reference builds from it); symmetric and
// Hermitian files are expanded as in the reference.  The COO's row, col and
// data pointers are overwritten with new malloc'd arrays without reading
// them, so an unconstructed COO from malloc (the reference driver's use) is
// valid input.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "format.hpp"

#define SUCCESS 0
#define MM_LOADING_FILE_ERROR -1
#define MM_READ_BANNER_ERROR -2
#define ERROR_READ_MTX 3
#define ERROR_ARGS 4

template <class DataType>
int read_from_mtx(const char* filename, COO<DataType>* coo, bool canonicalize = false) {
  accspmm_host_csr h{};
  int code = 0;
  accspmm_status st = accspmm_read_mtx(filename, canonicalize ? 1 : 0, &h, &code);
  if (st != ACCSPMM_OK) {
    if (code == -1) {
      std::perror("Error opening file.");
      return MM_LOADING_FILE_ERROR;
    }
    if (code == -2) {
      std::printf("Couldn't process Matrix Market banner. \n");
      return MM_READ_BANNER_ERROR;
    }
    return ERROR_READ_MTX;
  }
  std::cout << "m_tmp: " << h.rows << ", n_tmp: " << h.cols << ", nnz_tmp: " << h.nnz
            << std::endl;
  // like the reference's handle_coo (include/mmio.hpp:136-138,160-162) the
  // array pointers are overwritten, never freed: the reference's driver
  // passes a COO from malloc whose pointers are uninitialised
  // (src/mma_tf32.cu:198-199).  Arrays a COO already holds are left to the
  // caller, as in the reference.
  coo->rows = h.rows;
  coo->cols = h.cols;
  coo->nnz = h.nnz;
  coo->row = (vint*)std::malloc(sizeof(vint) * (h.nnz ? h.nnz : 1));
  coo->col = (vint*)std::malloc(sizeof(vint) * (h.nnz ? h.nnz : 1));
  coo->data = (DataType*)std::malloc(sizeof(DataType) * (h.nnz ? h.nnz : 1));
  for (vint r = 0; r < h.rows; ++r)
    for (vint e = h.row_ptr[r]; e < h.row_ptr[r + 1]; ++e) coo->row[e] = r;
  std::memcpy(coo->col, h.col_idx, sizeof(vint) * h.nnz);
  for (vint e = 0; e < h.nnz; ++e) coo->data[e] = (DataType)h.val[e];
  accspmm_host_csr_free(&h);
  return SUCCESS;
}
