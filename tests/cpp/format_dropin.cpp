// Reference-style use of the format API (include/format.hpp, include/mmio.hpp,
// include/utils.hpp of AccSpMM), compiled unchanged against the drop-in
// headers.  Prints the METCF and AdpBME arrays as JSON for the test harness.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "accspmm/mmio.hpp"

template <class V>
static void dump(const char* name, const V& v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i)
    std::printf(i ? ", %llu" : "%llu", (unsigned long long)v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char** argv) {
  if (argc < 2) return ERROR_ARGS;
  // the reference driver's sequence (src/mma_tf32.cu:198-213): a COO from
  // malloc, never constructed, filled by read_from_mtx; CSR from it; freed
  // without destruction after the CSR is built
  COO<MAT_VAL_TYPE>* coo = (COO<MAT_VAL_TYPE>*)malloc(sizeof(COO<MAT_VAL_TYPE>));
  std::memset((void*)coo, 0xA5, sizeof(COO<MAT_VAL_TYPE>));  // garbage pointers
  if (read_from_mtx<MAT_VAL_TYPE>(argv[1], coo) != SUCCESS) return ERROR_READ_MTX;
  CSR<MAT_VAL_TYPE> csr(coo);
  std::free(coo->row);
  std::free(coo->col);
  std::free(coo->data);
  std::free(coo);
  METCFBit<MAT_VAL_TYPE> metcf_bit;
  metcf_bit.convertFromCSR(csr);
  bool load_balance = balanceStrategy(metcf_bit, metcf_bit.tcOffset.size() - 1,
                                      metcf_bit.rowWindowOffset.size() - 1);
  AdpBME<MAT_VAL_TYPE> adpbme;
  adpbme.CSR2AdpBME(csr, 3, 1, 2);
  std::vector<unsigned long long> data_bits;
  for (float f : metcf_bit.data) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    data_bits.push_back(u);
  }
  std::printf("{\n");
  dump("rowWindowOffset", metcf_bit.rowWindowOffset);
  dump("tcOffset", metcf_bit.tcOffset);
  dump("tcLocalBit", metcf_bit.tcLocalBit);
  dump("sparseA2B", metcf_bit.sparseA2B);
  dump("data_bits", data_bits);
  dump("groupOffset", adpbme.groupOffset);
  dump("rowIndices", adpbme.rowIndices);
  std::printf("\"balanced\": %d\n}\n", load_balance ? 1 : 0);
  return 0;
}
