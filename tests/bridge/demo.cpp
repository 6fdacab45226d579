// The reference driver's sequence (main.cpp:59-118) written against the
// reference's C++ names, served by include/partc_bridge.hpp:
//   read A (.tns), A.populate_meta_info(meta), A_copy free-first, B := A,
//   HT_B = buildHashTable(B), HT_A = buildHashTable(A_copy),
//   O = tc_fksCuckoo(HT_A, HT_B, A_copy, B)
// plus tc_fksCuckoo<Hash_Table_SPA> and matricize_spGEMM on the same input.
// Writes the three results as .tns files: <out>.tc.tns, <out>.spa.tns,
// <out>.mat.tns.  Built with -DTENSOR_HEADER=<the caller's tensor.hpp>.
#include <cstdlib>
#include <cstring>
#include <string>

#include TENSOR_HEADER
#include "partc_bridge.hpp"

static Tensor read_tns(const char* path) {
  ptc_host_tensor h{};
  if (ptc_read_tns(path, &h) != PTC_OK) {
    std::fprintf(stderr, "cannot read %s\n", path);
    std::exit(1);
  }
  Tensor T;
  T.d = h.d;
  T.N = (unsigned)h.N;
  T.hedges_array = (unsigned*)std::malloc(sizeof(unsigned) * h.N * h.d);
  T.val_array = (double*)std::malloc(sizeof(double) * h.N);
  T.dimensions_array = (unsigned*)std::malloc(sizeof(unsigned) * h.d);
  std::memcpy(T.hedges_array, h.hedges, sizeof(unsigned) * h.N * h.d);
  std::memcpy(T.val_array, h.vals, sizeof(double) * h.N);
  for (unsigned j = 0; j < h.d; ++j) T.dimensions_array[j] = h.dims[j];
  ptc_host_tensor_free(&h);
  return T;
}

static void write(const Tensor& O, const std::string& path) {
  if (ptc_write_tns(path.c_str(), O.d, O.N, O.hedges_array, O.val_array) != PTC_OK) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s tensor.tns tensor.meta out_prefix\n", argv[0]);
    return 2;
  }
  Tensor A = read_tns(argv[1]);
  A.populate_meta_info(argv[2]);
  // A_copy: the same nonzeros keyed on the free modes (main.cpp:68-109)
  Tensor A_copy = A;
  const unsigned c = A.num_key_dimensions;
  A_copy.num_key_dimensions = A.d - c;
  A_copy.dim_tc = (unsigned*)std::malloc(sizeof(unsigned) * A.d);
  for (unsigned i = c; i < A.d; ++i) A_copy.dim_tc[i - c] = A.dim_tc[i];
  for (unsigned i = 0; i < c; ++i) A_copy.dim_tc[A.d - c + i] = A.dim_tc[i];
  Tensor B = A;  // get_sparse_tensor: the same tensor, contracted-first
  Hash_Table* HT_B = buildHashTable(B);
  Hash_Table* HT_A = buildHashTable(A_copy);
  Tensor O = tc_fksCuckoo(HT_A, HT_B, A_copy, B);
  Tensor S = tc_fksCuckoo<Hash_Table_SPA>(HT_A, HT_B, A_copy, B);
  Tensor M = matricize_spGEMM(A, B);
  const std::string out = argv[3];
  write(O, out + ".tc.tns");
  write(S, out + ".spa.tns");
  write(M, out + ".mat.tns");
  std::printf("nnz %u %u %u\n", O.N, S.N, M.N);
  freeHashTable(HT_A);
  freeHashTable(HT_B);
  partc::shutdown();
  return 0;
}
