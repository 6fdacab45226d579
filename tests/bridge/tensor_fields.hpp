// A tensor type with the field layout of the reference's struct Tensor
// (tensor.hpp:19-26), for building the bridge demo where the reference tree
// is absent.  populate_meta_info reads a .meta file with ptc_meta_parse.
#ifndef PTC_TEST_TENSOR_FIELDS_HPP
#define PTC_TEST_TENSOR_FIELDS_HPP
#include <cstdio>
#include <cstdlib>
#include <string>

#include "partc.h"

struct Tensor {
  unsigned d = 0;
  unsigned N = 0;
  unsigned num_key_dimensions = 0;
  unsigned* hedges_array = nullptr;
  unsigned* dimensions_array = nullptr;
  double* val_array = nullptr;
  unsigned* dim_tc = nullptr;

  void populate_meta_info(std::string filename_meta) {
    uint32_t c = 0;
    dim_tc = (unsigned*)std::malloc(sizeof(unsigned) * d);
    if (ptc_meta_parse(filename_meta.c_str(), d, &c, dim_tc) != PTC_OK) {
      std::fprintf(stderr, "Could not read \"%s\"\n", filename_meta.c_str());
      std::exit(1);
    }
    num_key_dimensions = c;
  }
};
#endif
