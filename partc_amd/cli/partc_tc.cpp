// partc-tc: command-line driver with the reference driver's arguments
// (main.cpp:41-66): a tensor (or matrix) file, its .meta file, a flag for
// Matrix Market input and a number of rounds.  Each round contracts the
// tensor with itself over the .meta modes, as main.cpp:68-118 does (A_copy
// free-first, B = A, HT_B then HT_A, tc_fksCuckoo), on the GPU through the C
// ABI, and prints the phase times and the output size.
//
//   partc-tc tensorA metaA isMtx rounds [two_pass|upper_bound|learned] [out.tns]
//
// The learned method needs a trained model file: the path in PTC_MODEL, or
// rowsize_mlp.bin in the models/ directory beside this program.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "partc.h"

static int fail(ptc_ctx* ctx, int rc, const char* what) {
  fprintf(stderr, "%s failed (%d)%s%s\n", what, rc, ctx ? ": " : "", ctx ? ptc_last_error(ctx) : "");
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr,
            "usage: %s tensorA metaA isMtx rounds [two_pass|upper_bound|learned] [out.tns]\n"
            "  tensorA   FROSTT .tns file (isMtx = 0) or Matrix Market file (isMtx = 1)\n"
            "  metaA     contraction spec: two text lines, c, then the mode ids (contracted first)\n"
            "  rounds    number of times to run the contraction\n",
            argv[0]);
    return 1;
  }
  const char* path = argv[1];
  const char* meta = argv[2];
  const int is_mtx = atoi(argv[3]);
  const int rounds = atoi(argv[4]);
  const std::string method = argc > 5 ? argv[5] : "two_pass";
  const char* out_path = argc > 6 ? argv[6] : nullptr;

  ptc_host_tensor T;
  int rc = is_mtx ? ptc_read_mtx(path, &T) : ptc_read_tns(path, &T);
  if (rc) return fail(nullptr, rc, is_mtx ? "reading the matrix" : "reading the tensor");
  printf("input: %s, %u modes, %llu nonzeros, dims", path, T.d, (unsigned long long)T.N);
  for (uint32_t j = 0; j < T.d; ++j) printf(" %u", T.dims[j]);
  printf("\n");
  uint32_t c = 0, perm[PTC_MAX_MODES];
  if ((rc = ptc_meta_parse(meta, T.d, &c, perm))) return fail(nullptr, rc, "reading the .meta file");

  ptc_ctx* ctx = nullptr;
  if ((rc = ptc_create(&ctx, 0))) return fail(nullptr, rc, "opening the GPU");
  // the per-phase times printed below (main.cpp prints its phases too)
  if ((rc = ptc_set_timing(ctx, 1))) return fail(ctx, rc, "phase timing");
  ptc_options opt;
  ptc_options_default(&opt);
  if (method == "upper_bound") {
    opt.method = PTC_METHOD_UPPER_BOUND;
  } else if (method == "learned") {
    opt.method = PTC_METHOD_LEARNED;
    std::string model;
    if (const char* m = getenv("PTC_MODEL")) {
      model = m;
    } else {
      std::string self(argv[0]);
      const size_t slash = self.find_last_of('/');
      model = (slash == std::string::npos ? std::string(".") : self.substr(0, slash)) + "/models/rowsize_mlp.bin";
    }
    ptc_mlp* mlp = nullptr;
    if ((rc = ptc_mlp_load(ctx, model.c_str(), &mlp))) return fail(ctx, rc, "loading the row-size model");
    ptc_learned_set_default(mlp);
  } else if (method != "two_pass") {
    fprintf(stderr, "unknown method %s\n", method.c_str());
    return 1;
  }
  // the tensor as main.cpp holds it after populate_meta_info: contracted
  // modes first, num_key_dimensions = c; B := A (get_sparse_tensor)
  ptc_tensor_view v{T.d, c, T.N, T.hedges, T.dims, T.vals, perm};
  ptc_dtensor* A = nullptr;
  if ((rc = ptc_dtensor_upload(ctx, &v, &A))) return fail(ctx, rc, "upload");
  for (int r = 0; r < rounds; ++r) {
    ptc_dtensor* O = nullptr;
    ptc_stats st;
    rc = ptc_contract(ctx, A, A, &opt, &O, &st);
    if (rc == PTC_ERR_EMPTY) {
      printf("tensor contraction not possible!!\n");
      return 1;
    }
    if (rc) return fail(ctx, rc, "contraction");
    printf("round %d (%s): HT_B %.3f ms, HT_A %.3f ms, preprocessing %.3f ms, symbolic %.3f ms, numeric %.3f ms, "
           "contraction total %.3f ms\n",
           r, method.c_str(), st.ms_build_b, st.ms_build_a, st.ms_preprocess, st.ms_symbolic, st.ms_numeric,
           st.ms_total);
    printf("multiply-adds %llu, nonzeros in the output tensor %llu, fallback rows %llu\n",
           (unsigned long long)st.flops, (unsigned long long)st.nnz_out, (unsigned long long)st.fallback_rows);
    if (out_path && r == rounds - 1) {
      uint32_t d = 0, nk = 0;
      uint64_t N = 0;
      ptc_dtensor_info(O, &d, &N, &nk);
      uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (N * d + 1));
      double* val = (double*)malloc(sizeof(double) * (N + 1));
      if (!h || !val) return fail(ctx, PTC_ERR_NOMEM, "output buffers");
      if ((rc = ptc_dtensor_download(ctx, O, h, val, nullptr))) return fail(ctx, rc, "download");
      if ((rc = ptc_write_tns(out_path, d, N, h, val))) return fail(ctx, rc, "writing the output");
      free(h);
      free(val);
    }
    ptc_dtensor_free(O);
  }
  ptc_dtensor_free(A);
  ptc_host_tensor_free(&T);
  ptc_destroy(ctx);
  return 0;
}
