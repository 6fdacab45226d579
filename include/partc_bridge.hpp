// partc_bridge.hpp -- the reference's C++ names over the C ABI (partc.h), so
// a caller of ssomesh/parTC's hash-based contraction keeps its code and
// swaps the includes: include the caller's own `struct Tensor` (tensor.hpp
// of the reference, or any struct with the same fields) first, then this
// header instead of fksCuckoo.hpp / fksCuckooTC.hpp / matricizedspGEMM*.hpp.
//
//   Hash_Table* buildHashTable(Tensor&)                 fksCuckoo.hpp:1044
//   Tensor tc_fksCuckoo(Hash_Table*, Hash_Table*,
//                       Tensor& A, Tensor& B)            fksCuckooTC.hpp:80
//   struct Hash_Table_SPA                               fksCuckooSPA.hpp:197
//       (here the accumulator policy: tc_fksCuckoo<Hash_Table_SPA>(...)
//        accumulates every output row in a hash table, as the reference
//        does; the default lets the engine choose dense LDS rows)
//   Tensor matricize_spGEMM(Tensor&, Tensor&)           matricizedspGEMM.hpp:172
//   Tensor matricize_sort_spGEMM(Tensor&, Tensor&)      matricizedspGEMM_sorting.hpp:194
//
// The functions are templates over the caller's tensor type, which needs the
// reference's fields (d, N, num_key_dimensions, hedges_array,
// dimensions_array, val_array, dim_tc; tensor.hpp:19-26) and a default
// constructor.  Preconditions and results are the reference's: A free-first
// and B contracted-first for tc_fksCuckoo, both contracted-first for the
// matricized products; the output has A's free modes then B's free modes,
// dimensions_array all zero, dim_tc the identity, and malloc'd arrays owned
// by the caller, exactly nnz long (the reference's are F long).  Errors
// print a message and exit(1), as the reference does; define
// PARTC_BRIDGE_THROW to get a partc::Error exception instead.
//
// One engine context per host thread is created on first use on the
// current HIP device.  Call partc::shutdown() before the process exits (it
// destroys the calling thread's context; HIP's own exit handlers may run
// before static destructors, so none is relied on).  Tables hold device
// memory until freeHashTable (the reference never frees them).
#ifndef PARTC_BRIDGE_HPP
#define PARTC_BRIDGE_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "partc.h"

namespace partc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& what, const ptc_ctx* ctx = nullptr) {
  std::string msg = what + " (partc status " + std::to_string(code) + ")";
  if (ctx && ptc_last_error(ctx)[0]) msg += ": " + std::string(ptc_last_error(ctx));
#ifdef PARTC_BRIDGE_THROW
  throw Error(code, msg);
#else
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(1);
#endif
}

inline ptc_ctx*& thread_ctx() {
  static thread_local ptc_ctx* ctx = nullptr;
  return ctx;
}

inline ptc_ctx* context() {
  ptc_ctx*& c = thread_ctx();
  if (!c) {
    const int rc = ptc_create(&c, -1);
    if (rc != PTC_OK) {
      c = nullptr;
      fail(rc, "partc: no engine context on the current HIP device");
    }
  }
  return c;
}

// Destroys the calling thread's context (free the tables first).
inline void shutdown() {
  ptc_ctx*& c = thread_ctx();
  if (c) ptc_destroy(c);
  c = nullptr;
}

// Accumulator policies for tc_fksCuckoo.
struct Auto_SPA {  // dense LDS rows where the column space fits, hash tables otherwise
  static constexpr int acc_mode = PTC_ACC_AUTO;
};
struct Hash_Table_SPA {  // a hash table per output row (fksCuckooSPA.hpp:197), in LDS or HBM
  static constexpr int acc_mode = PTC_ACC_HASH;
};

// The device copy of a host tensor, with the key split it had when copied.
template <class TensorT>
ptc_dtensor* upload(ptc_ctx* ctx, const TensorT& T) {
  ptc_tensor_view v{};
  v.d = T.d;
  v.num_key_dimensions = T.num_key_dimensions;
  v.N = T.N;
  v.hedges = T.hedges_array;
  v.dims = T.dimensions_array;
  v.vals = T.val_array;
  v.dim_tc = T.dim_tc;
  ptc_dtensor* dt = nullptr;
  const int rc = ptc_dtensor_upload(ctx, &v, &dt);
  if (rc != PTC_OK) fail(rc, "partc: tensor upload", ctx);
  return dt;
}

// A device result as a caller-owned host tensor (reference layout).
template <class TensorT>
TensorT download(ptc_ctx* ctx, ptc_dtensor* O) {
  uint32_t d = 0, nk = 0;
  uint64_t n = 0;
  ptc_dtensor_info(O, &d, &n, &nk);
  if (n > 0xFFFFFFFFull) {
    ptc_dtensor_free(O);
    fail(PTC_ERR_RANGE, "partc: the output has more nonzeros than a Tensor's 32-bit N holds");
  }
  TensorT R;
  R.d = d;
  R.N = (unsigned)n;
  R.num_key_dimensions = d;  // tensor.hpp:124
  R.hedges_array = (unsigned*)std::malloc(sizeof(unsigned) * (n * d ? n * d : 1));
  R.val_array = (double*)std::malloc(sizeof(double) * (n ? n : 1));
  R.dimensions_array = (unsigned*)std::calloc(d ? d : 1, sizeof(unsigned));  // zeros, tensor.hpp:121-122
  R.dim_tc = (unsigned*)std::malloc(sizeof(unsigned) * (d ? d : 1));
  for (unsigned i = 0; i < d; ++i) R.dim_tc[i] = i;  // identity, tensor.hpp:128-129
  if (!R.hedges_array || !R.val_array || !R.dimensions_array || !R.dim_tc) {
    ptc_dtensor_free(O);
    fail(PTC_ERR_NOMEM, "partc: host allocation of the output");
  }
  const int rc = ptc_dtensor_download(ctx, O, R.hedges_array, R.val_array, nullptr);
  ptc_dtensor_free(O);
  if (rc != PTC_OK) fail(rc, "partc: output download", ctx);
  return R;
}

}  // namespace partc

// The reference's Hash_Table (fksCuckoo.hpp), opaque here: a key table on the
// device over the tensor it was built on.
struct Hash_Table {
  ptc_ctx* ctx = nullptr;
  ptc_dtensor* tensor = nullptr;    // device copy of the source tensor
  ptc_hashtable* table = nullptr;
  const void* source = nullptr;     // the host tensor it was built on
};

using partc::Hash_Table_SPA;

// buildHashTable (fksCuckoo.hpp:1044-1060): group T's nonzeros by its key
// modes, dim_tc[0 .. num_key_dimensions).
template <class TensorT>
Hash_Table* buildHashTable(TensorT& T) {
  ptc_ctx* ctx = partc::context();
  Hash_Table* H = new Hash_Table();
  H->ctx = ctx;
  H->source = &T;
  H->tensor = partc::upload(ctx, T);
  const int rc = ptc_hashtable_build(ctx, H->tensor, &H->table);
  if (rc != PTC_OK) partc::fail(rc, "partc: buildHashTable", ctx);
  return H;
}

// Releases a table's device memory (no reference counterpart).
inline void freeHashTable(Hash_Table* H) {
  if (!H) return;
  ptc_hashtable_free(H->table);
  ptc_dtensor_free(H->tensor);
  delete H;
}

// tc_fksCuckoo (fksCuckooTC.hpp:80-371): O = A x B; HT_A built on A
// (free-first), HT_B on B (contracted-first).
template <class SPA = partc::Auto_SPA, class TensorT>
TensorT tc_fksCuckoo(Hash_Table* HT_A, Hash_Table* HT_B, TensorT& A, TensorT& B) {
  if (!HT_A || !HT_B || HT_A->source != &A || HT_B->source != &B)
    partc::fail(PTC_ERR_ARG, "partc: tc_fksCuckoo needs HT_A built on A and HT_B built on B");
  ptc_ctx* ctx = HT_A->ctx;
  ptc_options opt;
  ptc_options_default(&opt);
  opt.acc_mode = SPA::acc_mode;
  ptc_dtensor* O = nullptr;
  const int rc = ptc_tc_fksCuckoo(ctx, HT_A->table, HT_B->table, HT_A->tensor, HT_B->tensor, &opt, &O, nullptr);
  if (rc == PTC_ERR_EMPTY) partc::fail(rc, "tensor contraction not possible!!", ctx);  // tensor.hpp:116-119
  if (rc != PTC_OK) partc::fail(rc, "partc: tc_fksCuckoo", ctx);
  return partc::download<TensorT>(ctx, O);
}

namespace partc {
template <class TensorT, class Fn>
TensorT matricized(TensorT& A, TensorT& B, Fn fn, const char* name) {
  ptc_ctx* ctx = context();
  ptc_dtensor* dA = upload(ctx, A);
  ptc_dtensor* dB = upload(ctx, B);
  ptc_dtensor* O = nullptr;
  const int rc = fn(ctx, dA, dB, &O, nullptr);
  ptc_dtensor_free(dA);
  ptc_dtensor_free(dB);
  if (rc == PTC_ERR_EMPTY) fail(rc, "tensor contraction not possible!!", ctx);
  if (rc != PTC_OK) fail(rc, std::string("partc: ") + name, ctx);
  return download<TensorT>(ctx, O);
}
}  // namespace partc

// matricize_spGEMM (matricizedspGEMM.hpp:172): both tensors contracted-first.
template <class TensorT>
TensorT matricize_spGEMM(TensorT& A, TensorT& B) {
  return partc::matricized(A, B, ptc_matricize_spGEMM, "matricize_spGEMM");
}

// matricize_sort_spGEMM (matricizedspGEMM_sorting.hpp:194, which exits before
// computing the product, :754): the same product, grouped by sorting.
template <class TensorT>
TensorT matricize_sort_spGEMM(TensorT& A, TensorT& B) {
  return partc::matricized(A, B, ptc_matricize_sort_spGEMM, "matricize_sort_spGEMM");
}

#endif  // PARTC_BRIDGE_HPP
