/* oracle.h -- CPU restatement of the reference contraction (test
 * infrastructure only: the checker and the CPU baseline; the product never
 * links it). */
#ifndef PTC_ORACLE_H
#define PTC_ORACLE_H
#include <stdint.h>

#define ORC_MAX_MODES 8

enum {
  ORC_OK = 0,
  ORC_ERR_ARG = -1,
  ORC_ERR_NOMEM = -3,
  ORC_ERR_EMPTY = -4,   /* "tensor contraction not possible!!" (tensor.hpp:116-119) */
  ORC_ERR_RANGE = -5,
  ORC_ERR_KEYS = -10,   /* "None of the keys in kDict worked" (fksCuckoo.hpp:439-441) */
  ORC_ERR_INTERNAL = -9
};

/* the reference Tensor (tensor.hpp:19-26) */
typedef struct {
  uint32_t d;
  uint32_t nk;            /* num_key_dimensions */
  uint64_t N;
  const uint32_t* hedges; /* N*d row-major */
  const double* vals;
  const uint32_t* dims;   /* d */
  const uint32_t* dim_tc; /* d */
} orc_tensor;

typedef struct {
  uint32_t d;
  uint64_t N;
  uint32_t* hedges;
  double* vals;
  uint64_t flops;      /* total multiply-adds (the preprocessing bound) */
  uint64_t rows;       /* nonempty rows of A */
  uint64_t rows_done;  /* rows the numeric loop processed */
  double t_build_b, t_build_a, t_preprocess, t_contract; /* seconds */
} orc_result;

void orc_seed(uint64_t s);
void orc_set_plant(int p);
uint64_t orc_fastmod_M(uint32_t d);
uint32_t orc_fastmod_u32(uint32_t a, uint64_t M, uint32_t d);
int orc_tc_fksCuckoo(const orc_tensor* A, const orc_tensor* B, uint64_t max_rows, orc_result* out);
int orc_contract_meta(const orc_tensor* A, const orc_tensor* B, uint64_t max_rows, orc_result* out);
int orc_matricize_spGEMM(const orc_tensor* A, const orc_tensor* B, orc_result* out);
int orc_flops(const orc_tensor* A, const orc_tensor* B, uint64_t* flops_out);
void orc_result_free(orc_result* r);
#endif
