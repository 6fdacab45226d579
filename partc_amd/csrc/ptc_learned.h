// Learned row-size prediction: features of every output row and the
// network that maps them to a predicted number of nonzeros.
#ifndef PTC_LEARNED_H
#define PTC_LEARNED_H

#include "ptc_internal.h"

namespace ptc {

struct LearnedInputs {
  uint64_t nrows;
  const KtSlice* rows;                 // A positions of each row
  const unsigned long long* row_flops; // sum of probed slice lengths
  const uint32_t* row_hits;            // distinct contracted keys that hit B
  const uint32_t* row_maxlen;          // longest slice
  uint64_t ncols;                      // column space
  // with row_hits NULL (capacities only) the kernel finds the hits and the
  // longest slice itself: from every position's slice length and the
  // TilePart of every TILE_PART positions (parts strictly inside a row)
  const uint32_t* pos_len;
  const TilePart* tiles;
};

// pred[r] = predicted nonzeros of row r (device array of nrows floats).
int learned_predict(ptc_ctx* ctx, const ptc_mlp* mlp, const LearnedInputs& in, float* pred);
// cap[r] = the learned capacity of row r (features, inference and the
// capacity rule of the single-pass methods in one launch).
int learned_caps(ptc_ctx* ctx, const ptc_mlp* mlp, const LearnedInputs& in, double scale, uint32_t* cap);
// Feature rows X (nrows x PTC_MLP_IN) and upper bounds ub (device arrays).
int learned_features(ptc_ctx* ctx, const LearnedInputs& in, float* X, float* ub);

}  // namespace ptc

extern "C" const ptc_mlp* ptc_learned_default(void);

#endif
