// Row-size network (placeholder until the trained predictor lands).
#include <hip/hip_runtime.h>

#include "ptc_learned.h"

namespace ptc {
int learned_predict(ptc_ctx* ctx, const ptc_mlp* mlp, const LearnedInputs& in, float* pred) {
  (void)mlp; (void)in; (void)pred;
  return set_err(ctx, PTC_ERR_UNSUPPORTED, "row-size network not available");
}
}  // namespace ptc

extern "C" const ptc_mlp* ptc_learned_default(void) { return nullptr; }
