// Mixed-precision evaluation (FP32 monomial products, FP64 Boys and sums).
#include "fe_internal.h"

int fe_launch_fwd_mixed(fe_table*, const FwdPlan&, const double*, int, double*, hipStream_t) {
  return fe_set_error(FE_ERR_UNSUPPORTED, "mixed precision is not built yet");
}
