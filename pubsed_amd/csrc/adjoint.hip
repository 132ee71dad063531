// adjoint.hip -- differentiable transport (continuous-absorption estimator).
#include "sgpu_internal.h"

extern "C" int sgpu_adjoint_forward(sgpu_ctx* ctx, const sgpu_grid*, const sgpu_locate*, const sgpu_locate*,
                                    const float*, const double*, const double*, const sgpu_transport_params*,
                                    const sgpu_packets*, int64_t, double*, int32_t*, double*)
{
  SGPU_REQUIRE(ctx && 0, "sgpu_adjoint_forward: not built yet");
  return SGPU_ERR_ARG;
}

extern "C" int sgpu_adjoint_backward(sgpu_ctx* ctx, const sgpu_grid*, const sgpu_locate*, const sgpu_locate*,
                                     const float*, const double*, const double*, const sgpu_transport_params*,
                                     const sgpu_packets*, int64_t, const double*, const int32_t*, const double*,
                                     double*)
{
  SGPU_REQUIRE(ctx && 0, "sgpu_adjoint_backward: not built yet");
  return SGPU_ERR_ARG;
}
