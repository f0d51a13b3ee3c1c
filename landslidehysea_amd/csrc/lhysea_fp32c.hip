// FP32-compensated step kernels (placeholder until the FP32c path lands).
#include "lhysea_internal.hpp"

namespace lhysea {

hipError_t f32c_initial_dt(const Consts&, const float*, Ctl*, hipStream_t) { return hipErrorNotSupported; }
hipError_t f32c_step_fused(const Consts&, const float*, float*, const double*, const StepArgs&, hipStream_t)
{
    return hipErrorNotSupported;
}

}  // namespace lhysea
