// Learned preconditioner correction (placeholder until the message-passing network
// lands): no correction is applied.
#include "dev_common.hpp"
#include "solvers.hpp"

namespace senk_amd {

int learned_apply_f32(Ctx *, const PrecView &, const float *, const double *, float *, const int *) {
    return SENK_ERR_UNSUPPORTED;
}
int learned_apply_f64(Ctx *, const PrecView &, const double *, const double *, double *, const int *) {
    return SENK_ERR_UNSUPPORTED;
}

}  // namespace senk_amd
