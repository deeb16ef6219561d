// stencil_3d.hip -- 3D instantiations of the fused stencil step.
#include "ch_internal.h"
#include "stencil_dispatch.h"

namespace chd {
template <> hipError_t stencil_step<3>(const Params &p, int precision, int integrator, int model, int mob, int K,
                                         const void *in, void *out, hipStream_t s) {
    return stencil_step_dims<3>(p, precision, integrator, model, mob, K, in, out, s);
}
}  // namespace chd
