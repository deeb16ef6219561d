// stencil_2d.hip -- 2D instantiations of the fused stencil step (euler); the
// euler_mobility instantiations live in stencil_2d_em.hip.
#include "ch_internal.h"
#include "stencil_dispatch.h"

namespace chd {
hipError_t stencil2d_em(const Params &p, int precision, int model, int mob, int K, const void *in, void *out,
                        hipStream_t s);

template <> hipError_t stencil_step<2>(const Params &p, int precision, int integrator, int model, int mob, int K,
                                       const void *in, void *out, hipStream_t s) {
    if(integrator != CH_EULER) return stencil2d_em(p, precision, model, mob, K, in, out, s);
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return stencil_k<double, 2, CH_LANDAU, CH_MOB_CONSTANT, false>(p, K, in, out, s);
        return stencil_k<double, 2, CH_SIMPLE_WERTHEIM, CH_MOB_CONSTANT, false>(p, K, in, out, s);
    }
    if(model == CH_LANDAU) return stencil_k<float, 2, CH_LANDAU, CH_MOB_CONSTANT, false>(p, K, in, out, s);
    return stencil_k<float, 2, CH_SIMPLE_WERTHEIM, CH_MOB_CONSTANT, false>(p, K, in, out, s);
}
}  // namespace chd
