// stencil_2d_em.hip -- 2D euler_mobility instantiations of the fused stencil step.
#include "ch_internal.h"
#include "stencil_dispatch.h"

namespace chd {
hipError_t stencil2d_em(const Params &p, int precision, int model, int mob, int K, const void *in, void *out,
                        hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return stencil_model<double, 2, CH_LANDAU>(p, CH_EULER_MOBILITY, mob, K, in, out, s);
        return stencil_model<double, 2, CH_SIMPLE_WERTHEIM>(p, CH_EULER_MOBILITY, mob, K, in, out, s);
    }
    if(model == CH_LANDAU) return stencil_model<float, 2, CH_LANDAU>(p, CH_EULER_MOBILITY, mob, K, in, out, s);
    return stencil_model<float, 2, CH_SIMPLE_WERTHEIM>(p, CH_EULER_MOBILITY, mob, K, in, out, s);
}
}  // namespace chd
