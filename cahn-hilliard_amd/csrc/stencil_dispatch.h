// stencil_dispatch.h -- runtime descriptor -> k_stencil instantiation.
#pragma once

#include "stencil.h"
#include "stencil_wave.h"

namespace chd {

// Tile shapes (TX, TY, TZ) per dimension; the largest that fits N is used.
//   1D: 256 | 64 | 16 | 4      2D: 64x16 | 16x16 | 4x4      3D: 16x8x8 | 4x4x4
// (3D, N >= 16: Params::tile3d selects 16x8x8 with 256 or 512 threads, or 16x16x4 with 512)
template <typename T, int DIMS, int MODEL, int MOB, bool EM, int K>
hipError_t stencil_tiles(const Params &p, const void *in, void *out, hipStream_t s) {
    const int N = p.N;
    if constexpr(DIMS == 1) {
        if(N >= 256) return launch_stencil_cfg<T, 1, MODEL, MOB, EM, 256, 1, 1, K>(p, in, out, s);
        if(N >= 64) return launch_stencil_cfg<T, 1, MODEL, MOB, EM, 64, 1, 1, K>(p, in, out, s);
        if(N >= 16) return launch_stencil_cfg<T, 1, MODEL, MOB, EM, 16, 1, 1, K>(p, in, out, s);
        return launch_stencil_cfg<T, 1, MODEL, MOB, EM, 4, 1, 1, K>(p, in, out, s);
    }
    else if constexpr(DIMS == 2) {
        if(N >= 64) return launch_stencil_cfg<T, 2, MODEL, MOB, EM, 64, 16, 1, K>(p, in, out, s);
        if(N >= 16) return launch_stencil_cfg<T, 2, MODEL, MOB, EM, 16, 16, 1, K>(p, in, out, s);
        return launch_stencil_cfg<T, 2, MODEL, MOB, EM, 4, 4, 1, K>(p, in, out, s);
    }
    else {
        if(N >= 16) {
            if(p.tile3d == 1) return launch_stencil_cfg<T, 3, MODEL, MOB, EM, 16, 8, 8, K, 512>(p, in, out, s);
            if(p.tile3d == 2) return launch_stencil_cfg<T, 3, MODEL, MOB, EM, 16, 16, 4, K, 512>(p, in, out, s);
            return launch_stencil_cfg<T, 3, MODEL, MOB, EM, 16, 8, 8, K>(p, in, out, s);
        }
        return launch_stencil_cfg<T, 3, MODEL, MOB, EM, 4, 4, 4, K>(p, in, out, s);
    }
}

// 2D grids with N >= 64 use the register/DPP wave kernel.  Region shapes
// (p.wave_cfg), columns per lane x rows per wave x waves (region W x H):
//   0 = 1 x 16 x 4 (64 x 64)    1 = 2 x 8 x 4 (128 x 32)   2 = 2 x 8 x 8 (128 x 64)
//   3 = 1 x 8 x 8 (64 x 64)     4 = 1 x 4 x 16 (64 x 64)   5 = 2 x 4 x 16 (128 x 64)
//   6 = 1 x 8 x 16 (64 x 128)   7 = 2 x 5 x 16 (128 x 80)  8 = 3 x 4 x 16 (192 x 64)
// Fewer cells per lane means more waves per CU to hide latency, at the price
// of more row exchanges through LDS.  K in {1, 2, 4, 8} steps per launch where
// the region leaves an interior.
using WaveCfg0 = WaveCfg<1, 16, 4>;
using WaveCfg1 = WaveCfg<2, 8, 4>;
using WaveCfg2 = WaveCfg<2, 8, 8>;
using WaveCfg3 = WaveCfg<1, 8, 8>;
using WaveCfg4 = WaveCfg<1, 4, 16>;
using WaveCfg5 = WaveCfg<2, 4, 16>;
using WaveCfg6 = WaveCfg<1, 8, 16>;
using WaveCfg7 = WaveCfg<2, 5, 16>;
using WaveCfg8 = WaveCfg<3, 4, 16>;
static_assert(kWaveCfgs == 9, "one WaveCfg per p.wave_cfg value");

template <class CFG, int K>
constexpr bool wave_ok() { return CFG::RW - 4 * K > 0 && CFG::RH - 4 * K > 0; }

template <typename T, int MODEL, int MOB, bool EM, class CFG>
hipError_t wave_k(const Params &p, int K, const void *in, void *out, hipStream_t s) {
    switch(K) {
    case 1: return launch_wave_cfg<T, MODEL, MOB, EM, 1, CFG>(p, in, out, s);
    case 2: return launch_wave_cfg<T, MODEL, MOB, EM, 2, CFG>(p, in, out, s);
    case 4: return launch_wave_cfg<T, MODEL, MOB, EM, 4, CFG>(p, in, out, s);
    case 8:
        if constexpr(wave_ok<CFG, 8>()) return launch_wave_cfg<T, MODEL, MOB, EM, 8, CFG>(p, in, out, s);
        return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template <typename T, int DIMS, int MODEL, int MOB, bool EM>
hipError_t stencil_k(const Params &p, int K, const void *in, void *out, hipStream_t s) {
    if constexpr(DIMS == 2) {
        if(p.N >= 64) {
            if(p.N >= 128) {
                switch(p.wave_cfg) {
                case 1: return wave_k<T, MODEL, MOB, EM, WaveCfg1>(p, K, in, out, s);
                case 2: return wave_k<T, MODEL, MOB, EM, WaveCfg2>(p, K, in, out, s);
                case 5: return wave_k<T, MODEL, MOB, EM, WaveCfg5>(p, K, in, out, s);
                case 6: return wave_k<T, MODEL, MOB, EM, WaveCfg6>(p, K, in, out, s);
                case 7: return wave_k<T, MODEL, MOB, EM, WaveCfg7>(p, K, in, out, s);
                case 8: return wave_k<T, MODEL, MOB, EM, WaveCfg8>(p, K, in, out, s);
                default: break;
                }
            }
            if(p.wave_cfg == 3) return wave_k<T, MODEL, MOB, EM, WaveCfg3>(p, K, in, out, s);
            if(p.wave_cfg == 4) return wave_k<T, MODEL, MOB, EM, WaveCfg4>(p, K, in, out, s);
            return wave_k<T, MODEL, MOB, EM, WaveCfg0>(p, K, in, out, s);
        }
    }
    if(K != 1) return hipErrorInvalidValue;
    return stencil_tiles<T, DIMS, MODEL, MOB, EM, 1>(p, in, out, s);
}

template <typename T, int DIMS, int MODEL>
hipError_t stencil_model(const Params &p, int integrator, int mob, int K, const void *in, void *out, hipStream_t s) {
    if(integrator == CH_EULER) return stencil_k<T, DIMS, MODEL, CH_MOB_CONSTANT, false>(p, K, in, out, s);
    switch(mob) {
    case CH_MOB_CONSTANT: return stencil_k<T, DIMS, MODEL, CH_MOB_CONSTANT, true>(p, K, in, out, s);
    case CH_MOB_REGULARISED: return stencil_k<T, DIMS, MODEL, CH_MOB_REGULARISED, true>(p, K, in, out, s);
    case CH_MOB_LINEAR: return stencil_k<T, DIMS, MODEL, CH_MOB_LINEAR, true>(p, K, in, out, s);
    case CH_MOB_FREE_ENERGY:
        if constexpr(MODEL == CH_SIMPLE_WERTHEIM)
            return stencil_k<T, DIMS, MODEL, CH_MOB_FREE_ENERGY, true>(p, K, in, out, s);
        return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template <int DIMS>
hipError_t stencil_step_dims(const Params &p, int precision, int integrator, int model, int mob, int K,
                             const void *in, void *out, hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return stencil_model<double, DIMS, CH_LANDAU>(p, integrator, mob, K, in, out, s);
        return stencil_model<double, DIMS, CH_SIMPLE_WERTHEIM>(p, integrator, mob, K, in, out, s);
    }
    if(model == CH_LANDAU) return stencil_model<float, DIMS, CH_LANDAU>(p, integrator, mob, K, in, out, s);
    return stencil_model<float, DIMS, CH_SIMPLE_WERTHEIM>(p, integrator, mob, K, in, out, s);
}

}  // namespace chd
