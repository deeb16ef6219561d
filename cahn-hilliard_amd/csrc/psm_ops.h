// psm_ops.h -- the per-entry operations of the pseudospectral_mobility
// divergence term (src/integrators/PseudospectralMobilityCPU.cpp:252-287),
// shared by the element-wise kernels (psm.hip) and the fused transform passes
// (fft.hip), so that both give the same bits.
#pragma once

#include "ch_device.h"

namespace chd {

// (i k_axis) deal z for the half-spectrum entry (ix, iy, iz) (:252-258)
template <typename T, typename T2>
__device__ __forceinline__ T2 psm_grad_term(const Params &p, int ix, int iy, int iz, int axis, T2 z) {
#pragma clang fp contract(off)
    double k2, deal;
    mode_q2(p, ix, iy, iz, k2, deal);
    if(!p.use_dealias) deal = 1.0;
    const T kd = (T)(p.k_unit * axis_wavenumber(p, ix, iy, iz, axis)), d = (T)deal;
    return T2{-kd * z.y * d, kd * z.x * d};
}

// div (=|+=) (i k_axis) deal F (:278-287); axis 0 starts the sum
template <typename T, typename T2>
__device__ __forceinline__ T2 psm_div_term(const Params &p, int ix, int iy, int iz, int axis, T2 F, T2 div) {
#pragma clang fp contract(off)
    double k2, deal;
    mode_q2(p, ix, iy, iz, k2, deal);
    if(!p.use_dealias) deal = 1.0;
    const T kd = (T)(p.k_unit * axis_wavenumber(p, ix, iy, iz, axis)), d = (T)deal;
    const T2 f{F.x * d, F.y * d};
    const T2 t{-kd * f.y, kd * f.x};
    if(axis == 0) return T2{(T)0 + t.x, (T)0 + t.y};
    return T2{div.x + t.x, div.y + t.y};
}

// the flux w g / N^d of one real cell, w = M - M0 (splitting) or M (:264-276)
template <typename T>
__device__ __forceinline__ T psm_flux(T g, T mob, T M0, T ib) {
#pragma clang fp contract(off)
    return (mob - M0) * (g * ib);
}

}  // namespace chd
