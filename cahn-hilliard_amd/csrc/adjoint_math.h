// adjoint_math.h -- derivatives used by the reverse sweeps (FP64).
#pragma once

#include "ch_internal.h"

namespace chd {

// d^2 f / d rho^2 in FP64 (Landau: -eps + 3 psi^2; SimpleWertheim: 2 B2 + 1/rho
// + valence X'/X with X' = -c X^2 / (2 c rho X + 1), c = valence delta)
template <int MODEL>
__device__ __forceinline__ double fsecond(const Params &p, double r) {
    if constexpr(MODEL == CH_LANDAU) {
        return -p.eps + 3.0 * r * r;
    }
    else {
        double d = 2 * p.B2 + ((r < p.reg) ? 1.0 / p.reg : 1.0 / r);
        if(r > 0.) {
            const double c = 0.5 * p.tvd, X = wertheim_X(p, r);
            d += p.valence * (-c * X * X / (2.0 * c * r * X + 1.0)) / X;
        }
        return d;
    }
}

// d f'/d delta (SimpleWertheim, rho > 0): valence X_delta / X with
// X_delta = -valence rho X^2 / (2 c rho X + 1), c = valence delta
template <int MODEL>
__device__ __forceinline__ double dfprime_ddelta(const Params &p, double r) {
    if constexpr(MODEL == CH_LANDAU) {
        return 0.0;
    }
    else {
        if(!(r > 0.)) return 0.0;
        const double c = 0.5 * p.tvd, X = wertheim_X(p, r);
        return -(double)p.valence * p.valence * r * X / (2.0 * c * r * X + 1.0);
    }
}

// d f/d delta (SimpleWertheim, rho > 0): valence rho X_delta (1/X - 1/2)
template <int MODEL>
__device__ __forceinline__ double dbulk_ddelta(const Params &p, double r) {
    if constexpr(MODEL == CH_LANDAU) {
        return 0.0;
    }
    else {
        if(!(r > 0.)) return 0.0;
        const double c = 0.5 * p.tvd, X = wertheim_X(p, r);
        const double Xd = -(double)p.valence * r * X * X / (2.0 * c * r * X + 1.0);
        return p.valence * r * Xd * (1.0 / X - 0.5);
    }
}

// mobility(rho) / M, d mobility / d rho and d mobility / d delta (FP64)
template <int MOB>
struct MobDeriv {
    double hat, drho, ddelta;
};

template <int MOB>
__device__ __forceinline__ MobDeriv<MOB> mob_deriv(const Params &p, double r) {
    MobDeriv<MOB> m{1.0, 0.0, 0.0};
    if constexpr(MOB == CH_MOB_FREE_ENERGY) {
        const double c = 0.5 * p.tvd, X = wertheim_X(p, r), v = p.valence;
        const double den = 2.0 * c * r * X + 1.0;
        const double Xv1 = pow(X, v - 1.0);
        m.hat = r * Xv1 * X;
        m.drho = p.M * (Xv1 * X + r * v * Xv1 * (-c * X * X / den));
        m.ddelta = p.M * r * v * Xv1 * (-v * r * X * X / den);
    }
    else if constexpr(MOB == CH_MOB_REGULARISED) {
        m.hat = r / (r + p.rho_min);
        m.drho = p.M * p.rho_min / ((r + p.rho_min) * (r + p.rho_min));
    }
    else if constexpr(MOB == CH_MOB_LINEAR) {
        m.hat = r / p.rho_ref;
        m.drho = p.M / p.rho_ref;
    }
    return m;
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// adds the block's sums of the kGradSlots per-thread values to its own slots
template <int NT>
__device__ __forceinline__ void block_accumulate(const double (&g)[kGradSlots], double (&red)[kGradSlots][NT / 64],
                                                 double *__restrict__ slots) {
    const int w = threadIdx.x >> 6;
#pragma unroll
    for(int i = 0; i < kGradSlots; i++) {
        const double v = wave_sum64(g[i]);
        if((threadIdx.x & 63) == 0) red[i][w] = v;
    }
    __syncthreads();
    if(threadIdx.x < kGradSlots) {
        double a = 0;
        for(int i = 0; i < NT / 64; i++) a += red[threadIdx.x][i];
        slots[threadIdx.x] += a;
    }
}

}  // namespace chd
