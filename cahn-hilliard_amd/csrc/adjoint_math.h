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

// Per-cell terms of the euler_mobility reverse step (EulerMobilityCPU.cpp:44-82
// reversed), written without contraction so that every kernel that shares
// them (the LDS tile and the three-launch form) gives the same bits.
// nu = D(Mob, lambda): the face-flux divergence of lambda with the mobility
// Mi of the cell and Mm / Mq of its -e_d / +e_d neighbours
template <typename T, int DIMS>
__device__ __forceinline__ T adj_nu(T l, T Mi, const T (&Lm)[DIMS], const T (&Lq)[DIMS], const T (&Mm)[DIMS],
                                    const T (&Mq)[DIMS], T inv_dx2) {
#pragma clang fp contract(off)
    const T half = (T)0.5;
    T s = 0;
#pragma unroll
    for(int d = 0; d < DIMS; d++) s = s + half * (Mi + Mq[d]) * (Lq[d] - l) - half * (Mm[d] + Mi) * (l - Lm[d]);
    return s * inv_dx2;
}
// eta = -1/2 sum_d [(mu_+ - mu)(lambda_+ - lambda) + (mu - mu_-)(lambda - lambda_-)] / dx^2:
// the mobility's share of the reverse step
template <typename T, int DIMS>
__device__ __forceinline__ T adj_eta(T mu, T l, const T (&Um)[DIMS], const T (&Uq)[DIMS], const T (&Lm)[DIMS],
                                     const T (&Lq)[DIMS], T inv_dx2) {
#pragma clang fp contract(off)
    const T half = (T)0.5;
    T se = 0;
#pragma unroll
    for(int d = 0; d < DIMS; d++) se = se + (Uq[d] - mu) * (Lq[d] - l) + (mu - Um[d]) * (l - Lm[d]);
    return -half * se * inv_dx2;
}
// lambda_n = lambda + dt (f'' nu - 2k lap nu + Mob' eta)
template <typename T>
__device__ __forceinline__ T adj_em_step(const Params &p, T l, T fpp, T nu, T lnu, T drho, T eta) {
#pragma clang fp contract(off)
    const T upd = fpp * nu - (T)p.two_k * lnu + drho * eta;
    return l + (T)p.dt * upd;
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
