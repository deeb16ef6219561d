// ch_device.h -- shared device-side definitions of the gfx950 library.
//
// Model and mobility functions are evaluated inside the stencil / FFT kernels
// (fused), instead of the reference's one-kernel-per-model-call pattern
// (src/CUDA/models/Landau.cu:5-19, SimpleWertheim.cu:10-42).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ch_hip.h"

namespace chd {

// Scalars every kernel needs, passed by value as a kernel argument.
struct Params {
    int N, mask, log2N, dims;
    int valence;
    int use_dealias;
    double eps;            // Landau epsilon
    double B2, reg, log_reg, tvd;   // SimpleWertheim: tvd = 2 valence delta
    double M, rho_min, rho_ref;     // mobility
    double k, two_k, dx, inv_dx, inv_dx2, dt, S;
    double vbin_over_bins;          // V_bin / bins
    double inv_bins;                // 1 / N^dims (c2r normalisation)
};

// ---------------------------------------------------------------- models
// Landau: f' = -eps psi + psi^3 (src/models/Landau.cpp:46-52), evaluated in T.
// SimpleWertheim: f' = 2 B2 rho + ln rho (or the linear branch below
// regularisation_delta) + valence ln X(rho) (src/models/SimpleWertheim.cpp:40-53);
// X(rho) and its logs are evaluated in FP64 in both precisions.
__device__ __forceinline__ double wertheim_X(const Params &p, double r) {
    return (-1.0 + sqrt(1.0 + 2.0 * p.tvd * r)) / (p.tvd * r);
}

template <typename T, int MODEL>
__device__ __forceinline__ T fprime(const Params &p, T r) {
    if constexpr(MODEL == CH_LANDAU) {
        return -(T)p.eps * r + r * r * r;
    }
    else {
        double rd = (double)r;
        double d = 2 * p.B2 * rd;
        if(rd < p.reg) d += p.log_reg + (rd - p.reg) / p.reg;
        else d += log(rd);
        double db = (rd > 0.) ? p.valence * log(wertheim_X(p, rd)) : 0.0;
        return (T)(d + db);
    }
}

// f(rho), FP64 (observables): Landau.cpp:54-57, SimpleWertheim.cpp:71-86
template <int MODEL>
__device__ __forceinline__ double bulk(const Params &p, double r) {
    if constexpr(MODEL == CH_LANDAU) {
        return -0.5 * p.eps * r * r + 0.25 * (r * r) * (r * r);
    }
    else {
        double f_ref = p.B2 * r * r;
        if(r < p.reg) {
            double a = 1 / (2.0 * p.reg);
            double b = p.log_reg - 1.0;
            double c = p.reg * (p.log_reg - 1.0) - a * p.reg * p.reg - b * p.reg;
            f_ref += a * r * r + b * r + c;
        }
        else {
            f_ref += r * (log(r) - 1.0);
        }
        double f_bond = 0.0;
        if(r > 0.) {
            double X = wertheim_X(p, r);
            f_bond = p.valence * r * (log(X) + 0.5 * (1. - X));
        }
        return f_ref + f_bond;
    }
}

// ---------------------------------------------------------------- mobility
// constant (ConstantMobility.h:19-23), free_energy (FreeEnergyMobility.h:25-34
// -> SimpleWertheim.cpp:104-111), regularised (RegularisedMobility.h:26-33),
// linear (new): M rho / rho_ref.
template <typename T, int MOB>
__device__ __forceinline__ T mobility(const Params &p, T r) {
    if constexpr(MOB == CH_MOB_CONSTANT) {
        return (T)p.M;
    }
    else if constexpr(MOB == CH_MOB_FREE_ENERGY) {
        double rd = (double)r;
        double X = wertheim_X(p, rd);
        double D = p.M * pow(X, (double)p.valence);
        return (T)(rd * D);
    }
    else if constexpr(MOB == CH_MOB_REGULARISED) {
        return (T)p.M * r / (r + (T)p.rho_min);
    }
    else {
        return (T)p.M * r / (T)p.rho_ref;
    }
}

// XCD-aware block order: gfx950 hands workgroups to its 8 XCDs round-robin by
// id, so block b runs on XCD b % 8.  Mapping XCD x to the x-th contiguous
// range of tiles keeps a band of the domain in one XCD's L2 from step to step.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    if(nblocks % 8 != 0) return b;
    int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename T> struct Complex;
template <> struct Complex<double> { using type = double2; };
template <> struct Complex<float> { using type = float2; };

}  // namespace chd
