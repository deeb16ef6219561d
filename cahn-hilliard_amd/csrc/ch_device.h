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
    int wave_cfg;          // region shape of the 2D wave kernel (0 .. kWaveCfgs - 1)
    int adj_cfg;           // region shape of the 2D wave adjoint (0, 3, 4 or 5)
    int tile3d;            // 3D forward step: LDS tile 0: 16x8x8 / 256 threads, 1: the same / 512, 2: 16x16x4 / 512; 3: two launches (stencil_split.hip)
    int adj_tile3d;        // 3D tile of the euler_mobility adjoint (0: 16x4x4 / 256, 1 (default): 16x8x4 / 512, 2: 16x8x8 / 512)
    int noise_tile;        // tile of the noise step (0: 2D 64x16, 3D 8x8x8 / 256 threads; 1: the same / 512; 2 (default): 2D 64x32, 3D 16x8x8 / 512)
    double eps;            // Landau epsilon
    double B2, reg, log_reg, tvd;   // SimpleWertheim: tvd = 2 valence delta
    double M, rho_min, rho_ref;     // mobility
    double k, two_k, dx, inv_dx, inv_dx2, dt, S;
    double Mdt;                     // M dt (constant-mobility Euler update)
    double mu_lap, upd_lap;         // -2 k / dx^2 and M dt / dx^2 (the folded cell arithmetic)
    double q2_unit;                 // (2 pi / (N dx))^2: q^2 of a unit wave vector
    double vbin_over_bins;          // V_bin / bins
    double inv_bins;                // 1 / N^dims (c2r normalisation)
    double k_unit;                  // 2 pi / (N dx): wave number of a unit index
    double dcf;                     // density conversion: ln(rho * dcf) in the Wertheim f
    double noise;                   // sqrt(2 / (dx^dims dt)) noise_rescale_factor (0: no noise)
    uint64_t seed;                  // key of the noise generator
};

// Every per-cell expression below is written with contraction off and
// explicit fused multiply-adds, so that all kernels which share these
// functions (any tile shape, any number of steps per launch) produce
// bit-identical fields.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ---------------------------------------------------------------- models
// Landau: f' = -eps psi + psi^3 (src/models/Landau.cpp:46-52), evaluated in T.
// SimpleWertheim: f' = 2 B2 rho + ln rho (or the linear branch below
// regularisation_delta) + valence ln X(rho) (src/models/SimpleWertheim.cpp:40-53);
// X(rho) and its logs are evaluated in FP64 in both precisions, above
// regularisation_delta as the single logarithm ln(rho X^valence).
// Natural logarithm for the per-cell free-energy derivative.  x = m 2^e with
// m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1) with
// |s| <= 3 - 2 sqrt(2), summed as its Maclaurin series 2 s sum_k s^(2k) / (2k + 1)
// through k = 10 (the first omitted term is below 1e-18 of the sum); m - 1 is
// exact, s comes from a reciprocal refined by Newton steps and one residual
// correction, and ln 2 is split into a 31-bit head (e ln2_head exact for any
// exponent) plus the rest of its value.  Within 2 ulp (worst near m = 1) in
// some 30 instructions, where the library log spends about 100; zero,
// negative, infinite and NaN arguments go to log().
__device__ __forceinline__ double log_series(double x) {
    if(!(x > 0.0) || x == __builtin_inf()) return log(x);
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1), denormals included
    int e = __builtin_amdgcn_frexp_exp(x);
    if(m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0, d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    double sv = f * r;
    sv = __builtin_fma(__builtin_fma(-d, sv, f), r, sv);
    const double t = sv * sv;
    double q = 1.0 / 21.0;
    q = __builtin_fma(q, t, 1.0 / 19.0);
    q = __builtin_fma(q, t, 1.0 / 17.0);
    q = __builtin_fma(q, t, 1.0 / 15.0);
    q = __builtin_fma(q, t, 1.0 / 13.0);
    q = __builtin_fma(q, t, 1.0 / 11.0);
    q = __builtin_fma(q, t, 1.0 / 9.0);
    q = __builtin_fma(q, t, 1.0 / 7.0);
    q = __builtin_fma(q, t, 1.0 / 5.0);
    q = __builtin_fma(q, t, 1.0 / 3.0);
    const double u = 2.0 * sv;
    const double lm = __builtin_fma(u * t, q, u);
    const double ed = (double)e;
    constexpr double ln2_head = 0.6931471803691238, ln2_rest = 1.9082149292705877e-10;
    return __builtin_fma(ed, ln2_head, __builtin_fma(ed, ln2_rest, lm));
}

__device__ __forceinline__ double wertheim_X(const Params &p, double r) {
    return (-1.0 + sqrt(1.0 + 2.0 * p.tvd * r)) / (p.tvd * r);
}

template <typename T, int MODEL>
__device__ __forceinline__ T fprime(const Params &p, T r) {
#pragma clang fp contract(off)
    if constexpr(MODEL == CH_LANDAU) {
        // psi (psi^2 - eps): two operations and two roundings for the
        // reference's -eps psi + psi^3
        return r * fma_t(r, r, -(T)p.eps);
    }
    else {
        const double rd = (double)r;
        double d = 2 * p.B2 * rd;
        if(rd < p.reg) {
            d += p.log_reg + (rd - p.reg) / p.reg;
            if(rd > 0.) d += p.valence * log_series(wertheim_X(p, rd));
        }
        else if(rd > 0.) {
            // ln rho + valence ln X = ln(rho X^valence): one logarithm
            const double X = wertheim_X(p, rd);
            double xv = rd;
            for(int i = 0; i < p.valence; i++) xv *= X;
            d += log_series(xv);
        }
        else {
            d += log(rd);
        }
        return (T)d;
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
            f_ref += r * (log(r * p.dcf) - 1.0);
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

// ---------------------------------------------------------------- cells
// Periodic (2d+1)-point Laplacian with the neighbours summed in the
// reference's order -x, +x, -y, +y, -z, +z (EulerCPU.cpp:71-84);
// m[d] / q[d] are the -e_d / +e_d neighbours.
template <typename T, int DIMS>
__device__ __forceinline__ T laplacian(T c, const T (&m)[DIMS], const T (&q)[DIMS], T inv_dx2) {
#pragma clang fp contract(off)
    T s = m[0] + q[0];
#pragma unroll
    for(int d = 1; d < DIMS; d++) {
        s = s + m[d];
        s = s + q[d];
    }
    return fma_t((T)(-2 * DIMS), c, s) * inv_dx2;
}

// The Laplacian without its 1 / dx^2: the neighbour sum minus 2 d times the
// centre.  The cells fold 1 / dx^2 into the constant that multiplies it
// (-2 k / dx^2 for mu, M dt / dx^2 for the update): one multiplication less
// per Laplacian, and the same bits as lap = sum / dx^2 when dx = 1.
template <typename T, int DIMS>
__device__ __forceinline__ T laplacian_sum(T c, const T (&m)[DIMS], const T (&q)[DIMS]) {
#pragma clang fp contract(off)
    T s = m[0] + q[0];
#pragma unroll
    for(int d = 1; d < DIMS; d++) {
        s = s + m[d];
        s = s + q[d];
    }
    return fma_t((T)(-2 * DIMS), c, s);
}

// mu = f'(psi) - 2 k lap(psi)  (EulerCPU.cpp:20-26)
template <typename T, int DIMS, int MODEL>
__device__ __forceinline__ T chem_pot(const Params &p, T c, const T (&m)[DIMS], const T (&q)[DIMS]) {
#pragma clang fp contract(off)
    return fma_t((T)p.mu_lap, laplacian_sum<T, DIMS>(c, m, q), fprime<T, MODEL>(p, c));
}

// dt M lap(mu), the increment of the constant-mobility Euler update
template <typename T, int DIMS>
__device__ __forceinline__ T euler_inc(const Params &p, T mu, const T (&m)[DIMS], const T (&q)[DIMS], T base) {
#pragma clang fp contract(off)
    return fma_t((T)p.upd_lap, laplacian_sum<T, DIMS>(mu, m, q), base);
}

// psi + dt M lap(mu)  (EulerCPU.cpp:29-35), constant mobility
template <typename T, int DIMS>
__device__ __forceinline__ T euler_update(const Params &p, T psi, T mu, const T (&m)[DIMS], const T (&q)[DIMS]) {
    return euler_inc<T, DIMS>(p, mu, m, q, psi);
}

// the same update as a compensated sum: the FP32 rounding lost in psi + dt M
// lap(mu) is kept in cmp and given back on the next step (Kahan), so that long
// FP32 runs do not accumulate it
template <typename T, int DIMS>
__device__ __forceinline__ T euler_update_comp(const Params &p, T psi, T mu, const T (&m)[DIMS], const T (&q)[DIMS],
                                               T &cmp) {
#pragma clang fp contract(off)
    const T y = euler_inc<T, DIMS>(p, mu, m, q, -cmp);
    const T t = psi + y;
    cmp = (t - psi) - y;
    return t;
}

// the update of a launch-compensated step: the increments of the launch so far
// are summed in acc (small beside psi, so their sum keeps nearly all their
// bits) and the field is psi0 + acc, psi0 being the value the launch loaded;
// psi is rounded once from the launch's total increment instead of once per step
template <typename T, int DIMS>
__device__ __forceinline__ void euler_increment(const Params &p, T mu, const T (&m)[DIMS], const T (&q)[DIMS], T &acc) {
    acc = euler_inc<T, DIMS>(p, mu, m, q, acc);
}

// An 8-bit low-order part of an FP32 value hi: lo = q 2^(E - 32) where
// hi = m 2^E with 0.5 <= |m| < 1 (frexp), i.e. q in units of ulp(hi) / 256.
// encode: the q nearest to a remainder e with |e| <= ulp(hi) / 2 (|q| <= 127).
__device__ __forceinline__ float low8_decode(float hi, int q) {
    return __builtin_ldexpf((float)q, __builtin_amdgcn_frexp_expf(hi) - 32);
}
__device__ __forceinline__ int low8_encode(float hi, float e) {
    const int q = (int)__builtin_rintf(__builtin_ldexpf(e, 32 - __builtin_amdgcn_frexp_expf(hi)));
    return q < -127 ? -127 : (q > 127 ? 127 : q);
}

// psi + dt div J with face fluxes J_d(i) = (M_i + M_{i+e_d})/2 (mu_{i+e_d} - mu_i)/dx
// (EulerMobilityCPU.cpp:44-82, divergence EulerCPU.cpp:114-128)
template <typename T, int DIMS>
__device__ __forceinline__ T mobility_update(const Params &p, T psi, T mu, const T (&mum)[DIMS], const T (&muq)[DIMS],
                                             T M, const T (&Mm)[DIMS], const T (&Mq)[DIMS]) {
#pragma clang fp contract(off)
    const T half = (T)0.5, inv_dx = (T)p.inv_dx;
    T res = 0;
#pragma unroll
    for(int d = 0; d < DIMS; d++) {
        const T Jq = half * (M + Mq[d]) * (muq[d] - mu) * inv_dx;
        const T Jm = half * (Mm[d] + M) * (mu - mum[d]) * inv_dx;
        res = res + (Jq - Jm) * inv_dx;
    }
    return fma_t(res, (T)p.dt, psi);
}

// A workgroup barrier that orders LDS only: the release and acquire fences are
// restricted to the local address space, so it waits for the wave's LDS
// operations (lgkmcnt) but not for its outstanding global loads and stores, as
// __syncthreads does (vmcnt(0)).  For exchanges through LDS whose global
// accesses need no ordering between the block's threads.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders the LDS accesses of one wave's lanes among themselves, for a buffer
// no other wave touches meanwhile: a wave's LDS instructions execute in issue
// order, so no wait and no s_barrier is needed, only that the compiler keep
// the accesses on their side of this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// XCD-aware block order: gfx950 hands workgroups to its 8 XCDs round-robin by
// id, so block b runs on XCD b % 8.  Mapping XCD x to the x-th contiguous
// range of tiles keeps a band of the domain in one XCD's L2 from step to step.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    if(nblocks % 8 != 0) return b;
    int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------- spectra
// q^2 and the dealias factor of a half-spectrum mode (ix: half-axis index).
// The reference's tables (PseudospectralCPU.cpp:34-83) hold q = 2 pi |n| / (N dx)
// for the integer wave vector n, and keep modes with q < k_cut (2D) or
// q <= k_cut (1D, 3D), k_cut = (2/3) pi / dx, i.e. 9 |n|^2 < N^2 (<=).  The
// cut is tested on the integers: N is a power of two, so 9 |n|^2 never equals
// N^2 and |9 |n|^2 - N^2| >= 1 keeps every floating-point evaluation of the
// reference's test on the same side.  q^2 = q2_unit |n|^2 with
// q2_unit = (2 pi / (N dx))^2: no square root.
__device__ __forceinline__ void mode_q2(const Params &p, int ix, int iy, int iz, double &k2, double &deal) {
    const int N = p.N;
    int64_t n2;
    if(p.dims == 1) {
        n2 = (int64_t)ix * ix;
    }
    else if(p.dims == 2) {
        const int ky = (iy < N / 2) ? iy : N - iy;
        n2 = (int64_t)ky * ky + (int64_t)ix * ix;
    }
    else {
        const int kx = (iz < N / 2) ? iz : iz - N;
        const int ky = (iy < N / 2) ? iy : iy - N;
        n2 = (int64_t)kx * kx + (int64_t)ky * ky + (int64_t)ix * ix;
    }
    k2 = p.q2_unit * (double)n2;
    const int64_t lhs = 9 * n2, rhs = (int64_t)N * N;
    deal = (p.dims == 2 ? lhs < rhs : lhs <= rhs) ? 1.0 : 0.0;
}

// number of 2D wave-kernel region shapes selectable through Params::wave_cfg
constexpr int kWaveCfgs = 9;
// default forward shape, measured on MI355X over shapes 0-8 at 8 steps per
// launch (FP64 and FP32): 64 x 64 regions of 16 waves up to N = 512, where more,
// smaller blocks fill the CUs; 192 x 64 at N = 1024 (224 blocks, one per CU);
// 128 x 80 from N = 2048
__host__ __device__ constexpr int wave_default_cfg(int N) { return N <= 512 ? 4 : (N == 1024 ? 8 : 7); }
// reverse sweep and its forward recompute (shapes 0, 3, 4, 5): 128 x 64 regions
// from N = 512, 64 x 64 below
__host__ __device__ constexpr int adj_default_cfg(int N) { return N >= 512 ? 5 : 4; }

// half-spectrum entry q = ((z N) + y) Nh + kx -> (half-axis index ix, iy, iz)
// as mode_q2 takes them
__device__ __forceinline__ void mode_index(const Params &p, int64_t q, int &ix, int &iy, int &iz) {
    const int Nh = p.N / 2 + 1;
    ix = (int)(q % Nh);
    const int64_t rest = q / Nh;
    if(p.dims == 3) {
        iy = (int)(rest % p.N);
        iz = (int)(rest / p.N);
    }
    else {
        iy = (int)rest;
        iz = 0;
    }
}

// signed wave number of `axis` (0 = slowest) of a half-spectrum entry, in
// units of 2 pi / (N dx): the x (half) axis runs 0 .. N/2, a full axis uses
// i < N/2 ? i : i - N, so its Nyquist entry is -N/2
// (PseudospectralMobilityCPU.cpp:56-125)
__device__ __forceinline__ int axis_wavenumber(const Params &p, int ix, int iy, int iz, int axis) {
    if(axis == p.dims - 1) return ix;
    const int i = (p.dims == 3 && axis == 0) ? iz : iy;
    return i < p.N / 2 ? i : i - p.N;
}

template <typename T> struct Complex;
template <> struct Complex<double> { using type = double2; };
template <> struct Complex<float> { using type = float2; };

}  // namespace chd
