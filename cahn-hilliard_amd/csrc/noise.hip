// noise.hip -- euler_mobility step with conserved noise.
//
// src/integrators/EulerMobilityCPU.cpp:10-18, 55-83: the flux through the +e_d
// face of cell i is
//     J_d(i) = M_f (mu_{i+e_d} - mu_i) / dx + sqrt(max(M_f, 0)) a eta_d(i),
//     M_f = (M_i + M_{i+e_d}) / 2,  a = sqrt(2 / (dx^dims dt)) noise_rescale_factor,
// and psi_i += dt sum_d (J_d(i) - J_d(i - e_d)) / dx, so the noise moves mass
// between neighbours and the total is conserved exactly.
//
// eta_d(i) is the d-th standard normal of cell i at step t: Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, SC 2011) of the counter (i, 0, t lo, t hi)
// under the key (seed lo, seed hi), the four 32-bit words mapped to uniforms
// u = (x + 1/2) 2^-32 and Box-Muller on (u0, u1), (u2, u3) (in FP64 with the
// series logarithm of ch_device.h and sincos_2pi below).  A face's value
// is a pure function of (seed, cell, step), so any tile computes it without
// coordination and a run does not depend on the launch shape.  The
// reference draws from one mt19937 stream (CPU) or per-cell curand states
// (EulerMobilityCUDA.cu:52-71); neither sequence is reproduced.
//
// One launch = one step on a tile + 2-cell halo in LDS:
//   load    psi and M(psi) on the region                       (A, C)
//   mu      on the box [1, R-1)                                (B)
//   flux    J_d = deterministic + stochastic on [1, R-2)       (F_d)
//   update  psi + dt div J on the tile, stored to `out`
// The step index is read from device memory (base + offset), so a captured
// chain of launches replays with fresh noise once the base is advanced.
#include "ch_internal.h"
#include "stencil.h"

namespace chd {

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for(int r = 0; r < 10; r++) {
        if(r > 0) {
            k0 += W0;
            k1 += W1;
        }
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = uint4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// sin(2 pi u) and cos(2 pi u) for u in [0, 1): with 8 u = k + 1/2 + f (k the
// octant, |f| <= 1/2, both exact), 2 pi u = (2k + 1) pi / 8 + r, r = f pi / 4
// (pi / 4 as head + tail), |r| <= pi / 8; sin r and cos r by their Maclaurin
// series through r^13 and r^14 (first omitted terms below 1e-18), then the
// angle sum with sin and cos of (2k + 1) pi / 8, which are +-sin(pi/8) or
// +-cos(pi/8).  Within 2e-16 of sin / cos of the exact 2 pi u, with no
// general argument reduction.
__device__ __forceinline__ void sincos_2pi(double u, double &sn, double &cs) {
#pragma clang fp contract(off)
    const double v = 8.0 * u;
    const int k = (int)v;
    const double f = (v - (double)k) - 0.5;
    const double r = __builtin_fma(f, 0.7853981633974483, f * 3.061616997868383e-17);
    const double t = r * r;
    double s = -1.0 / 6227020800.0;  // -1/13!
    s = __builtin_fma(s, t, 1.0 / 39916800.0);
    s = __builtin_fma(s, t, -1.0 / 362880.0);
    s = __builtin_fma(s, t, 1.0 / 5040.0);
    s = __builtin_fma(s, t, -1.0 / 120.0);
    s = __builtin_fma(s, t, 1.0 / 6.0);
    const double sr = __builtin_fma(-r * t, s, r);  // r - r^3 (1/3! - r^2/5! + ...)
    double c = 1.0 / 87178291200.0;                 // 1/14!
    c = __builtin_fma(c, t, -1.0 / 479001600.0);
    c = __builtin_fma(c, t, 1.0 / 3628800.0);
    c = __builtin_fma(c, t, -1.0 / 40320.0);
    c = __builtin_fma(c, t, 1.0 / 720.0);
    c = __builtin_fma(c, t, -1.0 / 24.0);
    c = __builtin_fma(c, t, 0.5);
    const double cr = __builtin_fma(-t, c, 1.0);  // 1 - r^2 (1/2! - r^2/4! + ...)
    constexpr double S1 = 0.3826834323650898, C1 = 0.9238795325112867;  // sin, cos of pi/8
    const bool swap = k == 1 || k == 2 || k == 5 || k == 6;
    double sk = swap ? C1 : S1, ck = swap ? S1 : C1;
    if(k >= 4) sk = -sk;
    if(k >= 2 && k <= 5) ck = -ck;
    sn = __builtin_fma(sk, cr, ck * sr);
    cs = __builtin_fma(ck, cr, -(sk * sr));
}

// standard normals DIMS of cell `cell` at step `step`, in T
template <typename T, int DIMS>
__device__ __forceinline__ void cell_normals(uint64_t seed, uint32_t cell, uint64_t step, T (&eta)[DIMS]) {
#pragma clang fp contract(off)
    const uint4 x = philox4x32_10(uint4{cell, 0u, (uint32_t)step, (uint32_t)(step >> 32)}, (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for(int h = 0; 2 * h < DIMS; h++) {
        if constexpr(sizeof(T) == 8) {
            const double u1 = ((double)w[2 * h] + 0.5) * 0x1p-32;
            const double u2 = ((double)w[2 * h + 1] + 0.5) * 0x1p-32;
            const double r = sqrt(-2.0 * log_series(u1));
            double sn, cs;
            sincos_2pi(u2, sn, cs);
            eta[2 * h] = r * cs;
            if(2 * h + 1 < DIMS) eta[2 * h + 1] = r * sn;
        }
        else {
            // (float)x keeps 24 significant bits; u <= 1 after rounding, so the
            // logarithm is never positive
            const float u1 = (float)w[2 * h] * 0x1p-32f + 0x1p-33f;
            const float u2 = (float)w[2 * h + 1] * 0x1p-32f + 0x1p-33f;
            const float r = sqrtf(-2.0f * logf(u1)), th = 6.28318530717958647692f * u2;
            float sn, cs;
            sincosf(th, &sn, &cs);
            eta[2 * h] = r * cs;
            if(2 * h + 1 < DIMS) eta[2 * h + 1] = r * sn;
        }
    }
}

template <typename T, int DIMS, int MODEL, int MOB, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_em_noise(const T *__restrict__ in, T *__restrict__ out, Params p, int tiles_x,
                                                 int tiles_y, const int64_t *__restrict__ step_base, int step_off) {
#pragma clang fp contract(off)
    using G = TileGeom<DIMS, TX, TY, TZ, 1>;
    constexpr int RX = G::RX, RY = G::RY, RZ = G::RZ, RXY = G::RXY, R = G::R, H = G::H;
    constexpr bool VARM = MOB != CH_MOB_CONSTANT;
    __shared__ T A[R];
    __shared__ T B[R];
    __shared__ T C[VARM ? R : 1];
    __shared__ T F[DIMS][R];

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % tiles_x;
    const int ty = (tile / tiles_x) % tiles_y;
    const int tz = tile / (tiles_x * tiles_y);
    const int x0 = tx * TX - H, y0 = ty * TY - H, z0 = tz * TZ - H;
    const int N = p.N, mask = p.mask;
    const uint64_t step = (uint64_t)(*step_base + step_off);

    for(int c = threadIdx.x; c < R; c += NT) {
        const int lx = c % RX, ly = (c / RX) % RY, lz = c / RXY;
        const int gx = (x0 + lx) & mask;
        const int gy = DIMS >= 2 ? (y0 + ly) & mask : 0;
        const int gz = DIMS == 3 ? (z0 + lz) & mask : 0;
        const T a = in[((size_t)gz * N + gy) * N + gx];
        A[c] = a;
        if constexpr(VARM) C[c] = mobility<T, MOB>(p, a);
    }
    __syncthreads();

    constexpr int off[3] = {1, RX, RXY};
    // mu on [1, R-1) per axis
    {
        constexpr int EX = RX - 2, EY = DIMS >= 2 ? RY - 2 : 1, EZ = DIMS == 3 ? RZ - 2 : 1;
        for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
            const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
            const int li = (1 + ex) + (DIMS >= 2 ? (1 + ey) * RX : 0) + (DIMS == 3 ? (1 + ez) * RXY : 0);
            T m[DIMS], q[DIMS];
#pragma unroll
            for(int d = 0; d < DIMS; d++) {
                m[d] = A[li - off[d]];
                q[d] = A[li + off[d]];
            }
            B[li] = chem_pot<T, DIMS, MODEL>(p, A[li], m, q);
        }
    }
    __syncthreads();
    // +e_d face fluxes on [1, R-2) per axis (mu of the +e_d neighbour is in [1, R-1))
    {
        constexpr int EX = RX - 3, EY = DIMS >= 2 ? RY - 3 : 1, EZ = DIMS == 3 ? RZ - 3 : 1;
        const T half = (T)0.5, inv_dx = (T)p.inv_dx, amp = (T)p.noise, Mc = (T)p.M;
        for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
            const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
            const int li = (1 + ex) + (DIMS >= 2 ? (1 + ey) * RX : 0) + (DIMS == 3 ? (1 + ez) * RXY : 0);
            const int gx = (x0 + 1 + ex) & mask;
            const int gy = DIMS >= 2 ? (y0 + 1 + ey) & mask : 0;
            const int gz = DIMS == 3 ? (z0 + 1 + ez) & mask : 0;
            T eta[DIMS];
            cell_normals<T, DIMS>(p.seed, (uint32_t)(((size_t)gz * N + gy) * N + gx), step, eta);
            const T mu = B[li];
#pragma unroll
            for(int d = 0; d < DIMS; d++) {
                const T Mf = VARM ? half * (C[li] + C[li + off[d]]) : Mc;
                const T det = Mf * (B[li + off[d]] - mu) * inv_dx;
                const T sto = sqrt(Mf > (T)0 ? Mf : (T)0) * amp * eta[d];
                F[d][li] = det + sto;
            }
        }
    }
    __syncthreads();
    // update on the tile interior [2, R-2), straight to `out`
    {
        constexpr int TY_ = DIMS >= 2 ? TY : 1, TZ_ = DIMS == 3 ? TZ : 1;
        const T inv_dx = (T)p.inv_dx, dt = (T)p.dt;
        for(int c = threadIdx.x; c < TX * TY_ * TZ_; c += NT) {
            const int lx = c % TX, ly = (c / TX) % TY_, lz = c / (TX * TY_);
            const int li = (H + lx) + (DIMS >= 2 ? (H + ly) * RX : 0) + (DIMS == 3 ? (H + lz) * RXY : 0);
            T res = 0;
#pragma unroll
            for(int d = 0; d < DIMS; d++) res = res + (F[d][li] - F[d][li - off[d]]) * inv_dx;
            const int gx = tx * TX + lx;
            const int gy = DIMS >= 2 ? ty * TY + ly : 0;
            const int gz = DIMS == 3 ? tz * TZ + lz : 0;
            out[((size_t)gz * N + gy) * N + gx] = fma_t(res, dt, A[li]);
        }
    }
}

template <typename T, int DIMS, int MODEL, int MOB, int TX, int TY, int TZ, int NT = 256>
hipError_t launch_noise_cfg(const Params &p, const void *in, void *out, const int64_t *step_base, int step_off,
                            hipStream_t s) {
    const int tiles_x = p.N / TX;
    const int tiles_y = DIMS >= 2 ? p.N / TY : 1;
    const int tiles_z = DIMS == 3 ? p.N / TZ : 1;
    hipLaunchKernelGGL((k_em_noise<T, DIMS, MODEL, MOB, TX, TY, TZ, NT>), dim3(tiles_x * tiles_y * tiles_z), dim3(NT),
                       0, s, (const T *)in, (T *)out, p, tiles_x, tiles_y, step_base, step_off);
    return hipGetLastError();
}

// tiles: 1D 256 | 4, 2D 64 x 32 | 4 x 4, 3D 16 x 8 x 8 | 8 x 8 x 8 | 4 x 4 x 4 (the largest N allows;
// Params::noise_tile selects the earlier 64 x 16 / 8 x 8 x 8 shapes)
template <typename T, int DIMS, int MODEL, int MOB>
hipError_t noise_tiles(const Params &p, const void *in, void *out, const int64_t *step_base, int step_off,
                       hipStream_t s) {
    if constexpr(DIMS == 1) {
        if(p.N >= 256) return launch_noise_cfg<T, 1, MODEL, MOB, 256, 1, 1>(p, in, out, step_base, step_off, s);
        return launch_noise_cfg<T, 1, MODEL, MOB, 4, 1, 1>(p, in, out, step_base, step_off, s);
    }
    else if constexpr(DIMS == 2) {
        if(p.N >= 64) {
            if(p.noise_tile == 1) return launch_noise_cfg<T, 2, MODEL, MOB, 64, 16, 1, 512>(p, in, out, step_base, step_off, s);
            if(p.noise_tile == 2) return launch_noise_cfg<T, 2, MODEL, MOB, 64, 32, 1, 512>(p, in, out, step_base, step_off, s);
            return launch_noise_cfg<T, 2, MODEL, MOB, 64, 16, 1>(p, in, out, step_base, step_off, s);
        }
        return launch_noise_cfg<T, 2, MODEL, MOB, 4, 4, 1>(p, in, out, step_base, step_off, s);
    }
    else {
        if(p.N >= 16) {
            if(p.noise_tile == 1) return launch_noise_cfg<T, 3, MODEL, MOB, 8, 8, 8, 512>(p, in, out, step_base, step_off, s);
            if(p.noise_tile == 2) return launch_noise_cfg<T, 3, MODEL, MOB, 16, 8, 8, 512>(p, in, out, step_base, step_off, s);
        }
        if(p.N >= 8) return launch_noise_cfg<T, 3, MODEL, MOB, 8, 8, 8>(p, in, out, step_base, step_off, s);
        return launch_noise_cfg<T, 3, MODEL, MOB, 4, 4, 4>(p, in, out, step_base, step_off, s);
    }
}

template <typename T, int DIMS, int MODEL>
hipError_t noise_mob(const Params &p, int mob, const void *in, void *out, const int64_t *step_base, int step_off,
                     hipStream_t s) {
    switch(mob) {
    case CH_MOB_CONSTANT: return noise_tiles<T, DIMS, MODEL, CH_MOB_CONSTANT>(p, in, out, step_base, step_off, s);
    case CH_MOB_REGULARISED: return noise_tiles<T, DIMS, MODEL, CH_MOB_REGULARISED>(p, in, out, step_base, step_off, s);
    case CH_MOB_LINEAR: return noise_tiles<T, DIMS, MODEL, CH_MOB_LINEAR>(p, in, out, step_base, step_off, s);
    case CH_MOB_FREE_ENERGY:
        if constexpr(MODEL == CH_SIMPLE_WERTHEIM)
            return noise_tiles<T, DIMS, MODEL, CH_MOB_FREE_ENERGY>(p, in, out, step_base, step_off, s);
        return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template <typename T, int DIMS>
hipError_t noise_model(const Params &p, int model, int mob, const void *in, void *out, const int64_t *step_base,
                       int step_off, hipStream_t s) {
    if(model == CH_LANDAU) return noise_mob<T, DIMS, CH_LANDAU>(p, mob, in, out, step_base, step_off, s);
    return noise_mob<T, DIMS, CH_SIMPLE_WERTHEIM>(p, mob, in, out, step_base, step_off, s);
}

template <typename T>
hipError_t noise_dims(const Params &p, int model, int mob, const void *in, void *out, const int64_t *step_base,
                      int step_off, hipStream_t s) {
    switch(p.dims) {
    case 1: return noise_model<T, 1>(p, model, mob, in, out, step_base, step_off, s);
    case 2: return noise_model<T, 2>(p, model, mob, in, out, step_base, step_off, s);
    default: return noise_model<T, 3>(p, model, mob, in, out, step_base, step_off, s);
    }
}

hipError_t em_noise_step(const Params &p, int precision, int model, int mob, const void *in, void *out,
                         const int64_t *step_base, int step_off, hipStream_t s) {
    if(precision == CH_FP64) return noise_dims<double>(p, model, mob, in, out, step_base, step_off, s);
    return noise_dims<float>(p, model, mob, in, out, step_base, step_off, s);
}

__global__ void k_step_advance(int64_t *step_base, int n) {
    if(threadIdx.x == 0 && blockIdx.x == 0) *step_base += n;
}

hipError_t noise_step_advance(int64_t *step_base, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, step_base, n);
    return hipGetLastError();
}

}  // namespace chd
