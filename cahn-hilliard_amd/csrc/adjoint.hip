// adjoint.hip -- reverse-mode sweep of the euler and euler_mobility integrators.
//
// euler (src/integrators/EulerCPU.cpp:18-36), constant mobility:
//     psi_{n+1} = psi_n + dt M L(mu_n),  mu_n = f'(psi_n) - 2k L psi_n
// with L the periodic (2d+1)-point Laplacian (symmetric).  Its transpose gives
//     lambda_n = lambda_{n+1} + dt M ( f''(psi_n) w - 2k L w ),  w = L lambda_{n+1}
// euler_mobility (src/integrators/EulerMobilityCPU.cpp:31-83), any mobility rule:
//     psi_{n+1} = psi_n + dt D(Mob_n, mu_n),  Mob_n = mobility(psi_n),
//     D(a, u)_i = sum_d [A(i,i+e)(u_{i+e} - u_i) - A(i-e,i)(u_i - u_{i-e})] / dx^2,
//     A(i,j) = (a_i + a_j) / 2 (the face mobility of the flux form).
// <lambda, D(Mob, mu)> is bilinear in (Mob, mu) and D(a, .) is symmetric, so
//     nu = D(Mob, lambda_{n+1}),
//     eta_j = -sum_d [(mu_{j+e}-mu_j)(l_{j+e}-l_j) + (mu_j-mu_{j-e})(l_j-l_{j-e})] / (2 dx^2),
//     lambda_n = lambda_{n+1} + dt ( f''(psi_n) nu - 2k L nu + Mob'(psi_n) eta ).
// Parameter gradients of the loss, summed over steps (nu = M w for euler):
//     dL/deps = -dt <nu, psi_n>        (Landau: d mu/d eps = -psi)
//     dL/dk   = -2 dt <nu, L psi_n>
//     dL/dM   = dt <eta, Mob_n / M>    (euler: dt <w, mu_n>)
//     dL/ddelta = dt (<nu, d f'/d delta> + <eta, d Mob/d delta>)   (SimpleWertheim)
//     dL/dB2  = 2 dt <nu, psi_n>       (SimpleWertheim: d f'/d B2 = 2 rho)
// One euler launch runs KA backward steps on a tile carrying a 2 KA halo of
// lambda; euler_mobility runs one step per launch.  The states psi_n come from
// the tape of the current checkpoint window.  Each block keeps FP64 partial
// sums of the inner products for its own tile (fixed tile -> slot mapping, no
// atomics: deterministic).  This reverse sweep has no reference counterpart;
// it is pinned by the oracle's FP64 adjoint and by finite differences of the
// forward map.
#include <algorithm>
#include <type_traits>

#include "adjoint_wave.h"
#include "ch_internal.h"
#include "stencil_dispatch.h"

namespace chd {

// slots: <nu, psi>, <nu, L psi>, the M term, the delta term (nu = w for euler;
// the host applies the factors dt, M)
template <typename T, int DIMS, int MODEL, int TX, int TY, int TZ, int KA, int NT>
__global__ __launch_bounds__(NT) void k_adjoint(const T *__restrict__ lam_in, T *__restrict__ lam_out,
                                                const T *__restrict__ tape, int s_top, Params p, int tiles_x,
                                                int tiles_y, double *__restrict__ gacc) {
    using G = TileGeom<DIMS, TX, TY, TZ, KA>;
    constexpr int RX = G::RX, RY = G::RY, RZ = G::RZ, RXY = G::RXY, R = G::R, H = G::H;
    constexpr int off[3] = {1, RX, RXY};
    __shared__ T Lm[R];   // lambda
    __shared__ T Wb[R];   // w = L lambda
    __shared__ T P[R];    // psi_n
    __shared__ double red[kGradSlots][NT / 64];

    const int tile = blockIdx.x;
    const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
    const int x0 = tx * TX - H, y0 = ty * TY - H, z0 = tz * TZ - H;
    const int N = p.N, mask = p.mask;
    const size_t bins = (size_t)1 << (p.log2N * DIMS);

    auto gidx = [&](int c) {
        const int lx = c % RX, ly = (c / RX) % RY, lz = c / RXY;
        const int gx = (x0 + lx) & mask;
        const int gy = DIMS >= 2 ? (y0 + ly) & mask : 0;
        const int gz = DIMS == 3 ? (z0 + lz) & mask : 0;
        return ((size_t)gz * N + gy) * N + gx;
    };
    for(int c = threadIdx.x; c < R; c += NT) Lm[c] = lam_in[gidx(c)];

    const T inv_dx2 = (T)p.inv_dx2, two_k = (T)p.two_k, Mdt = (T)p.Mdt;
    double ge = 0, gk = 0, gm = 0, gd = 0;
#pragma unroll 1
    for(int j = 0; j < KA; j++) {
        const T *psi = tape + (size_t)(s_top - j) * bins;
        for(int c = threadIdx.x; c < R; c += NT) P[c] = psi[gidx(c)];
        __syncthreads();
        {   // w = L lambda on [2j+1, R-2j-1)
            const int lo = 2 * j + 1;
            const int EX = RX - 2 * lo, EY = DIMS >= 2 ? RY - 2 * lo : 1, EZ = DIMS == 3 ? RZ - 2 * lo : 1;
            for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
                const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
                const int li = (lo + ex) + (DIMS >= 2 ? (lo + ey) * RX : 0) + (DIMS == 3 ? (lo + ez) * RXY : 0);
                T m[DIMS], q[DIMS];
#pragma unroll
                for(int d = 0; d < DIMS; d++) { m[d] = Lm[li - off[d]]; q[d] = Lm[li + off[d]]; }
                Wb[li] = laplacian<T, DIMS>(Lm[li], m, q, inv_dx2);
            }
        }
        __syncthreads();
        {   // gradient terms on the tile interior (each cell counted once)
            constexpr int TY_ = DIMS >= 2 ? TY : 1, TZ_ = DIMS == 3 ? TZ : 1;
            for(int c = threadIdx.x; c < TX * TY_ * TZ_; c += NT) {
                const int lx = c % TX, ly = (c / TX) % TY_, lz = c / (TX * TY_);
                const int li = (H + lx) + (DIMS >= 2 ? (H + ly) * RX : 0) + (DIMS == 3 ? (H + lz) * RXY : 0);
                T m[DIMS], q[DIMS];
#pragma unroll
                for(int d = 0; d < DIMS; d++) { m[d] = P[li - off[d]]; q[d] = P[li + off[d]]; }
                const double w = (double)Wb[li], ps = (double)P[li];
                const double lps = (double)laplacian<T, DIMS>(P[li], m, q, inv_dx2);
                const double mu = (double)fprime<T, MODEL>(p, P[li]) - p.two_k * lps;
                ge += w * ps;
                gk += w * lps;
                gm += w * mu;
                if constexpr(MODEL == CH_SIMPLE_WERTHEIM) gd += w * dfprime_ddelta<MODEL>(p, ps);
            }
        }
        {   // lambda += dt M (f''(psi) w - 2k L w) on [2j+2, R-2j-2)
            const int lo = 2 * j + 2;
            const int EX = RX - 2 * lo, EY = DIMS >= 2 ? RY - 2 * lo : 1, EZ = DIMS == 3 ? RZ - 2 * lo : 1;
            for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
                const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
                const int li = (lo + ex) + (DIMS >= 2 ? (lo + ey) * RX : 0) + (DIMS == 3 ? (lo + ez) * RXY : 0);
                T m[DIMS], q[DIMS];
#pragma unroll
                for(int d = 0; d < DIMS; d++) { m[d] = Wb[li - off[d]]; q[d] = Wb[li + off[d]]; }
                const T w = Wb[li];
                const T lw = laplacian<T, DIMS>(w, m, q, inv_dx2);
                const T fpp = (T)fsecond<MODEL>(p, (double)P[li]);
                Lm[li] = Lm[li] + Mdt * (fpp * w - two_k * lw);
            }
        }
        __syncthreads();
    }

    constexpr int TY_ = DIMS >= 2 ? TY : 1, TZ_ = DIMS == 3 ? TZ : 1;
    for(int c = threadIdx.x; c < TX * TY_ * TZ_; c += NT) {
        const int lx = c % TX, ly = (c / TX) % TY_, lz = c / (TX * TY_);
        const int li = (H + lx) + (DIMS >= 2 ? (H + ly) * RX : 0) + (DIMS == 3 ? (H + lz) * RXY : 0);
        const int gx = tx * TX + lx, gy = DIMS >= 2 ? ty * TY + ly : 0, gz = DIMS == 3 ? tz * TZ + lz : 0;
        lam_out[((size_t)gz * N + gy) * N + gx] = Lm[li];
    }

    const double g[kGradSlots] = {ge, gk, gm, gd};
    block_accumulate<NT>(g, red, gacc + kGradSlots * tile);
}

// one euler_mobility backward step: lambda_in = lambda_{n+1} -> lambda_out =
// lambda_n, psi = psi_n; the tile carries a 2-cell halo
template <typename T, int DIMS, int MODEL, int MOB, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_adjoint_em(const T *__restrict__ lam_in, T *__restrict__ lam_out,
                                                   const T *__restrict__ psi, Params p, int tiles_x, int tiles_y,
                                                   double *__restrict__ gacc) {
    using G = TileGeom<DIMS, TX, TY, TZ, 1>;
    constexpr int RX = G::RX, RY = G::RY, RZ = G::RZ, RXY = G::RXY, R = G::R, H = G::H;
    constexpr int off[3] = {1, RX, RXY};
    constexpr bool VARM = MOB != CH_MOB_CONSTANT;
    __shared__ T Lm[R];            // lambda_{n+1}
    __shared__ T P[R];             // psi_n
    __shared__ T Mu[R];            // mu_n
    __shared__ T Nu[R];            // nu = D(Mob, lambda)
    __shared__ T Mb[VARM ? R : 1];  // Mob_n
    __shared__ double red[kGradSlots][NT / 64];

    const int tile = blockIdx.x;
    const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
    const int x0 = tx * TX - H, y0 = ty * TY - H, z0 = tz * TZ - H;
    const int N = p.N, mask = p.mask;
    for(int c = threadIdx.x; c < R; c += NT) {
        const int lx = c % RX, ly = (c / RX) % RY, lz = c / RXY;
        const int gx = (x0 + lx) & mask;
        const int gy = DIMS >= 2 ? (y0 + ly) & mask : 0;
        const int gz = DIMS == 3 ? (z0 + lz) & mask : 0;
        const size_t g = ((size_t)gz * N + gy) * N + gx;
        Lm[c] = lam_in[g];
        const T ps = psi[g];
        P[c] = ps;
        if constexpr(VARM) Mb[c] = mobility<T, MOB>(p, ps);
    }
    __syncthreads();
    const T inv_dx2 = (T)p.inv_dx2, Mc = (T)p.M;
    auto cell = [&](int lo, int c) {
        const int EX = RX - 2 * lo, EY = DIMS >= 2 ? RY - 2 * lo : 1;
        const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
        return (lo + ex) + (DIMS >= 2 ? (lo + ey) * RX : 0) + (DIMS == 3 ? (lo + ez) * RXY : 0);
    };
    auto box = [&](int lo) {
        return (RX - 2 * lo) * (DIMS >= 2 ? RY - 2 * lo : 1) * (DIMS == 3 ? RZ - 2 * lo : 1);
    };
    for(int c = threadIdx.x; c < box(1); c += NT) {  // mu and nu on [1, R-1)
        const int li = cell(1, c);
        T m[DIMS], q[DIMS];
#pragma unroll
        for(int d = 0; d < DIMS; d++) { m[d] = P[li - off[d]]; q[d] = P[li + off[d]]; }
        Mu[li] = chem_pot<T, DIMS, MODEL>(p, P[li], m, q);
        const T l = Lm[li];
        const T Mi = VARM ? Mb[li] : Mc;
        T lm[DIMS], lq[DIMS], Mm[DIMS], Mq[DIMS];
#pragma unroll
        for(int d = 0; d < DIMS; d++) {
            Mm[d] = VARM ? Mb[li - off[d]] : Mc;
            Mq[d] = VARM ? Mb[li + off[d]] : Mc;
            lm[d] = Lm[li - off[d]];
            lq[d] = Lm[li + off[d]];
        }
        Nu[li] = adj_nu<T, DIMS>(l, Mi, lm, lq, Mm, Mq, inv_dx2);
    }
    __syncthreads();
    double ge = 0, gk = 0, gm = 0, gd = 0;
    for(int c = threadIdx.x; c < box(2); c += NT) {  // the tile interior [2, R-2)
        const int li = cell(2, c);
        T nm[DIMS], nq[DIMS], pm[DIMS], pq[DIMS];
#pragma unroll
        for(int d = 0; d < DIMS; d++) {
            nm[d] = Nu[li - off[d]]; nq[d] = Nu[li + off[d]];
            pm[d] = P[li - off[d]]; pq[d] = P[li + off[d]];
        }
        const T nu = Nu[li], l = Lm[li], mu = Mu[li];
        T um[DIMS], uq[DIMS], lm[DIMS], lq[DIMS];
#pragma unroll
        for(int d = 0; d < DIMS; d++) {
            um[d] = Mu[li - off[d]];
            uq[d] = Mu[li + off[d]];
            lm[d] = Lm[li - off[d]];
            lq[d] = Lm[li + off[d]];
        }
        const T eta = adj_eta<T, DIMS>(mu, l, um, uq, lm, lq, inv_dx2);
        const T lnu = laplacian<T, DIMS>(nu, nm, nq, inv_dx2);
        const double ps = (double)P[li];
        const MobDeriv<MOB> md = mob_deriv<MOB>(p, ps);
        const int lx = li % RX, ly = (li / RX) % RY, lz = li / RXY;
        const int gx = tx * TX + lx - H, gy = DIMS >= 2 ? ty * TY + ly - H : 0, gz = DIMS == 3 ? tz * TZ + lz - H : 0;
        lam_out[((size_t)gz * N + gy) * N + gx] =
            adj_em_step<T>(p, l, (T)fsecond<MODEL>(p, ps), nu, lnu, (T)md.drho, eta);
        const double lps = (double)laplacian<T, DIMS>(P[li], pm, pq, inv_dx2);
        ge += (double)nu * ps;
        gk += (double)nu * lps;
        gm += (double)eta * md.hat;
        if constexpr(MODEL == CH_SIMPLE_WERTHEIM) gd += (double)nu * dfprime_ddelta<MODEL>(p, ps) + (double)eta * md.ddelta;
    }
    const double g[kGradSlots] = {ge, gk, gm, gd};
    block_accumulate<NT>(g, red, gacc + kGradSlots * tile);
}

__global__ void k_fill_zero(double *__restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) a[i] = 0.0;
}

// F = V_bin / bins sum_i [f(rho_i) + k sum_d ((rho_{i+e} - rho_i) / dx)^2]
// (src/CahnHilliard.cpp:300-319): dF/drho_i = V_bin / bins (f'(rho_i) - 2k L rho_i),
// plus the sums behind the explicit parameter derivatives.  FP64 throughout.
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_fe_grad(const T *__restrict__ rho, Params p, double *__restrict__ grad,
                                                 double *__restrict__ partials) {
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int mask = p.mask, lg = p.log2N;
    double acc[kFeSlots] = {0, 0, 0, 0, 0};
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < bins; i += (int64_t)gridDim.x * 256) {
        const double r = (double)rho[i];
        double nsum = 0, g2 = 0;
        for(int d = 0; d < p.dims; d++) {
            const int sh = d * lg;
            const int c = (int)((i >> sh) & mask);
            const int64_t base = i - ((int64_t)c << sh);
            const double rm = (double)rho[base + ((int64_t)((c - 1) & mask) << sh)];
            const double rp = (double)rho[base + ((int64_t)((c + 1) & mask) << sh)];
            nsum += rm + rp;
            const double g = (rp - r) * p.inv_dx;
            g2 += g * g;
        }
        const double lap = (nsum - 2.0 * p.dims * r) * p.inv_dx2;
        grad[i] = p.vbin_over_bins * ((double)fprime<double, MODEL>(p, r) - p.two_k * lap);
        acc[0] += bulk<MODEL>(p, r);
        acc[1] += g2;
        acc[2] += -0.5 * r * r;
        acc[3] += dbulk_ddelta<MODEL>(p, r);
        acc[4] += r * r;
    }
    __shared__ double red[kFeSlots][4];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for(int j = 0; j < kFeSlots; j++) {
        const double v = wave_sum64(acc[j]);
        if((threadIdx.x & 63) == 0) red[j][w] = v;
    }
    __syncthreads();
    if(threadIdx.x < kFeSlots)
        partials[kFeSlots * blockIdx.x + threadIdx.x] =
            (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

hipError_t free_energy_grad(const Params &p, int precision, int model, const void *rho, double *grad,
                            double *partials, int max_parts, int *nb, hipStream_t s) {
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int n = (int)std::min<int64_t>(max_parts, (bins + 255) / 256);
    *nb = n;
    auto go = [&](auto tag) {
        using T = decltype(tag);
        if(model == CH_LANDAU)
            hipLaunchKernelGGL((k_fe_grad<T, CH_LANDAU>), dim3(n), dim3(256), 0, s, (const T *)rho, p, grad, partials);
        else
            hipLaunchKernelGGL((k_fe_grad<T, CH_SIMPLE_WERTHEIM>), dim3(n), dim3(256), 0, s, (const T *)rho, p, grad,
                               partials);
    };
    if(precision == CH_FP64) go(double{});
    else go(float{});
    return hipGetLastError();
}

// ---------------------------------------------------------------- host side
template <typename T, int DIMS, int MODEL, int TX, int TY, int TZ, int KA>
static hipError_t launch_adj(const Params &p, const void *lam_in, void *lam_out, const void *tape, int s_top,
                             double *gacc, hipStream_t s) {
    constexpr int NT = 256;
    const int tiles_x = p.N / TX, tiles_y = DIMS >= 2 ? p.N / TY : 1, tiles_z = DIMS == 3 ? p.N / TZ : 1;
    hipLaunchKernelGGL((k_adjoint<T, DIMS, MODEL, TX, TY, TZ, KA, NT>), dim3(tiles_x * tiles_y * tiles_z), dim3(NT), 0, s,
                       (const T *)lam_in, (T *)lam_out, (const T *)tape, s_top, p, tiles_x, tiles_y, gacc);
    return hipGetLastError();
}

template <typename T, int DIMS, int MODEL>
static hipError_t adj_tiles(const Params &p, int KA, const void *li, void *lo, const void *tape, int s_top, double *gacc,
                            hipStream_t s) {
    const int N = p.N;
    if constexpr(DIMS == 1) {
        if(KA == 4 && N >= 64) return launch_adj<T, 1, MODEL, 64, 1, 1, 4>(p, li, lo, tape, s_top, gacc, s);
        if(N >= 64) return launch_adj<T, 1, MODEL, 64, 1, 1, 1>(p, li, lo, tape, s_top, gacc, s);
        if(N >= 16) return launch_adj<T, 1, MODEL, 16, 1, 1, 1>(p, li, lo, tape, s_top, gacc, s);
        return launch_adj<T, 1, MODEL, 4, 1, 1, 1>(p, li, lo, tape, s_top, gacc, s);
    }
    else if constexpr(DIMS == 2) {
        if(KA == 4 && N >= 64) return launch_adj<T, 2, MODEL, 64, 16, 1, 4>(p, li, lo, tape, s_top, gacc, s);
        if(N >= 64) return launch_adj<T, 2, MODEL, 64, 16, 1, 1>(p, li, lo, tape, s_top, gacc, s);
        if(N >= 16) return launch_adj<T, 2, MODEL, 16, 16, 1, 1>(p, li, lo, tape, s_top, gacc, s);
        return launch_adj<T, 2, MODEL, 4, 4, 1, 1>(p, li, lo, tape, s_top, gacc, s);
    }
    else {
        if(N >= 16) return launch_adj<T, 3, MODEL, 16, 8, 8, 1>(p, li, lo, tape, s_top, gacc, s);
        return launch_adj<T, 3, MODEL, 4, 4, 4, 1>(p, li, lo, tape, s_top, gacc, s);
    }
}

template <typename T, int DIMS, int MODEL, int MOB, int TX, int TY, int TZ, int NT = 256>
static hipError_t launch_adj_em(const Params &p, const void *lam_in, void *lam_out, const void *psi, double *gacc,
                                hipStream_t s) {
    const int tiles_x = p.N / TX, tiles_y = DIMS >= 2 ? p.N / TY : 1, tiles_z = DIMS == 3 ? p.N / TZ : 1;
    hipLaunchKernelGGL((k_adjoint_em<T, DIMS, MODEL, MOB, TX, TY, TZ, NT>), dim3(tiles_x * tiles_y * tiles_z), dim3(NT),
                       0, s, (const T *)lam_in, (T *)lam_out, (const T *)psi, p, tiles_x, tiles_y, gacc);
    return hipGetLastError();
}

// euler_mobility adjoint tiles: one step per launch, LDS of five tile fields
template <typename T, int DIMS, int MODEL, int MOB>
static hipError_t adj_em_tiles(const Params &p, const void *li, void *lo, const void *psi, double *gacc,
                               hipStream_t s) {
    const int N = p.N;
    if constexpr(DIMS == 1) {
        if(N >= 64) return launch_adj_em<T, 1, MODEL, MOB, 64, 1, 1>(p, li, lo, psi, gacc, s);
        if(N >= 16) return launch_adj_em<T, 1, MODEL, MOB, 16, 1, 1>(p, li, lo, psi, gacc, s);
        return launch_adj_em<T, 1, MODEL, MOB, 4, 1, 1>(p, li, lo, psi, gacc, s);
    }
    else if constexpr(DIMS == 2) {
        if(N >= 64) return launch_adj_em<T, 2, MODEL, MOB, 64, 16, 1>(p, li, lo, psi, gacc, s);
        if(N >= 16) return launch_adj_em<T, 2, MODEL, MOB, 16, 16, 1>(p, li, lo, psi, gacc, s);
        return launch_adj_em<T, 2, MODEL, MOB, 4, 4, 1>(p, li, lo, psi, gacc, s);
    }
    else {
        // Params::adj_tile3d: 16x4x4 / 256 threads, 16x8x4 / 512, 16x8x8 / 512 (gradient slots are
        // sized for the first, the tile count of the others is smaller)
        if(N >= 16) {
            if(p.adj_tile3d == 1) return launch_adj_em<T, 3, MODEL, MOB, 16, 8, 4, 512>(p, li, lo, psi, gacc, s);
            if(p.adj_tile3d == 2) return launch_adj_em<T, 3, MODEL, MOB, 16, 8, 8, 512>(p, li, lo, psi, gacc, s);
            return launch_adj_em<T, 3, MODEL, MOB, 16, 4, 4>(p, li, lo, psi, gacc, s);
        }
        return launch_adj_em<T, 3, MODEL, MOB, 4, 4, 4>(p, li, lo, psi, gacc, s);
    }
}

// 2D euler with N >= 64: the register/DPP adjoint on the forward's default
// region shape for this N
template <class CFG>
static int wave_adj_tiles(int N, int KA) {
    if(CFG::RW == CFG::RH && N == CFG::RW) return 1;
    const int W = CFG::RW - 4 * KA, TH = CFG::RH - 4 * KA;
    return ((N + W - 1) / W) * ((N + TH - 1) / TH);
}

template <typename T, int MODEL, int KA, class CFG>
static hipError_t launch_adj_wave(const Params &p, const void *li, void *lo, const void *tape, int s_top, double *gacc,
                                  hipStream_t s) {
    constexpr int NT = 64 * CFG::NW;
    if(CFG::RW == CFG::RH && p.N == CFG::RW) {
        hipLaunchKernelGGL((k_adjoint2d_wave<T, MODEL, KA, CFG, true>), dim3(1), dim3(NT), 0, s, (const T *)li, (T *)lo,
                           (const T *)tape, s_top, p, 1, gacc);
        return hipGetLastError();
    }
    constexpr int W = CFG::RW - 4 * KA;
    const int tiles_x = (p.N + W - 1) / W;
    hipLaunchKernelGGL((k_adjoint2d_wave<T, MODEL, KA, CFG, false>), dim3(wave_adj_tiles<CFG>(p.N, KA)), dim3(NT), 0, s,
                       (const T *)li, (T *)lo, (const T *)tape, s_top, p, tiles_x, gacc);
    return hipGetLastError();
}

template <typename T, int MODEL, class CFG>
static hipError_t adj_wave_cfg(const Params &p, int KA, const void *li, void *lo, const void *tape, int s_top,
                               double *gacc, hipStream_t s) {
    switch(KA) {
    case 1: return launch_adj_wave<T, MODEL, 1, CFG>(p, li, lo, tape, s_top, gacc, s);
    case 2: return launch_adj_wave<T, MODEL, 2, CFG>(p, li, lo, tape, s_top, gacc, s);
    case 4: return launch_adj_wave<T, MODEL, 4, CFG>(p, li, lo, tape, s_top, gacc, s);
    case 8: return launch_adj_wave<T, MODEL, 8, CFG>(p, li, lo, tape, s_top, gacc, s);
    }
    return hipErrorInvalidValue;
}

template <typename T, int MODEL>
static hipError_t adj_wave(const Params &p, int KA, const void *li, void *lo, const void *tape, int s_top,
                           double *gacc, hipStream_t s) {
    switch(p.adj_cfg) {
    case 0: return adj_wave_cfg<T, MODEL, WaveCfg0>(p, KA, li, lo, tape, s_top, gacc, s);
    case 3: return adj_wave_cfg<T, MODEL, WaveCfg3>(p, KA, li, lo, tape, s_top, gacc, s);
    case 5: return adj_wave_cfg<T, MODEL, WaveCfg5>(p, KA, li, lo, tape, s_top, gacc, s);
    default: return adj_wave_cfg<T, MODEL, WaveCfg4>(p, KA, li, lo, tape, s_top, gacc, s);
    }
}

static bool adj_is_wave(int dims, int N, bool em) { return dims == 2 && !em && N >= 64; }
static bool tape_is_wave(int dims, int N) { return dims == 2 && N >= 64; }

// backward steps of the next adjoint launch when `remaining` are left: the
// wave kernel has depths 8, 4, 2, 1; the LDS kernel 4 (1D / 2D euler, N >= 64)
// and 1
int adjoint_depth(int dims, int N, bool em, int64_t remaining) {
    if(adj_is_wave(dims, N, em)) {
        int k = 8;
        while(k > remaining) k /= 2;
        return k;
    }
    return (!em && dims <= 2 && N >= 64 && remaining >= 4) ? 4 : 1;
}

// states of the next recompute launch when `remaining` are left: wave kernel
// 8, 4, 2, 1; the LDS kernel 4 (1D euler, N >= 64) and 1
int tape_depth(int dims, int N, bool em, int64_t remaining) {
    if(tape_is_wave(dims, N)) {
        int k = 8;
        while(k > remaining) k /= 2;
        return k;
    }
    return (dims == 1 && !em && N >= 64 && remaining >= 4) ? 4 : 1;
}

// number of tiles (gradient accumulator slots) of the adjoint launches: the
// most any launch depth of this shape uses
int adjoint_tiles(const Params &p, bool em) {
    const int dims = p.dims, N = p.N;
    if(dims == 1) return N >= 64 ? N / 64 : (N >= 16 ? N / 16 : N / 4);
    if(adj_is_wave(dims, N, em)) {
        int m = 0;
        for(int k = 1; k <= 8; k *= 2) {
            int t;
            switch(p.adj_cfg) {
            case 0: t = wave_adj_tiles<WaveCfg0>(N, k); break;
            case 3: t = wave_adj_tiles<WaveCfg3>(N, k); break;
            case 5: t = wave_adj_tiles<WaveCfg5>(N, k); break;
            default: t = wave_adj_tiles<WaveCfg4>(N, k); break;
            }
            m = std::max(m, t);
        }
        return m;
    }
    if(dims == 2) return N >= 64 ? (N / 64) * (N / 16) : (N >= 16 ? (N / 16) * (N / 16) : (N / 4) * (N / 4));
    if(em) return N >= 16 ? (N / 16) * (N / 4) * (N / 4) : (N / 4) * (N / 4) * (N / 4);
    return N >= 16 ? (N / 16) * (N / 8) * (N / 8) : (N / 4) * (N / 4) * (N / 4);
}

// dims x model (x mobility rule) dispatch; free_energy mobility only exists for
// SimpleWertheim (the descriptor check rejects it for Landau)
template <typename T, typename F>
static hipError_t by_dims_model_mob(int dims, int model, int mob, F &&f) {
    auto with_model = [&](auto dtag, auto mtag) -> hipError_t {
        constexpr int D = decltype(dtag)::value, MO = decltype(mtag)::value;
        switch(mob) {
        case CH_MOB_CONSTANT: return f(dtag, mtag, std::integral_constant<int, CH_MOB_CONSTANT>{});
        case CH_MOB_REGULARISED: return f(dtag, mtag, std::integral_constant<int, CH_MOB_REGULARISED>{});
        case CH_MOB_LINEAR: return f(dtag, mtag, std::integral_constant<int, CH_MOB_LINEAR>{});
        default:
            if constexpr(MO == CH_SIMPLE_WERTHEIM) return f(dtag, mtag, std::integral_constant<int, CH_MOB_FREE_ENERGY>{});
            (void)D;
            return hipErrorInvalidValue;
        }
    };
    auto with_dims = [&](auto mtag) -> hipError_t {
        if(dims == 1) return with_model(std::integral_constant<int, 1>{}, mtag);
        if(dims == 2) return with_model(std::integral_constant<int, 2>{}, mtag);
        return with_model(std::integral_constant<int, 3>{}, mtag);
    };
    if(model == CH_LANDAU) return with_dims(std::integral_constant<int, CH_LANDAU>{});
    return with_dims(std::integral_constant<int, CH_SIMPLE_WERTHEIM>{});
}

hipError_t adjoint_launch(const Params &p, int precision, int model, int mob, bool em, int KA, const void *lam_in,
                          void *lam_out, const void *tape, int s_top, double *gacc, hipStream_t s) {
    const size_t bins = (size_t)1 << (p.log2N * p.dims);
    auto go = [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        if(em) {
            const T *psi = (const T *)tape + (size_t)s_top * bins;
            return by_dims_model_mob<T>(p.dims, model, mob, [&](auto dt, auto mt, auto bt) {
                return adj_em_tiles<T, decltype(dt)::value, decltype(mt)::value, decltype(bt)::value>(
                    p, lam_in, lam_out, psi, gacc, s);
            });
        }
        if(adj_is_wave(p.dims, p.N, em)) {
            if(model == CH_LANDAU) return adj_wave<T, CH_LANDAU>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
            return adj_wave<T, CH_SIMPLE_WERTHEIM>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
        }
        if(KA != 1 && !(KA == 4 && p.dims <= 2 && p.N >= 64)) return hipErrorInvalidValue;
        return by_dims_model_mob<T>(p.dims, model, CH_MOB_CONSTANT, [&](auto dt, auto mt, auto) {
            return adj_tiles<T, decltype(dt)::value, decltype(mt)::value>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
        });
    };
    if(precision == CH_FP64) return go(double{});
    return go(float{});
}

// recompute of a checkpoint window: K steps from `in`, all K states to tape_out
template <typename T, int DIMS, int MODEL, int MOB, bool EM, int TX, int TY, int TZ, int K, int NT = 256>
static hipError_t launch_tape(const Params &p, const void *in, void *tape_out, hipStream_t s) {
    const int tiles_x = p.N / TX, tiles_y = DIMS >= 2 ? p.N / TY : 1, tiles_z = DIMS == 3 ? p.N / TZ : 1;
    hipLaunchKernelGGL((k_stencil<T, DIMS, MODEL, MOB, EM, TX, TY, TZ, K, NT, true>), dim3(tiles_x * tiles_y * tiles_z),
                       dim3(NT), 0, s, (const T *)in, (T *)tape_out, p, tiles_x, tiles_y);
    return hipGetLastError();
}

template <typename T, int MODEL, int MOB, bool EM, class CFG>
static hipError_t tape_wave(const Params &p, int K, const void *in, void *out, hipStream_t s) {
    switch(K) {
    case 1: return launch_wave_cfg<T, MODEL, MOB, EM, 1, CFG, true>(p, in, out, s);
    case 2: return launch_wave_cfg<T, MODEL, MOB, EM, 2, CFG, true>(p, in, out, s);
    case 4: return launch_wave_cfg<T, MODEL, MOB, EM, 4, CFG, true>(p, in, out, s);
    case 8: return launch_wave_cfg<T, MODEL, MOB, EM, 8, CFG, true>(p, in, out, s);
    }
    return hipErrorInvalidValue;
}

// 2D grids with N >= 64 recompute with the register/DPP wave kernel (the
// forward's kernel family: same bits), 1, 2, 4 or 8 states per launch
template <typename T, int DIMS, int MODEL, int MOB, bool EM>
static hipError_t tape_tiles(const Params &p, int K, const void *in, void *out, hipStream_t s) {
    const int N = p.N;
    if constexpr(DIMS == 2) {
        if(N >= 64) {
            if(adj_default_cfg(N) == 5) return tape_wave<T, MODEL, MOB, EM, WaveCfg5>(p, K, in, out, s);
            return tape_wave<T, MODEL, MOB, EM, WaveCfg4>(p, K, in, out, s);
        }
    }
    if constexpr(DIMS == 1 && !EM) {
        if(K == 4 && N >= 64) return launch_tape<T, 1, MODEL, MOB, EM, 64, 1, 1, 4>(p, in, out, s);
    }
    if(K != 1) return hipErrorInvalidValue;
    if constexpr(DIMS == 1) {
        if(N >= 64) return launch_tape<T, 1, MODEL, MOB, EM, 64, 1, 1, 1>(p, in, out, s);
        if(N >= 16) return launch_tape<T, 1, MODEL, MOB, EM, 16, 1, 1, 1>(p, in, out, s);
        return launch_tape<T, 1, MODEL, MOB, EM, 4, 1, 1, 1>(p, in, out, s);
    }
    else if constexpr(DIMS == 2) {
        if(N >= 16) return launch_tape<T, 2, MODEL, MOB, EM, 16, 16, 1, 1>(p, in, out, s);
        return launch_tape<T, 2, MODEL, MOB, EM, 4, 4, 1, 1>(p, in, out, s);
    }
    else {
        if(N >= 16) {
            // the forward's 3D tile (Params::tile3d; the two-launch forward, 3,
            // records its states with the 512-thread tile: the same bits)
            if(p.tile3d == 1 || p.tile3d == 3) return launch_tape<T, 3, MODEL, MOB, EM, 16, 8, 8, 1, 512>(p, in, out, s);
            if(p.tile3d == 2) return launch_tape<T, 3, MODEL, MOB, EM, 16, 16, 4, 1, 512>(p, in, out, s);
            return launch_tape<T, 3, MODEL, MOB, EM, 16, 8, 8, 1>(p, in, out, s);
        }
        return launch_tape<T, 3, MODEL, MOB, EM, 4, 4, 4, 1>(p, in, out, s);
    }
}

hipError_t tape_launch(const Params &p, int precision, int model, int mob, bool em, int K, const void *in,
                       void *tape_out, hipStream_t s) {
    auto go = [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        if(em)
            return by_dims_model_mob<T>(p.dims, model, mob, [&](auto dt, auto mt, auto bt) {
                return tape_tiles<T, decltype(dt)::value, decltype(mt)::value, decltype(bt)::value, true>(p, K, in,
                                                                                                         tape_out, s);
            });
        return by_dims_model_mob<T>(p.dims, model, CH_MOB_CONSTANT, [&](auto dt, auto mt, auto) {
            return tape_tiles<T, decltype(dt)::value, decltype(mt)::value, CH_MOB_CONSTANT, false>(p, K, in, tape_out, s);
        });
    };
    if(precision == CH_FP64) return go(double{});
    return go(float{});
}

hipError_t fill_zero(double *a, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_zero, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

}  // namespace chd
