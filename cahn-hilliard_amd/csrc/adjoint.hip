// adjoint.hip -- reverse-mode sweep of the Euler integrator (constant mobility).
//
// Forward step (src/integrators/EulerCPU.cpp:18-36):
//     psi_{n+1} = psi_n + dt M L(mu_n),  mu_n = f'(psi_n) - 2k L psi_n
// with L the periodic (2d+1)-point Laplacian (symmetric).  Its transpose gives
//     lambda_n = lambda_{n+1} + dt M ( f''(psi_n) w - 2k L w ),  w = L lambda_{n+1}
// and the parameter gradients of the loss, summed over steps,
//     dL/deps = -dt M <w, psi_n>,  dL/dk = -2 dt M <w, L psi_n>,  dL/dM = dt <w, mu_n>
// (Landau: d mu/d eps = -psi).  One launch runs KA backward steps on a tile
// carrying a 2 KA halo of lambda; the states psi_n come from the tape of the
// current checkpoint window.  Each block keeps FP64 partial sums of the three
// inner products for its own tile (no atomics: fixed tile -> slot mapping, so
// the result is deterministic).  This reverse sweep has no reference
// counterpart; it is pinned by the oracle's FP64 adjoint and by finite
// differences of the forward map.
#include "ch_internal.h"
#include "stencil.h"

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

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

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
    __shared__ double red[3][NT / 64];

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
    double ge = 0, gk = 0, gm = 0;
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

    ge = wave_sum64(ge);
    gk = wave_sum64(gk);
    gm = wave_sum64(gm);
    const int w = threadIdx.x >> 6;
    if((threadIdx.x & 63) == 0) { red[0][w] = ge; red[1][w] = gk; red[2][w] = gm; }
    __syncthreads();
    if(threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for(int i = 0; i < NT / 64; i++) { a += red[0][i]; b += red[1][i]; c += red[2][i]; }
        gacc[3 * tile] += a;
        gacc[3 * tile + 1] += b;
        gacc[3 * tile + 2] += c;
    }
}

__global__ void k_fill_zero(double *__restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) a[i] = 0.0;
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

int adjoint_steps_per_launch(int dims, int N) { return (dims <= 2 && N >= 64) ? 4 : 1; }

// number of tiles (gradient accumulator slots) of the adjoint launches
int adjoint_tiles(int dims, int N) {
    if(dims == 1) return N >= 64 ? N / 64 : (N >= 16 ? N / 16 : N / 4);
    if(dims == 2) return N >= 64 ? (N / 64) * (N / 16) : (N >= 16 ? (N / 16) * (N / 16) : (N / 4) * (N / 4));
    return N >= 16 ? (N / 16) * (N / 8) * (N / 8) : (N / 4) * (N / 4) * (N / 4);
}

hipError_t adjoint_launch(const Params &p, int precision, int model, int KA, const void *lam_in, void *lam_out,
                          const void *tape, int s_top, double *gacc, hipStream_t s) {
    const int d = p.dims;
    auto go = [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        if(model == CH_LANDAU) {
            if(d == 1) return adj_tiles<T, 1, CH_LANDAU>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
            if(d == 2) return adj_tiles<T, 2, CH_LANDAU>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
            return adj_tiles<T, 3, CH_LANDAU>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
        }
        if(d == 1) return adj_tiles<T, 1, CH_SIMPLE_WERTHEIM>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
        if(d == 2) return adj_tiles<T, 2, CH_SIMPLE_WERTHEIM>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
        return adj_tiles<T, 3, CH_SIMPLE_WERTHEIM>(p, KA, lam_in, lam_out, tape, s_top, gacc, s);
    };
    if(precision == CH_FP64) return go(double{});
    return go(float{});
}

// recompute of a checkpoint window: K Euler steps from `in`, all K states to tape_out
template <typename T, int DIMS, int MODEL, int TX, int TY, int TZ, int K>
static hipError_t launch_tape(const Params &p, const void *in, void *tape_out, hipStream_t s) {
    constexpr int NT = 256;
    const int tiles_x = p.N / TX, tiles_y = DIMS >= 2 ? p.N / TY : 1, tiles_z = DIMS == 3 ? p.N / TZ : 1;
    hipLaunchKernelGGL((k_stencil<T, DIMS, MODEL, CH_MOB_CONSTANT, false, TX, TY, TZ, K, NT, true>),
                       dim3(tiles_x * tiles_y * tiles_z), dim3(NT), 0, s, (const T *)in, (T *)tape_out, p, tiles_x, tiles_y);
    return hipGetLastError();
}

template <typename T, int DIMS, int MODEL>
static hipError_t tape_tiles(const Params &p, int K, const void *in, void *out, hipStream_t s) {
    const int N = p.N;
    if constexpr(DIMS == 1) {
        if(K == 4 && N >= 64) return launch_tape<T, 1, MODEL, 64, 1, 1, 4>(p, in, out, s);
        if(N >= 64) return launch_tape<T, 1, MODEL, 64, 1, 1, 1>(p, in, out, s);
        if(N >= 16) return launch_tape<T, 1, MODEL, 16, 1, 1, 1>(p, in, out, s);
        return launch_tape<T, 1, MODEL, 4, 1, 1, 1>(p, in, out, s);
    }
    else if constexpr(DIMS == 2) {
        if(K == 4 && N >= 64) return launch_tape<T, 2, MODEL, 64, 16, 1, 4>(p, in, out, s);
        if(N >= 64) return launch_tape<T, 2, MODEL, 64, 16, 1, 1>(p, in, out, s);
        if(N >= 16) return launch_tape<T, 2, MODEL, 16, 16, 1, 1>(p, in, out, s);
        return launch_tape<T, 2, MODEL, 4, 4, 1, 1>(p, in, out, s);
    }
    else {
        if(N >= 16) return launch_tape<T, 3, MODEL, 16, 8, 8, 1>(p, in, out, s);
        return launch_tape<T, 3, MODEL, 4, 4, 4, 1>(p, in, out, s);
    }
}

hipError_t tape_launch(const Params &p, int precision, int model, int K, const void *in, void *tape_out,
                       hipStream_t s) {
    const int d = p.dims;
    auto go = [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        if(model == CH_LANDAU) {
            if(d == 1) return tape_tiles<T, 1, CH_LANDAU>(p, K, in, tape_out, s);
            if(d == 2) return tape_tiles<T, 2, CH_LANDAU>(p, K, in, tape_out, s);
            return tape_tiles<T, 3, CH_LANDAU>(p, K, in, tape_out, s);
        }
        if(d == 1) return tape_tiles<T, 1, CH_SIMPLE_WERTHEIM>(p, K, in, tape_out, s);
        if(d == 2) return tape_tiles<T, 2, CH_SIMPLE_WERTHEIM>(p, K, in, tape_out, s);
        return tape_tiles<T, 3, CH_SIMPLE_WERTHEIM>(p, K, in, tape_out, s);
    };
    if(precision == CH_FP64) return go(double{});
    return go(float{});
}

hipError_t fill_zero(double *a, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_zero, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

}  // namespace chd
