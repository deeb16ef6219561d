// stencil.h -- fused Euler / EulerMobility step kernels (templates).
//
// One launch advances a tile of the periodic grid by K steps:
//   load   psi on the tile + a 2K-cell halo into LDS (A)
//   repeat K times, each time on a box one ring smaller:
//     mu  = f'(psi) - 2k lap(psi)          -> LDS B     (EulerCPU.cpp:20-26)
//     M   = mobility(psi)  (euler_mobility) -> LDS C    (CahnHilliard.cpp:280)
//     psi += dt M lap(mu)                    (EulerCPU.cpp:29-35)
//       or psi += dt div(J), J_d = (M_i + M_{i+e_d})/2 (mu_{i+e_d} - mu_i)/dx
//                                            (EulerMobilityCPU.cpp:44-82)
//   store  the tile interior
// so the reference's 3-4 launches and ~48 B/cell-step of HBM traffic become
// one read and one write of psi per K steps.  The neighbour sum keeps the
// reference's order (-x, +x, -y, +y, -z, +z; EulerCPU.cpp:71-84).
#pragma once

#include "ch_device.h"

namespace chd {

template <int DIMS, int TX, int TY, int TZ, int K>
struct TileGeom {
    static constexpr int H = 2 * K;
    static constexpr int RX = TX + 2 * H;
    static constexpr int RY = DIMS >= 2 ? TY + 2 * H : 1;
    static constexpr int RZ = DIMS == 3 ? TZ + 2 * H : 1;
    static constexpr int RXY = RX * RY;
    static constexpr int R = RXY * RZ;
};


// store the tile interior of the LDS region A into the field `out`
template <typename T, int DIMS, int TX, int TY, int TZ, int RX, int RXY, int H>
__device__ __forceinline__ void store_tile(const T *A, T *__restrict__ out, int N, int tx, int ty, int tz, int NT) {
    constexpr int TY_ = DIMS >= 2 ? TY : 1, TZ_ = DIMS == 3 ? TZ : 1;
    for(int c = threadIdx.x; c < TX * TY_ * TZ_; c += NT) {
        const int lx = c % TX, ly = (c / TX) % TY_, lz = c / (TX * TY_);
        const int li = (H + lx) + (DIMS >= 2 ? (H + ly) * RX : 0) + (DIMS == 3 ? (H + lz) * RXY : 0);
        const int gx = tx * TX + lx;
        const int gy = DIMS >= 2 ? ty * TY + ly : 0;
        const int gz = DIMS == 3 ? tz * TZ + lz : 0;
        out[((size_t)gz * N + gy) * N + gx] = A[li];
    }
}

// TAPE = true: `out` holds K consecutive fields and the state after every one
// of the K steps is stored (the adjoint's recompute of a checkpoint window).
template <typename T, int DIMS, int MODEL, int MOB, bool EM, int TX, int TY, int TZ, int K, int NT, bool TAPE = false>
__global__ __launch_bounds__(NT) void k_stencil(const T *__restrict__ in, T *__restrict__ out, Params p,
                                                int tiles_x, int tiles_y) {
    using G = TileGeom<DIMS, TX, TY, TZ, K>;
    constexpr int RX = G::RX, RY = G::RY, RZ = G::RZ, RXY = G::RXY, R = G::R, H = G::H;
    __shared__ T A[R];
    __shared__ T B[R];
    __shared__ T C[EM && MOB != CH_MOB_CONSTANT ? R : 1];

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % tiles_x;
    const int ty = (tile / tiles_x) % tiles_y;
    const int tz = tile / (tiles_x * tiles_y);
    const int x0 = tx * TX - H, y0 = ty * TY - H, z0 = tz * TZ - H;
    const int N = p.N, mask = p.mask;

    for(int c = threadIdx.x; c < R; c += NT) {
        const int lx = c % RX, ly = (c / RX) % RY, lz = c / RXY;
        const int gx = (x0 + lx) & mask;
        const int gy = DIMS >= 2 ? (y0 + ly) & mask : 0;
        const int gz = DIMS == 3 ? (z0 + lz) & mask : 0;
        A[c] = in[((size_t)gz * N + gy) * N + gx];
    }
    __syncthreads();

    constexpr int off[3] = {1, RX, RXY};
    const T Mc = (T)p.M;

#pragma unroll
    for(int s = 0; s < K; s++) {
        {   // chemical potential (and mobility) on the box [2s+1, R-2s-1)
            const int lo = 2 * s + 1;
            const int EX = RX - 2 * lo;
            const int EY = DIMS >= 2 ? RY - 2 * lo : 1;
            const int EZ = DIMS == 3 ? RZ - 2 * lo : 1;
            for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
                const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
                const int li = (lo + ex) + (DIMS >= 2 ? (lo + ey) * RX : 0) + (DIMS == 3 ? (lo + ez) * RXY : 0);
                T m[DIMS], q[DIMS];
#pragma unroll
                for(int d = 0; d < DIMS; d++) { m[d] = A[li - off[d]]; q[d] = A[li + off[d]]; }
                const T a = A[li];
                B[li] = chem_pot<T, DIMS, MODEL>(p, a, m, q);
                if constexpr(EM && MOB != CH_MOB_CONSTANT) C[li] = mobility<T, MOB>(p, a);
            }
        }
        __syncthreads();
        {   // update on the box [2s+2, R-2s-2), in place in A
            const int lo = 2 * s + 2;
            const int EX = RX - 2 * lo;
            const int EY = DIMS >= 2 ? RY - 2 * lo : 1;
            const int EZ = DIMS == 3 ? RZ - 2 * lo : 1;
            for(int c = threadIdx.x; c < EX * EY * EZ; c += NT) {
                const int ex = c % EX, ey = (c / EX) % EY, ez = c / (EX * EY);
                const int li = (lo + ex) + (DIMS >= 2 ? (lo + ey) * RX : 0) + (DIMS == 3 ? (lo + ez) * RXY : 0);
                T m[DIMS], q[DIMS];
#pragma unroll
                for(int d = 0; d < DIMS; d++) { m[d] = B[li - off[d]]; q[d] = B[li + off[d]]; }
                const T mu = B[li];
                if constexpr(!EM) {
                    A[li] = euler_update<T, DIMS>(p, A[li], mu, m, q);
                }
                else {
                    T Mm[DIMS], Mq[DIMS], Mi;
                    if constexpr(MOB == CH_MOB_CONSTANT) {
                        Mi = Mc;
#pragma unroll
                        for(int d = 0; d < DIMS; d++) { Mm[d] = Mc; Mq[d] = Mc; }
                    }
                    else {
                        Mi = C[li];
#pragma unroll
                        for(int d = 0; d < DIMS; d++) { Mm[d] = C[li - off[d]]; Mq[d] = C[li + off[d]]; }
                    }
                    A[li] = mobility_update<T, DIMS>(p, A[li], mu, m, q, Mi, Mm, Mq);
                }
            }
        }
        __syncthreads();
        if constexpr(TAPE) {
            if(s < K - 1) store_tile<T, DIMS, TX, TY, TZ, RX, RXY, H>(A, out + (size_t)s * ((size_t)1 << (p.log2N * DIMS)), N,
                                                                     tx, ty, tz, NT);
        }
    }

    T *dst = TAPE ? out + (size_t)(K - 1) * ((size_t)1 << (p.log2N * DIMS)) : out;
    store_tile<T, DIMS, TX, TY, TZ, RX, RXY, H>(A, dst, N, tx, ty, tz, NT);
}

template <typename T, int DIMS, int MODEL, int MOB, bool EM, int TX, int TY, int TZ, int K, int NT = 256>
hipError_t launch_stencil_cfg(const Params &p, const void *in, void *out, hipStream_t s) {
    const int tiles_x = p.N / TX;
    const int tiles_y = DIMS >= 2 ? p.N / TY : 1;
    const int tiles_z = DIMS == 3 ? p.N / TZ : 1;
    const int nb = tiles_x * tiles_y * tiles_z;
    hipLaunchKernelGGL((k_stencil<T, DIMS, MODEL, MOB, EM, TX, TY, TZ, K, NT>), dim3(nb), dim3(NT), 0, s,
                       (const T *)in, (T *)out, p, tiles_x, tiles_y);
    return hipGetLastError();
}

}  // namespace chd
