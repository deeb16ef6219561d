// stencil_wave.h -- register/DPP 2D stencil kernel (Euler, EulerMobility).
//
// A block is NW waves.  Lane l of every wave owns grid column x0 + l (a wave
// spans 64 consecutive x), and wave w owns RPW consecutive rows of the block's
// region, held in registers.  Per step:
//   x neighbours  DPP wave shifts (v_mov_b32_dpp wave_shr:1 / wave_shl:1)
//   y neighbours  the lane's own registers; the first / last row of each wave
//                 goes through LDS (2 rows per field per phase)
// K steps run per launch on a region 64 x (NW RPW) that carries a 2K halo on
// every side; each phase invalidates one more ring, so after K steps the
// (64 - 4K) x (NW RPW - 4K) interior is exact and is stored.  Tiles that do not
// divide N are clipped at the store.  All per-cell arithmetic is the shared
// code of ch_device.h, so results are bit-identical to k_stencil's.
#pragma once

#include "ch_device.h"

namespace chd {

// value held by lane - 1 (the x - 1 neighbour); lane 0 reads 0
__device__ __forceinline__ double from_left(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// value held by lane + 1 (the x + 1 neighbour); lane 63 reads 0
__device__ __forceinline__ double from_right(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
}

template <typename T, int MODEL, int MOB, bool EM, int K, int RPW, int NW>
__global__ __launch_bounds__(64 * NW) void k_stencil2d_wave(const T *__restrict__ in, T *__restrict__ out, Params p,
                                                          int tiles_x) {
    constexpr int W = 64 - 4 * K;        // interior columns per tile
    constexpr int RH = RPW * NW;         // region rows
    constexpr int TH = RH - 4 * K;       // interior rows per tile
    constexpr bool HASMOB = EM && MOB != CH_MOB_CONSTANT;
    constexpr int NF = HASMOB ? 3 : 2;   // exchanged fields: psi, mu (, M)
    static_assert(W > 0 && TH > 0, "tile too small for K");
    __shared__ T ex_first[NF][NW][64];   // first row of each wave
    __shared__ T ex_last[NF][NW][64];    // last row of each wave

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int gx0 = tx * W - 2 * K, gy0 = ty * TH - 2 * K;
    const int N = p.N, mask = p.mask;
    const int col = (gx0 + lane) & mask;

    T psi[RPW], mu[RPW], mob[HASMOB ? RPW : 1];
#pragma unroll
    for(int j = 0; j < RPW; j++) psi[j] = in[(size_t)((gy0 + w * RPW + j) & mask) * N + col];

    const T Mc = (T)p.M;
#pragma unroll 1
    for(int s = 0; s < K; s++) {
        // ---- phase 1: mu (and M) from psi
        ex_first[0][w][lane] = psi[0];
        ex_last[0][w][lane] = psi[RPW - 1];
        __syncthreads();
        const T above = (w > 0) ? ex_last[0][w - 1][lane] : psi[0];
        const T below = (w < NW - 1) ? ex_first[0][w + 1][lane] : psi[RPW - 1];
#pragma unroll
        for(int j = 0; j < RPW; j++) {
            const T c = psi[j];
            const T m[2] = {from_left(c), j == 0 ? above : psi[j - 1]};
            const T q[2] = {from_right(c), j == RPW - 1 ? below : psi[j + 1]};
            mu[j] = chem_pot<T, 2, MODEL>(p, c, m, q);
            if constexpr(HASMOB) mob[j] = mobility<T, MOB>(p, c);
        }
        // ---- phase 2: psi update from mu (and M)
        ex_first[1][w][lane] = mu[0];
        ex_last[1][w][lane] = mu[RPW - 1];
        if constexpr(HASMOB) {
            ex_first[2][w][lane] = mob[0];
            ex_last[2][w][lane] = mob[RPW - 1];
        }
        __syncthreads();
        const T mu_above = (w > 0) ? ex_last[1][w - 1][lane] : mu[0];
        const T mu_below = (w < NW - 1) ? ex_first[1][w + 1][lane] : mu[RPW - 1];
        T M_above = Mc, M_below = Mc;
        if constexpr(HASMOB) {
            M_above = (w > 0) ? ex_last[2][w - 1][lane] : mob[0];
            M_below = (w < NW - 1) ? ex_first[2][w + 1][lane] : mob[RPW - 1];
        }
#pragma unroll
        for(int j = 0; j < RPW; j++) {
            const T c = mu[j];
            const T m[2] = {from_left(c), j == 0 ? mu_above : mu[j - 1]};
            const T q[2] = {from_right(c), j == RPW - 1 ? mu_below : mu[j + 1]};
            if constexpr(!EM) {
                psi[j] = euler_update<T, 2>(p, psi[j], c, m, q);
            }
            else if constexpr(!HASMOB) {
                const T Mm[2] = {Mc, Mc}, Mq[2] = {Mc, Mc};
                psi[j] = mobility_update<T, 2>(p, psi[j], c, m, q, Mc, Mm, Mq);
            }
            else {
                const T Mi = mob[j];
                const T Mm[2] = {from_left(Mi), j == 0 ? M_above : mob[j - 1]};
                const T Mq[2] = {from_right(Mi), j == RPW - 1 ? M_below : mob[j + 1]};
                psi[j] = mobility_update<T, 2>(p, psi[j], c, m, q, Mi, Mm, Mq);
            }
        }
    }

    const int gx = gx0 + lane;
    const bool col_ok = lane >= 2 * K && lane < 64 - 2 * K && gx < N;
#pragma unroll
    for(int j = 0; j < RPW; j++) {
        const int r = w * RPW + j;
        const int gy = gy0 + r;
        if(col_ok && r >= 2 * K && r < RH - 2 * K && gy < N) out[(size_t)gy * N + gx] = psi[j];
    }
}

template <typename T, int MODEL, int MOB, bool EM, int K, int RPW, int NW>
hipError_t launch_wave_cfg(const Params &p, const void *in, void *out, hipStream_t s) {
    constexpr int W = 64 - 4 * K, TH = RPW * NW - 4 * K;
    const int tiles_x = (p.N + W - 1) / W;
    const int tiles_y = (p.N + TH - 1) / TH;
    hipLaunchKernelGGL((k_stencil2d_wave<T, MODEL, MOB, EM, K, RPW, NW>), dim3(tiles_x * tiles_y), dim3(64 * NW), 0, s,
                       (const T *)in, (T *)out, p, tiles_x);
    return hipGetLastError();
}

}  // namespace chd
