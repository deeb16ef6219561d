// stencil_wave.h -- register/DPP 2D stencil kernels (Euler, EulerMobility).
//
// A block is NW waves.  Lane l of every wave owns C adjacent grid columns
// x0 + C l + c (a wave spans 64 C consecutive x), and wave w owns RPW
// consecutive rows of the block's region, held in registers.  Per step:
//   x neighbours  the lane's own registers, and for its first / last column
//                 the neighbouring lane's last / first column through DPP wave
//                 shifts (v_mov_b32_dpp wave_shr:1 / wave_shl:1): 2 DPP moves
//                 per row per field and phase for C cells
//   y neighbours  the lane's own registers; the first / last row of each wave
//                 goes through LDS
// Halo mode: K steps run per launch on a region 64C x (NW RPW) that carries a
// 2K halo on every side; each phase invalidates one more ring, so after K
// steps the (64C - 4K) x (NW RPW - 4K) interior is exact and is stored.  Tiles
// that do not divide N are clipped at the store.  A phase computes only the
// rows that later phases read: psi after step s is read within 2(K-1-s) rows
// of the interior and mu of step s within 2(K-s)-1, so a wave whose rows all
// lie further out skips the phase.  Every row such a band reads lies in the
// band of the phase before, so a skipped wave's stale rows reach only rows
// that are themselves discarded.
// Whole mode (N == 64 C == NW RPW): the region is the entire periodic grid; x
// neighbours use DPP wave rotates and y neighbours wrap through LDS, so no
// halo is recomputed and every cell is stored.
// All per-cell arithmetic is the shared code of ch_device.h, so results are
// bit-identical to k_stencil's whatever the mode, K, C or tile shape.
#pragma once

#include "ch_device.h"

namespace chd {

// value held by lane - 1; lane 0 reads 0 (ROT: lane 63's value)
template <bool ROT>
__device__ __forceinline__ double from_left(double v) {
    constexpr int ctl = ROT ? 0x13C : 0x138;  // wave_ror:1 / wave_shr:1
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// value held by lane + 1; lane 63 reads 0 (ROT: lane 0's value)
template <bool ROT>
__device__ __forceinline__ double from_right(double v) {
    constexpr int ctl = ROT ? 0x134 : 0x130;  // wave_rol:1 / wave_shl:1
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <bool ROT>
__device__ __forceinline__ float from_left(float v) {
    constexpr int ctl = ROT ? 0x13C : 0x138;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctl, 0xF, 0xF, true));
}
template <bool ROT>
__device__ __forceinline__ float from_right(float v) {
    constexpr int ctl = ROT ? 0x134 : 0x130;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctl, 0xF, 0xF, true));
}

// region shape: C columns per lane, RPW rows per wave, NW waves
template <int C_, int RPW_, int NW_>
struct WaveCfg {
    static constexpr int C = C_, RPW = RPW_, NW = NW_;
    static constexpr int RW = 64 * C, RH = RPW * NW;
};

template <typename T, int NF, int NW, int C>
struct WaveExchange {
    T first[NF][NW][C][64];  // first row of each wave
    T last[NF][NW][C][64];   // last row of each wave
};

template <int MOB, bool EM>
constexpr int exchange_fields() { return (EM && MOB != CH_MOB_CONSTANT) ? 3 : 2; }

// the 4 neighbours of every cell of row j of field v (C columns per lane)
template <typename T, int C, int RPW, bool ROT>
__device__ __forceinline__ void row_neighbours(const T (&v)[C][RPW], int j, const T (&above)[C], const T (&below)[C],
                                               T (&m)[C][2], T (&q)[C][2]) {
    const T xl = from_left<ROT>(v[C - 1][j]);
    const T xr = from_right<ROT>(v[0][j]);
#pragma unroll
    for(int c = 0; c < C; c++) {
        m[c][0] = c == 0 ? xl : v[c - 1][j];
        q[c][0] = c == C - 1 ? xr : v[c + 1][j];
        m[c][1] = j == 0 ? above[c] : v[c][j - 1];
        q[c][1] = j == RPW - 1 ? below[c] : v[c][j + 1];
    }
}

// K steps on one region whose top-left grid cell is (gx0, gy0) (halo mode) or
// which is the whole grid (WHOLE).
// TAPE: `out` holds K consecutive fields and the state after every step is
// stored (the tile interior is valid after each of the K steps).
// UPD (Euler only) selects how the update psi + dt M lap(mu) is rounded:
//   UPD_PLAIN   rounded every step
//   UPD_LAUNCH  the launch keeps each cell's loaded value psi0 and the sum of
//               its increments in registers (euler_increment); the field used
//               for mu is psi0 + sum, so the only rounding of psi that stays
//               is the one at the store: one per launch instead of one per
//               step, at one extra add per cell and step and no extra traffic
//   UPD_STORED  compensated sums whose lost low-order parts are read from cin
//               and written to cout beside the field (euler_update_comp)
//   UPD_LOW8    UPD_LAUNCH over a field that carries an 8-bit low-order part
//               beside psi (low8_decode / low8_encode, FP32 only): the launch
//               starts its increment sum from the stored part, and the store
//               keeps the rounding remainder of psi in it, so the field holds
//               about 32 significant bits and no rounding accumulates
// XW (x-whole): the region spans the whole periodic x axis (N == 64 C, gx0 =
// 0): x neighbours use DPP wave rotates and carry no halo, y keeps its 2K halo.
enum { UPD_PLAIN = 0, UPD_LAUNCH = 1, UPD_STORED = 2, UPD_LOW8 = 3 };
template <typename T, int MODEL, int MOB, bool EM, int K, class CFG, bool WHOLE, bool TAPE = false, int UPD = UPD_PLAIN,
          bool XW = false>
__device__ __forceinline__ void wave_region(const T *__restrict__ in, T *__restrict__ out, const Params &p, int gx0,
                                            int gy0, WaveExchange<T, exchange_fields<MOB, EM>(), CFG::NW, CFG::C> &ex,
                                            const void *__restrict__ cin = nullptr, void *__restrict__ cout = nullptr) {
    static_assert(UPD == UPD_PLAIN || (!EM && !TAPE), "compensated updates for the Euler integrator only");
    static_assert(UPD != UPD_LOW8 || sizeof(T) == 4, "8-bit low-order parts of FP32 fields only");
    static_assert(!(XW && WHOLE), "x-whole regions keep a y halo");
    constexpr bool COMP = UPD == UPD_STORED, LOW8 = UPD == UPD_LOW8, ACC = UPD == UPD_LAUNCH || LOW8;
    constexpr int C = CFG::C, RPW = CFG::RPW, NW = CFG::NW, RW = CFG::RW, RH = CFG::RH;
    constexpr bool HASMOB = EM && MOB != CH_MOB_CONSTANT;
    constexpr int H = WHOLE ? 0 : 2 * K;          // y halo
    constexpr int HX = (WHOLE || XW) ? 0 : 2 * K;  // x halo
    constexpr bool ROTX = WHOLE || XW;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int N = p.N, mask = p.mask;
    // waves above and below, wrapping: in halo mode the first and last waves'
    // outer rows are halo rows whose results are discarded, so what they read
    // there does not matter and the exchange needs no branch
    const int wa = (w + NW - 1) % NW, wb = (w + 1) % NW;

    T psi[C][RPW], mu[C][RPW], mob[HASMOB ? C : 1][HASMOB ? RPW : 1], cmp[COMP ? C : 1][COMP ? RPW : 1];
    T psi0[ACC ? C : 1][ACC ? RPW : 1], acc[ACC ? C : 1][ACC ? RPW : 1];
#pragma unroll
    for(int j = 0; j < RPW; j++) {
        const size_t row = (size_t)((gy0 + w * RPW + j) & mask) * N;
#pragma unroll
        for(int c = 0; c < C; c++) {
            const size_t at = row + ((gx0 + C * lane + c) & mask);
            psi[c][j] = in[at];
            mu[c][j] = 0;
            if constexpr(HASMOB) mob[c][j] = 0;
            if constexpr(COMP) cmp[c][j] = static_cast<const T *>(cin)[at];
            if constexpr(ACC) {
                psi0[c][j] = psi[c][j];
                if constexpr(LOW8) acc[c][j] = low8_decode(psi[c][j], static_cast<const int8_t *>(cin)[at]);
                else acc[c][j] = 0;
            }
        }
    }
    // wave w computes a phase whose results are read up to d rows beyond the interior
    auto needed = [&](int d) {
        if constexpr(WHOLE) return true;
        else return w * RPW + RPW > H - d && w * RPW < RH - H + d;
    };

    // the tile interior (the whole grid in WHOLE mode) of psi into dst
    auto store_interior = [&](T *__restrict__ dst) {
#pragma unroll
        for(int j = 0; j < RPW; j++) {
            const int r = w * RPW + j;
            const int gy = gy0 + r;
            if(!(WHOLE || (r >= H && r < RH - H && gy < N))) continue;
#pragma unroll
            for(int c = 0; c < C; c++) {
                const int lx = C * lane + c;
                const int gx = gx0 + lx;
                if(WHOLE || XW || (lx >= HX && lx < RW - HX && gx < N)) {
                    const size_t at = (size_t)gy * N + gx;
                    if constexpr(ACC) {
                        const T v = psi0[c][j] + acc[c][j];
                        dst[at] = v;
                        // the remainder of the rounding (Fast2Sum: psi0 is the larger term
                        // wherever the remainder is not negligible)
                        if constexpr(LOW8)
                            static_cast<int8_t *>(cout)[at] =
                                (int8_t)low8_encode(v, acc[c][j] - (v - psi0[c][j]));
                    }
                    else {
                        dst[at] = psi[c][j];
                    }
                    if constexpr(COMP) static_cast<T *>(cout)[at] = cmp[c][j];
                }
            }
        }
    };

    const T Mc = (T)p.M;
#pragma unroll 1
    for(int s = 0; s < K; s++) {
        // ---- phase 1: mu (and M) from psi
        if constexpr(ACC) {
            // psi of this step from the launch's start value and increments
#pragma unroll
            for(int j = 0; j < RPW; j++)
#pragma unroll
                for(int c = 0; c < C; c++) psi[c][j] = psi0[c][j] + acc[c][j];
        }
#pragma unroll
        for(int c = 0; c < C; c++) {
            ex.first[0][w][c][lane] = psi[c][0];
            ex.last[0][w][c][lane] = psi[c][RPW - 1];
        }
        __syncthreads();
        if(needed(2 * (K - s) - 1)) {
            T above[C], below[C];
#pragma unroll
            for(int c = 0; c < C; c++) {
                above[c] = ex.last[0][wa][c][lane];
                below[c] = ex.first[0][wb][c][lane];
            }
#pragma unroll
            for(int j = 0; j < RPW; j++) {
                T m[C][2], q[C][2];
                row_neighbours<T, C, RPW, ROTX>(psi, j, above, below, m, q);
#pragma unroll
                for(int c = 0; c < C; c++) {
                    mu[c][j] = chem_pot<T, 2, MODEL>(p, psi[c][j], m[c], q[c]);
                    if constexpr(HASMOB) mob[c][j] = mobility<T, MOB>(p, psi[c][j]);
                }
            }
        }
        // ---- phase 2: psi update from mu (and M)
#pragma unroll
        for(int c = 0; c < C; c++) {
            ex.first[1][w][c][lane] = mu[c][0];
            ex.last[1][w][c][lane] = mu[c][RPW - 1];
            if constexpr(HASMOB) {
                ex.first[2][w][c][lane] = mob[c][0];
                ex.last[2][w][c][lane] = mob[c][RPW - 1];
            }
        }
        __syncthreads();
        if(needed(2 * (K - 1 - s))) {
            T mu_a[C], mu_b[C], M_a[C], M_b[C];
#pragma unroll
            for(int c = 0; c < C; c++) {
                mu_a[c] = ex.last[1][wa][c][lane];
                mu_b[c] = ex.first[1][wb][c][lane];
                if constexpr(HASMOB) {
                    M_a[c] = ex.last[2][wa][c][lane];
                    M_b[c] = ex.first[2][wb][c][lane];
                }
            }
#pragma unroll
            for(int j = 0; j < RPW; j++) {
                T m[C][2], q[C][2];
                row_neighbours<T, C, RPW, ROTX>(mu, j, mu_a, mu_b, m, q);
                if constexpr(HASMOB) {
                    T Mm[C][2], Mq[C][2];
                    row_neighbours<T, C, RPW, ROTX>(mob, j, M_a, M_b, Mm, Mq);
#pragma unroll
                    for(int c = 0; c < C; c++)
                        psi[c][j] = mobility_update<T, 2>(p, psi[c][j], mu[c][j], m[c], q[c], mob[c][j], Mm[c], Mq[c]);
                }
                else {
#pragma unroll
                    for(int c = 0; c < C; c++) {
                        if constexpr(COMP) {
                            psi[c][j] = euler_update_comp<T, 2>(p, psi[c][j], mu[c][j], m[c], q[c], cmp[c][j]);
                        }
                        else if constexpr(ACC) {
                            euler_increment<T, 2>(p, mu[c][j], m[c], q[c], acc[c][j]);
                        }
                        else if constexpr(!EM) {
                            psi[c][j] = euler_update<T, 2>(p, psi[c][j], mu[c][j], m[c], q[c]);
                        }
                        else {
                            const T Mm[2] = {Mc, Mc}, Mq[2] = {Mc, Mc};
                            psi[c][j] = mobility_update<T, 2>(p, psi[c][j], mu[c][j], m[c], q[c], Mc, Mm, Mq);
                        }
                    }
                }
            }
        }
        if constexpr(TAPE) {
            if(s < K - 1) store_interior(out + (size_t)s * N * N);
        }
    }
    store_interior(TAPE ? out + (size_t)(K - 1) * N * N : out);
}

template <typename T, int MODEL, int MOB, bool EM, int K, class CFG, bool WHOLE, bool TAPE = false>
__global__ __launch_bounds__(64 * CFG::NW) void k_stencil2d_wave(const T *__restrict__ in, T *__restrict__ out, Params p,
                                                               int tiles_x) {
    constexpr int W = CFG::RW - 4 * K, TH = CFG::RH - 4 * K;
    static_assert(WHOLE || (W > 0 && TH > 0), "tile too small for K");
    __shared__ WaveExchange<T, exchange_fields<MOB, EM>(), CFG::NW, CFG::C> ex;
    if constexpr(WHOLE) {
        wave_region<T, MODEL, MOB, EM, K, CFG, true, TAPE>(in, out, p, 0, 0, ex);
    }
    else {
        const int tile = xcd_tile(blockIdx.x, gridDim.x);
        const int tx = tile % tiles_x, ty = tile / tiles_x;
        wave_region<T, MODEL, MOB, EM, K, CFG, false, TAPE>(in, out, p, tx * W - 2 * K, ty * TH - 2 * K, ex);
    }
}

template <typename T, int MODEL, int MOB, bool EM, int K, class CFG, bool TAPE = false>
hipError_t launch_wave_cfg(const Params &p, const void *in, void *out, hipStream_t s) {
    if(CFG::RW == CFG::RH && p.N == CFG::RW) {
        hipLaunchKernelGGL((k_stencil2d_wave<T, MODEL, MOB, EM, K, CFG, true, TAPE>), dim3(1), dim3(64 * CFG::NW), 0, s,
                           (const T *)in, (T *)out, p, 1);
        return hipGetLastError();
    }
    constexpr int W = CFG::RW - 4 * K, TH = CFG::RH - 4 * K;
    const int tiles_x = (p.N + W - 1) / W;
    const int tiles_y = (p.N + TH - 1) / TH;
    hipLaunchKernelGGL((k_stencil2d_wave<T, MODEL, MOB, EM, K, CFG, false, TAPE>), dim3(tiles_x * tiles_y),
                       dim3(64 * CFG::NW), 0, s, (const T *)in, (T *)out, p, tiles_x);
    return hipGetLastError();
}

}  // namespace chd
