// adjoint_wave.h -- register/DPP reverse sweep of the 2D euler integrator.
//
// The transpose of psi_{n+1} = psi_n + dt M L(f'(psi_n) - 2k L psi_n) is
//     w = L lambda_{n+1},  lambda_n = lambda_{n+1} + dt M (f''(psi_n) w - 2k L w)
// with the gradient terms <w, psi_n>, <w, L psi_n>, <w, mu_n>, <w, d f'/d delta>
// summed over the tile interior.  L is symmetric on the periodic grid, so
// <w, L psi> = <L w, psi> and <w, mu> = <w, f'(psi)> - 2k <L w, psi>: the terms
// use the L w of the update and need no Laplacian of psi.  Same geometry as the forward wave kernel
// (stencil_wave.h): lane l owns C adjacent columns, wave w owns RPW rows, x
// neighbours through DPP wave shifts, y neighbours from registers and, for the
// first / last row of a wave, through LDS.  KA backward steps run per launch on
// a region carrying a 2 KA halo of lambda; psi_n comes from the tape, the rows
// of later steps in flight while a step computes.  As in the forward kernel, a
// phase computes only the rows later phases read (lambda after backward step
// st within 2(KA-1-st) rows of the interior, w of step st within 2(KA-st)-1):
// a wave whose rows all lie further out skips the phase.
#pragma once

#include "adjoint_math.h"
#include "stencil_wave.h"

namespace chd {

// tape buffers of the reverse sweep's ring (Landau tiles): up to 16 tape
// values per lane in flight, at least two buffers, at most one per step
constexpr int adj_ring_depth(int KA, int cells) {
    const int d = 16 / cells < 2 ? 2 : 16 / cells;
    return d < KA ? d : KA;
}

template <typename T, int MODEL, int KA, class CFG, bool WHOLE>
__global__ __launch_bounds__(64 * CFG::NW) void k_adjoint2d_wave(const T *__restrict__ lam_in, T *__restrict__ lam_out,
                                                               const T *__restrict__ tape, int s_top, Params p,
                                                               int tiles_x, double *__restrict__ gacc) {
    constexpr int C = CFG::C, RPW = CFG::RPW, NW = CFG::NW, RW = CFG::RW, RH = CFG::RH;
    constexpr int H = WHOLE ? 0 : 2 * KA;
    constexpr int W = RW - 4 * KA, TH = RH - 4 * KA;
    static_assert(WHOLE || (W > 0 && TH > 0), "region too small for KA");
    __shared__ WaveExchange<T, 2, NW, C> ex;  // lambda, w
    __shared__ double red[kGradSlots][NW];

    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int N = p.N, mask = p.mask;
    int tile = 0, gx0 = 0, gy0 = 0;
    if constexpr(!WHOLE) {
        tile = xcd_tile(blockIdx.x, gridDim.x);
        gx0 = (tile % tiles_x) * W - H;
        gy0 = (tile / tiles_x) * TH - H;
    }
    // waves above and below, wrapping: in halo mode the first and last waves'
    // outer rows are halo rows whose results are discarded, so what they read
    // there does not matter and the exchange needs no branch
    const int wa = (w + NW - 1) % NW, wb = (w + 1) % NW;
    const size_t NN = (size_t)N * N;

    auto load = [&](const T *__restrict__ f, T(&dst)[C][RPW]) {
#pragma unroll
        for(int j = 0; j < RPW; j++) {
            const size_t row = (size_t)((gy0 + w * RPW + j) & mask) * N;
#pragma unroll
            for(int c = 0; c < C; c++) dst[c][j] = f[row + ((gx0 + C * lane + c) & mask)];
        }
    };
    // the tile interior owns each grid cell once
    auto row_in = [&](int j) {
        const int r = w * RPW + j;
        return WHOLE || (r >= H && r < RH - H && gy0 + r < N);
    };
    auto col_in = [&](int c) {
        const int lx = C * lane + c;
        return WHOLE || (lx >= H && lx < RW - H && gx0 + lx < N);
    };

    // wave w computes a phase whose results are read up to d rows beyond the interior
    auto needed = [&](int d) {
        if constexpr(WHOLE) return true;
        else return w * RPW + RPW > H - d && w * RPW < RH - H + d;
    };

    // backward step st reads psi where its lambda update is needed
    auto reads_tape = [&](int st) { return needed(2 * (KA - 1 - st)); };

    T lam[C][RPW], wv[C][RPW];
    load(lam_in, lam);
#pragma unroll
    for(int j = 0; j < RPW; j++)
#pragma unroll
        for(int c = 0; c < C; c++) wv[c][j] = 0;
    const T inv_dx2 = (T)p.inv_dx2, two_k = (T)p.two_k, Mdt = (T)p.Mdt;
    // gradient terms of this thread's cells over the launch's steps (at most
    // KA C RPW of them), summed in FP64 across threads and launches: in T, so
    // that an FP32 sweep spends no FP64 instructions per cell on them
    T ge = 0, gk = 0, gm = 0, gd = 0;
    // backward step st with psi_n in psi
    auto step = [&](int st, const T(&psi)[C][RPW]) {
        // ---- phase 1: w = L lambda
#pragma unroll
        for(int c = 0; c < C; c++) {
            ex.first[0][w][c][lane] = lam[c][0];
            ex.last[0][w][c][lane] = lam[c][RPW - 1];
        }
        __syncthreads();
        if(needed(2 * (KA - st) - 1)) {
            T la[C], lb[C];
#pragma unroll
            for(int c = 0; c < C; c++) {
                la[c] = ex.last[0][wa][c][lane];
                lb[c] = ex.first[0][wb][c][lane];
            }
#pragma unroll
            for(int j = 0; j < RPW; j++) {
                T m[C][2], q[C][2];
                row_neighbours<T, C, RPW, WHOLE>(lam, j, la, lb, m, q);
#pragma unroll
                for(int c = 0; c < C; c++) wv[c][j] = laplacian<T, 2>(lam[c][j], m[c], q[c], inv_dx2);
            }
        }
        // ---- phase 2: lambda += dt M (f''(psi) w - 2k L w), gradient terms
#pragma unroll
        for(int c = 0; c < C; c++) {
            ex.first[1][w][c][lane] = wv[c][0];
            ex.last[1][w][c][lane] = wv[c][RPW - 1];
        }
        __syncthreads();
        if(needed(2 * (KA - 1 - st))) {
            T wa_[C], wb_[C];
#pragma unroll
            for(int c = 0; c < C; c++) {
                wa_[c] = ex.last[1][wa][c][lane];
                wb_[c] = ex.first[1][wb][c][lane];
            }
#pragma unroll
            for(int j = 0; j < RPW; j++) {
                T m[C][2], q[C][2];
                row_neighbours<T, C, RPW, WHOLE>(wv, j, wa_, wb_, m, q);
                const bool rin = row_in(j);
#pragma unroll
                for(int c = 0; c < C; c++) {
                    const T lw = laplacian<T, 2>(wv[c][j], m[c], q[c], inv_dx2);
                    // f'' in T for Landau FP32 fields (-eps + 3 psi^2), else FP64
                    T fpp;
                    if constexpr(sizeof(T) == 4 && MODEL == CH_LANDAU) fpp = (T)3 * psi[c][j] * psi[c][j] - (T)p.eps;
                    else fpp = (T)fsecond<MODEL>(p, (double)psi[c][j]);
                    if(rin && col_in(c)) {
                        const T wd = wv[c][j], ps = psi[c][j];
                        ge += wd * ps;
                        gk += lw * ps;
                        gm += wd * fprime<T, MODEL>(p, ps) - two_k * (lw * ps);
                        if constexpr(MODEL == CH_SIMPLE_WERTHEIM)
                            gd += (T)((double)wd * dfprime_ddelta<MODEL>(p, (double)ps));
                    }
                    lam[c][j] = lam[c][j] + Mdt * (fpp * wv[c][j] - two_k * lw);
                }
            }
        }
    };

    if constexpr(MODEL == CH_LANDAU && !WHOLE) {
        // psi_n of backward step st sits in ring buffer st % D: the loads of
        // the first D steps issue together with lambda's, and a step's buffer
        // is refilled with the rows of step st + D once the step has read it.
        // The step loop is unrolled, so every buffer index is a constant: no
        // register copies between steps, and a step waits only for its own
        // rows (a wave's loads complete in issue order), D - 1 steps after
        // they were issued.  Every wave loads its rows of every step, also
        // those of a phase it skips (rows wrap, so all addresses lie in the
        // field): a load under a per-wave branch would leave the compiler
        // unable to count the loads in flight, and it would wait for all of
        // them at the first barrier.
        constexpr int D = adj_ring_depth(KA, C * RPW);
        T ring[D][C][RPW];
#pragma unroll
        for(int b = 0; b < D; b++) load(tape + (size_t)(s_top - b) * NN, ring[b]);
#pragma unroll
        for(int st = 0; st < KA; st++) {
            step(st, ring[st % D]);
            if(st + D < KA) load(tape + (size_t)(s_top - st - D) * NN, ring[st % D]);
        }
    }
    else {
        // SimpleWertheim and whole-grid regions, whose registers an unrolled
        // loop would overflow: psi_n of the next step (and, with <= 4 cells
        // per lane, of the one after) is in flight while a step computes
        constexpr bool AHEAD2 = C * RPW <= 4;
        T psi[C][RPW], nxt[C][RPW], nxt2[AHEAD2 ? C : 1][AHEAD2 ? RPW : 1];
#pragma unroll
        for(int j = 0; j < RPW; j++)
#pragma unroll
            for(int c = 0; c < C; c++) {
                nxt[c][j] = 0;
                if constexpr(AHEAD2) nxt2[c][j] = 0;
            }
        if(reads_tape(0)) load(tape + (size_t)s_top * NN, nxt);
        if constexpr(AHEAD2)
            if(KA > 1 && reads_tape(1)) load(tape + (size_t)(s_top - 1) * NN, nxt2);
#pragma unroll 1
        for(int st = 0; st < KA; st++) {
            if constexpr(AHEAD2) {
#pragma unroll
                for(int j = 0; j < RPW; j++)
#pragma unroll
                    for(int c = 0; c < C; c++) {
                        psi[c][j] = nxt[c][j];
                        nxt[c][j] = nxt2[c][j];
                    }
                if(st + 2 < KA && reads_tape(st + 2)) load(tape + (size_t)(s_top - st - 2) * NN, nxt2);
            }
            else {
#pragma unroll
                for(int j = 0; j < RPW; j++)
#pragma unroll
                    for(int c = 0; c < C; c++) psi[c][j] = nxt[c][j];
                if(st + 1 < KA && reads_tape(st + 1)) load(tape + (size_t)(s_top - st - 1) * NN, nxt);
            }
            step(st, psi);
        }
    }

#pragma unroll
    for(int j = 0; j < RPW; j++) {
        if(!row_in(j)) continue;
        const size_t row = (size_t)(gy0 + w * RPW + j) * N;
#pragma unroll
        for(int c = 0; c < C; c++)
            if(col_in(c)) lam_out[row + gx0 + C * lane + c] = lam[c][j];
    }

    const double g[kGradSlots] = {(double)ge, (double)gk, (double)gm, (double)gd};
#pragma unroll
    for(int i = 0; i < kGradSlots; i++) {
        const double v = wave_sum64(g[i]);
        if(lane == 0) red[i][w] = v;
    }
    __syncthreads();
    if(threadIdx.x < kGradSlots) {
        double a = 0;
        for(int i = 0; i < NW; i++) a += red[threadIdx.x][i];
        gacc[kGradSlots * tile + threadIdx.x] += a;
    }
}

}  // namespace chd
