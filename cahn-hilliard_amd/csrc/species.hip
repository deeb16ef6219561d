// species.hip -- multi-species Wertheim models and their integrators.
//
// Models (per cell, on the species vector rho_s, FP64 in both precisions):
//   SalehWertheim   src/models/SalehWertheim.cpp:47-110    3 species, closed-form X
//   RicciWertheim   src/models/RicciWertheim.cpp:43-98     2 species, closed-form X
//   GenericWertheim src/models/GenericWertheim.cpp:184-291 any patch topology:
//       X_p = 1 / (1 + sum_terms rho_s X_q c), Gauss-Seidel sweeps over the
//       patch types in ascending order until a sweep moves no X by 1e-8
//       (at most 10000 sweeps), warm-started from the cell's X of the
//       previous evaluation (X lives in HBM, patch-major, n_patches per cell).
// A species whose density is exactly 0 gets f' = 0 (the reference leaves its
// previous value in a static buffer).  f' takes its logarithms from
// log_series (ch_device.h, within 2 ulp); the free energy keeps the library's.
//
// Integrators over species-major fields rho[s * bins + i]:
//   euler / euler_mobility: k_sp_mu (mobility from the previous X, X update,
//     f', mu_s = f'_s - 2 k_s lap rho_s) then k_sp_update (rho_s += dt M lap mu_s,
//     or the flux form with M_s per species), in place: 2 launches per step
//     (EulerCPU.cpp:18-36, EulerMobilityCPU.cpp:31-83 over species);
//   pseudospectral: k_sp_fprime, then per species r2c, k_sp_ps_update (k_s in
//     the denominator, PseudospectralCPU.cpp:112-136) and c2r.
#include <algorithm>
#include <cmath>

#include "ch_internal.h"

namespace chd {

namespace {

template <typename T>
using C2 = typename Complex<T>::type;

constexpr int kSpBlocks = 2048;

__device__ void generic_update_X(const SpModel &m, const double *rhos, double *X) {
#pragma clang fp contract(off)
    for(int iter = 0; iter < 10000; iter++) {
        double max_delta = 0.0;
        for(int u = 0; u < m.n_unique; u++) {
            const int patch = m.unique[u];
            double sum = 0.0;
            for(int t = m.term_start[u]; t < m.term_start[u + 1]; t++) {
                const double rho = rhos[m.t_species[t]];
                if(rho != 0.0) sum += rho * X[m.t_other[t]] * m.t_coeff[t];
            }
            const double new_X = 1.0 / (1.0 + sum);
            const double d = fabs(new_X - X[patch]);
            max_delta = (max_delta < d) ? d : max_delta;
            X[patch] = new_X;
        }
        if(max_delta < 1e-8) return;
    }
}

__device__ double saleh_contribution(const SpModel &m, const double *rhos, int species) {
#pragma clang fp contract(off)
    const double delta = species == 0 ? m.dAA : m.dBB;
    const double rho_factor = delta * (m.val[species] * rhos[species] + m.lh * rhos[2]);
    const double X = (-1.0 + sqrt(1.0 + 4.0 * rho_factor)) / (2.0 * rho_factor);
    return rho_factor >= 0 ? log_series(X) : 0.0;
}

// f' of every species at one cell; X (Generic) in/out
__device__ void sp_fprime(const SpModel &m, const double *rhos, double *X, double *out) {
#pragma clang fp contract(off)
    const int S = m.S;
    double logX[kSpMaxPatches];
    if(m.model == CH_GENERIC_WERTHEIM) {
        generic_update_X(m, rhos, X);
        for(int i = 0; i < m.P; i++) logX[i] = log_series(X[i]);
    }
    for(int s = 0; s < S; s++) {
        out[s] = 0.0;
        if(rhos[s] == 0.) continue;
        double rho_tot = 0.;
        for(int t = 0; t < S; t++) rho_tot += rhos[t];
        double ref, bond = 0.0;
        if(m.model == CH_SALEH_WERTHEIM) {
            ref = log_series(rhos[s]) + 2.0 * m.sB2 * rho_tot + 3.0 * m.sB3 * (rho_tot * rho_tot);
            if(s == 0) bond = m.val[0] * saleh_contribution(m, rhos, 0);
            else if(s == 1) bond = m.val[1] * saleh_contribution(m, rhos, 1);
            else bond = m.lh * (saleh_contribution(m, rhos, 0) + saleh_contribution(m, rhos, 1));
        }
        else if(m.model == CH_RICCI_WERTHEIM) {
            ref = log_series(rhos[s]) + 2.0 * m.rB2 * rho_tot;
            const double A = rhos[1] * m.d12, B = rhos[0] * m.d12;
            if(s == 0) {
                const double X0 = (-1 + sqrt((1 + 12 * rhos[0] * m.d00))) / (6 * rhos[0] * m.d00);
                const double u = 1 + 2 * A - B;
                const double X1 = (B - 1 - 2 * A + sqrt(u * u + 4 * B)) / (2 * B);
                bond = 3 * log_series(X0) + log_series(X1);
            }
            else {
                const double v = B - 2 * A + 1;
                const double X2 = (-B + 2 * A - 1 + sqrt(v * v + 8 * A)) / (4 * A);
                bond = 2 * log_series(X2);
            }
        }
        else {
            double B2c = 0.;
            for(int o = 0; o < S; o++) B2c += 2.0 * rhos[o] * m.B2[s][o];
            ref = log_series(rhos[s]) + B2c;
            for(int q = 0; q < m.P; q++)
                if(m.pc[s][q] > 0) bond += m.pc[s][q] * logX[q];
        }
        out[s] = ref + bond;
    }
}

// bulk free-energy density at one cell (SalehWertheim.cpp:47-78,
// RicciWertheim.cpp:43-70, GenericWertheim.cpp:210-248; Generic X from 0)
__device__ double sp_bulk(const SpModel &m, double dcf, const double *rhos) {
#pragma clang fp contract(off)
    const int S = m.S;
    double rho = 0.;
    for(int i = 0; i < S; i++) rho += rhos[i];
    double mixing_S = 0., B2c = 0.;
    for(int i = 0; i < S; i++) {
        const double x_i = rhos[i] / rho;
        mixing_S += (x_i > 0) ? x_i * log(x_i) : 0;
        if(m.model == CH_GENERIC_WERTHEIM)
            for(int j = 0; j < S; j++) B2c += rhos[i] * rhos[j] * m.B2[i][j];
    }
    auto term = [](double X) { return log(X) - X / 2.0 + 0.5; };
    if(m.model == CH_SALEH_WERTHEIM) {
        const double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S + m.sB2 * rho + 0.5 * m.sB3 * (rho * rho));
        double rf = m.dAA * (m.val[0] * rhos[0] + m.lh * rhos[2]);
        const double X1A = (-1.0 + sqrt(1.0 + 4.0 * rf)) / (2.0 * rf);
        const double p1 = term(X1A);
        rf = m.dBB * (m.val[1] * rhos[1] + m.lh * rhos[2]);
        const double X2B = (-1.0 + sqrt(1.0 + 4.0 * rf)) / (2.0 * rf);
        const double p2 = term(X2B);
        return f_ref + (rhos[0] * m.val[0] * p1 + rhos[1] * m.val[1] * p2 + rhos[2] * m.lh * (p1 + p2));
    }
    if(m.model == CH_RICCI_WERTHEIM) {
        const double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S + m.rB2 * rho);
        const double X0 = (-1 + sqrt((1 + 12 * rhos[0] * m.d00))) / (6 * rhos[0] * m.d00);
        const double A = rhos[1] * m.d12, B = rhos[0] * m.d12;
        const double u = 1 + 2 * A - B, v = B - 2 * A + 1;
        const double X1 = (B - 1 - 2 * A + sqrt(u * u + 4 * B)) / (2 * B);
        const double X2 = (-B + 2 * A - 1 + sqrt(v * v + 8 * A)) / (4 * A);
        double bond = rhos[0] * (3 * (log(X0) - X0 / 2) + log(X1) - X1 / 2 + 2);
        bond += rhos[1] * (2 * log(X2) - X2 + 1);
        return f_ref + bond;
    }
    const double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S) + B2c;
    double X[kSpMaxPatches];
    for(int q = 0; q < kSpMaxPatches; q++) X[q] = 0.0;
    generic_update_X(m, rhos, X);
    double bond = 0;
    for(int s = 0; s < S; s++) {
        double sf = 0;
        for(int q = 0; q < m.P; q++)
            if(m.pc[s][q] > 0) sf += m.pc[s][q] * term(X[q]);
        bond += rhos[s] * sf;
    }
    return f_ref + bond;
}

// mobility of species s (RegularisedMobility.h:26-33; GenericWertheim.cpp:276-291)
__device__ double sp_mobility(const Params &p, const SpModel &m, int mob, int s, double r, const double *X) {
#pragma clang fp contract(off)
    switch(mob) {
    case CH_MOB_CONSTANT: return p.M;
    case CH_MOB_REGULARISED: return p.M * r / (r + p.rho_min);
    case CH_MOB_LINEAR: return p.M * r / p.rho_ref;
    default: {
        double D = p.M;
        for(int q = 0; q < m.P; q++)
            if(m.pc[s][q] > 0) D *= pow(X[q], (double)m.pc[s][q]);
        return r * D;
    }
    }
}

// cell index -> the 2 DIMS neighbour indices (-x, +x, -y, +y, -z, +z)
template <int DIMS>
__device__ __forceinline__ void neighbours(const Params &p, int64_t i, int64_t (&m)[DIMS], int64_t (&q)[DIMS]) {
    const int N = p.N, L = p.log2N, mask = p.mask;
    const int x = (int)(i & mask);
    const int y = DIMS >= 2 ? (int)((i >> L) & mask) : 0;
    const int z = DIMS == 3 ? (int)(i >> (2 * L)) : 0;
    auto idx = [&](int a, int b, int c) { return (int64_t)a + ((int64_t)b << L) + ((int64_t)c << (2 * L)); };
    m[0] = idx((x - 1) & mask, y, z);
    q[0] = idx((x + 1) & mask, y, z);
    if constexpr(DIMS >= 2) {
        m[1] = idx(x, (y - 1) & mask, z);
        q[1] = idx(x, (y + 1) & mask, z);
    }
    if constexpr(DIMS == 3) {
        m[2] = idx(x, y, (z - 1) & mask);
        q[2] = idx(x, y, (z + 1) & mask);
    }
    (void)N;
}

#define SP_CELLS for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < bins; i += (int64_t)gridDim.x * 256)

// X update only: the model evaluation of CahnHilliard.cpp:226-228 at set-up
template <typename T>
__global__ __launch_bounds__(256) void k_sp_warm(const T *__restrict__ rho, double *__restrict__ Xf,
                                                 const SpModel *__restrict__ mp, int64_t bins) {
    const SpModel &m = *mp;
    SP_CELLS {
        double rhos[kSpMaxSpecies], X[kSpMaxPatches], out[kSpMaxSpecies];
        for(int s = 0; s < m.S; s++) rhos[s] = (double)rho[s * bins + i];
        for(int q = 0; q < m.P; q++) X[q] = Xf[q * bins + i];
        sp_fprime(m, rhos, X, out);
        for(int q = 0; q < m.P; q++) Xf[q * bins + i] = X[q];
    }
}

// mu_s = f'_s - 2 k_s lap rho_s (and M_s from the X before this evaluation)
template <typename T, int DIMS>
__global__ __launch_bounds__(256) void k_sp_mu(const T *__restrict__ rho, double *__restrict__ Xf, T *__restrict__ mu,
                                               T *__restrict__ mobf, Params p, const SpModel *__restrict__ mp,
                                               int64_t bins, int mob) {
#pragma clang fp contract(off)
    const SpModel &m = *mp;
    SP_CELLS {
        double rhos[kSpMaxSpecies], X[kSpMaxPatches], f[kSpMaxSpecies];
        for(int s = 0; s < m.S; s++) rhos[s] = (double)rho[s * bins + i];
        for(int q = 0; q < m.P; q++) X[q] = Xf[q * bins + i];
        if(mob >= 0)
            for(int s = 0; s < m.S; s++) mobf[s * bins + i] = (T)sp_mobility(p, m, mob, s, rhos[s], X);
        sp_fprime(m, rhos, X, f);
        for(int q = 0; q < m.P; q++) Xf[q * bins + i] = X[q];
        int64_t nm[DIMS], nq[DIMS];
        neighbours<DIMS>(p, i, nm, nq);
        const T inv_dx2 = (T)p.inv_dx2;
        for(int s = 0; s < m.S; s++) {
            const T *r = rho + s * bins;
            T a[DIMS], b[DIMS];
            for(int d = 0; d < DIMS; d++) {
                a[d] = r[nm[d]];
                b[d] = r[nq[d]];
            }
            const T lap = laplacian<T, DIMS>(r[i], a, b, inv_dx2);
            mu[s * bins + i] = (T)f[s] - (T)(2 * m.k[s]) * lap;
        }
    }
}

// rho_s += dt M lap mu_s (euler) or dt div J_s (euler_mobility), in place
template <typename T, int DIMS, bool EM>
__global__ __launch_bounds__(256) void k_sp_update(T *__restrict__ rho, const T *__restrict__ mu,
                                                   const T *__restrict__ mobf, Params p, int S, int64_t bins) {
#pragma clang fp contract(off)
    SP_CELLS {
        int64_t nm[DIMS], nq[DIMS];
        neighbours<DIMS>(p, i, nm, nq);
        for(int s = 0; s < S; s++) {
            const T *u = mu + s * bins;
            T a[DIMS], b[DIMS];
            for(int d = 0; d < DIMS; d++) {
                a[d] = u[nm[d]];
                b[d] = u[nq[d]];
            }
            T *r = rho + s * bins;
            if constexpr(!EM) {
                r[i] = euler_update<T, DIMS>(p, r[i], u[i], a, b);
            }
            else {
                const T *M = mobf + s * bins;
                T Ma[DIMS], Mb[DIMS];
                for(int d = 0; d < DIMS; d++) {
                    Ma[d] = M[nm[d]];
                    Mb[d] = M[nq[d]];
                }
                r[i] = mobility_update<T, DIMS>(p, r[i], u[i], a, b, M[i], Ma, Mb);
            }
        }
    }
}

// f'_s into out (X updated), the pseudospectral path's model evaluation
template <typename T>
__global__ __launch_bounds__(256) void k_sp_fprime(const T *__restrict__ rho, double *__restrict__ Xf,
                                                   T *__restrict__ out, const SpModel *__restrict__ mp, int64_t bins) {
    const SpModel &m = *mp;
    SP_CELLS {
        double rhos[kSpMaxSpecies], X[kSpMaxPatches], f[kSpMaxSpecies];
        for(int s = 0; s < m.S; s++) rhos[s] = (double)rho[s * bins + i];
        for(int q = 0; q < m.P; q++) X[q] = Xf[q * bins + i];
        sp_fprime(m, rhos, X, f);
        for(int q = 0; q < m.P; q++) Xf[q * bins + i] = X[q];
        for(int s = 0; s < m.S; s++) out[s * bins + i] = (T)f[s];
    }
}

// mobility field of every species from rho and the current X
template <typename T>
__global__ __launch_bounds__(256) void k_sp_mob(const T *__restrict__ rho, const double *__restrict__ Xf,
                                                double *__restrict__ out, Params p, const SpModel *__restrict__ mp,
                                                int64_t bins, int mob) {
    const SpModel &m = *mp;
    SP_CELLS {
        double X[kSpMaxPatches];
        for(int q = 0; q < m.P; q++) X[q] = Xf[q * bins + i];
        for(int s = 0; s < m.S; s++) out[s * bins + i] = sp_mobility(p, m, mob, s, (double)rho[s * bins + i], X);
    }
}
#undef SP_CELLS

// the pseudospectral update of one species (PseudospectralCPU.cpp:118-133)
template <typename T>
__global__ __launch_bounds__(256) void k_sp_ps_update(C2<T> *__restrict__ rh, C2<T> *__restrict__ fh,
                                                      C2<T> *__restrict__ copy, Params p, int64_t hat, double ks) {
#pragma clang fp contract(off)
    for(int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < hat; q += (int64_t)gridDim.x * 256) {
        int ix, iy, iz;
        mode_index(p, q, ix, iy, iz);
        double k2, deal;
        mode_q2(p, ix, iy, iz, k2, deal);
        C2<T> f = fh[q];
        if(p.use_dealias) f = C2<T>{f.x * (T)deal, f.y * (T)deal};
        const double k4 = k2 * k2;
        const T den = (T)(1.0 + p.dt * p.M * (p.S * k2 + 2.0 * ks * k4));
        const T a = (T)(p.dt * p.M * k2), S = (T)p.S, ib = (T)p.inv_bins;
        const C2<T> r = rh[q];
        const T re = (r.x - a * (f.x - S * r.x)) / den;
        const T im = (r.y - a * (f.y - S * r.y)) / den;
        rh[q] = C2<T>{re, im};
        copy[q] = C2<T>{re * ib, im * ib};
    }
}

// FP64 mass and free energy partials over cells (CahnHilliard.cpp:285-319)
template <typename T, int DIMS>
__global__ __launch_bounds__(256) void k_sp_observables(const T *__restrict__ rho, Params p,
                                                        const SpModel *__restrict__ mp, int64_t bins,
                                                        double *__restrict__ partials) {
#pragma clang fp contract(off)
    const SpModel &m = *mp;
    __shared__ double red[2][256];
    double ms = 0., fe = 0.;
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < bins; i += (int64_t)gridDim.x * 256) {
        double rhos[kSpMaxSpecies], cell = 0.;
        for(int s = 0; s < m.S; s++) {
            rhos[s] = (double)rho[s * bins + i];
            cell += rhos[s];
        }
        ms += cell;
        int64_t nm[DIMS], nq[DIMS];
        neighbours<DIMS>(p, i, nm, nq);
        double inter = 0.;
        for(int s = 0; s < m.S; s++)
            for(int d = 0; d < DIMS; d++) {
                const double g = ((double)rho[s * bins + nq[d]] - rhos[s]) / p.dx;
                inter += m.k[s] * g * g;
            }
        fe += sp_bulk(m, p.dcf, rhos) + inter;
    }
    red[0][threadIdx.x] = ms;
    red[1][threadIdx.x] = fe;
    __syncthreads();
    for(int k = 128; k > 0; k >>= 1) {
        if(threadIdx.x < k) {
            red[0][threadIdx.x] += red[0][threadIdx.x + k];
            red[1][threadIdx.x] += red[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if(threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0];
        partials[2 * blockIdx.x + 1] = red[1][0];
    }
}

template <int DIMS, typename F>
hipError_t by_precision(int precision, F &&f) {
    if(precision == CH_FP64) return f(double{});
    return f(float{});
}

int grid_for(int64_t n) { return (int)std::min<int64_t>(kSpBlocks, (n + 255) / 256); }

}  // namespace

void sp_build_model(const ch_species_model &in, SpModel &m) {
    m = SpModel{};
    m.model = in.model;
    m.S = in.n_species;
    for(int s = 0; s < kSpMaxSpecies; s++) m.k[s] = in.k[s];
    for(int i = 0; i < 3; i++) m.val[i] = in.saleh_valence[i];
    m.lh = in.saleh_valence[2] / 2;
    m.sB2 = in.saleh_B2;
    m.sB3 = in.saleh_B3;
    m.dAA = in.saleh_delta_AA;
    m.dBB = in.saleh_delta_BB;
    m.rB2 = in.ricci_B2;
    m.d00 = in.ricci_delta_00;
    m.d12 = in.ricci_delta_12;
    m.P = in.model == CH_GENERIC_WERTHEIM ? in.n_patches : 0;
    for(int s = 0; s < kSpMaxSpecies; s++)
        for(int q = 0; q < kSpMaxPatches; q++) m.pc[s][q] = in.model == CH_GENERIC_WERTHEIM ? in.patch_count[s][q] : 0;
    for(int s = 0; s < kSpMaxSpecies; s++)
        for(int t = 0; t < kSpMaxSpecies; t++) m.B2[s][t] = in.B2[s][t];
    if(in.model != CH_GENERIC_WERTHEIM) return;
    // GenericWertheim.cpp:123-149: unique patch types ascending; per type the
    // terms over species in order and their patch types ascending
    int used[kSpMaxPatches] = {0};
    for(int s = 0; s < m.S; s++)
        for(int q = 0; q < m.P; q++)
            if(m.pc[s][q] > 0) used[q] = 1;
    int nt = 0;
    for(int q = 0; q < m.P; q++) {
        if(!used[q]) continue;
        const int u = m.n_unique++;
        m.unique[u] = q;
        m.term_start[u] = nt;
        for(int s = 0; s < m.S; s++)
            for(int o = 0; o < m.P; o++)
                if(m.pc[s][o] > 0 && in.delta[q][o] > 0.0) {
                    m.t_species[nt] = s;
                    m.t_other[nt] = o;
                    m.t_coeff[nt] = (double)m.pc[s][o] * in.delta[q][o];
                    nt++;
                }
    }
    m.term_start[m.n_unique] = nt;
}

hipError_t sp_warm(const SpArgs &a, hipStream_t s) {
    if(a.m_host->model != CH_GENERIC_WERTHEIM) return hipSuccess;
    return by_precision<1>(a.precision, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL((k_sp_warm<T>), dim3(grid_for(a.bins)), dim3(256), 0, s, (const T *)a.rho, a.X, a.m, a.bins);
        return hipGetLastError();
    });
}

hipError_t sp_step_stencil(const SpArgs &a, int integrator, hipStream_t s) {
    return by_precision<1>(a.precision, [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        const int g = grid_for(a.bins);
        const int mob = integrator == CH_EULER_MOBILITY ? a.mob_rule : -1;
        auto go = [&](auto dims_tag) {
            constexpr int D = decltype(dims_tag)::value;
            hipLaunchKernelGGL((k_sp_mu<T, D>), dim3(g), dim3(256), 0, s, (const T *)a.rho, a.X, (T *)a.mu, (T *)a.mob,
                               a.p, a.m, a.bins, mob);
            if(integrator == CH_EULER_MOBILITY)
                hipLaunchKernelGGL((k_sp_update<T, D, true>), dim3(g), dim3(256), 0, s, (T *)a.rho, (const T *)a.mu,
                                   (const T *)a.mob, a.p, a.m_host->S, a.bins);
            else
                hipLaunchKernelGGL((k_sp_update<T, D, false>), dim3(g), dim3(256), 0, s, (T *)a.rho, (const T *)a.mu,
                                   (const T *)a.mob, a.p, a.m_host->S, a.bins);
            return hipGetLastError();
        };
        if(a.p.dims == 1) return go(std::integral_constant<int, 1>{});
        if(a.p.dims == 2) return go(std::integral_constant<int, 2>{});
        return go(std::integral_constant<int, 3>{});
    });
}

hipError_t sp_step_spectral(const SpArgs &a, const SpectralArgs &sa, hipStream_t s) {
    return by_precision<1>(a.precision, [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        const int S = a.m_host->S;
        hipLaunchKernelGGL((k_sp_fprime<T>), dim3(grid_for(a.bins)), dim3(256), 0, s, (const T *)a.rho, a.X, (T *)a.mu,
                           a.m, a.bins);
        hipError_t e = hipGetLastError();
        for(int sp = 0; sp < S && e == hipSuccess; sp++) {
            SpectralArgs b = sa;
            b.rho = (T *)a.rho + sp * a.bins;
            C2<T> *rh = (C2<T> *)a.rho_hat + sp * a.hat;
            e = spec_r2c(b, (const T *)a.mu + sp * a.bins, a.spec[0], false, s);
            if(e != hipSuccess) break;
            hipLaunchKernelGGL((k_sp_ps_update<T>), dim3(grid_for(a.hat)), dim3(256), 0, s, rh, (C2<T> *)a.spec[0],
                               (C2<T> *)a.spec[1], a.p, a.hat, a.m_host->k[sp]);
            e = hipGetLastError();
            if(e == hipSuccess) e = spec_c2r(b, a.spec[1], (T *)a.rho + sp * a.bins, s);
        }
        return e;
    });
}

hipError_t sp_init_spectral(const SpArgs &a, const SpectralArgs &sa, hipStream_t s) {
    return by_precision<1>(a.precision, [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        hipError_t e = hipSuccess;
        for(int sp = 0; sp < a.m_host->S && e == hipSuccess; sp++)
            e = spec_r2c(sa, (const T *)a.rho + sp * a.bins, (C2<T> *)a.rho_hat + sp * a.hat, false, s);
        return e;
    });
}

hipError_t sp_mobility_field(const SpArgs &a, double *out, hipStream_t s) {
    return by_precision<1>(a.precision, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL((k_sp_mob<T>), dim3(grid_for(a.bins)), dim3(256), 0, s, (const T *)a.rho, a.X, out, a.p, a.m,
                           a.bins, a.mob_rule);
        return hipGetLastError();
    });
}

hipError_t sp_observables(const SpArgs &a, double *partials, int nparts, double *result, hipStream_t s, double *ring,
                          unsigned *count) {
    return by_precision<1>(a.precision, [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        const int nb = (int)std::min<int64_t>(nparts, (a.bins + 255) / 256);
        auto go = [&](auto dims_tag) {
            constexpr int D = decltype(dims_tag)::value;
            hipLaunchKernelGGL((k_sp_observables<T, D>), dim3(nb), dim3(256), 0, s, (const T *)a.rho, a.p, a.m, a.bins,
                               partials);
            return hipGetLastError();
        };
        hipError_t e = a.p.dims == 1 ? go(std::integral_constant<int, 1>{})
                                     : (a.p.dims == 2 ? go(std::integral_constant<int, 2>{})
                                                      : go(std::integral_constant<int, 3>{}));
        if(e != hipSuccess) return e;
        return ring ? finalize_pairs_ring(partials, nb, ring, count, s) : finalize_pairs(partials, nb, result, s);
    });
}

}  // namespace chd
