// stencil_split.hip -- the 3D euler / euler_mobility step as two launches per
// step (Params::tile3d == 3):
//   k_split_mu      mu = f'(psi) - 2k lap(psi) of every cell -> work field
//   k_split_update  psi + dt M lap(mu), or psi + dt div J with the face fluxes
//                   of EulerMobilityCPU.cpp:44-82 from the cells' mobilities
// A thread per column of a few cells along z, neighbours through the caches
// (blocks in XCD-aware order, so an XCD's cells form a contiguous slab whose
// z-neighbour planes stay in its L2).  Each cell's f' is evaluated once per
// step, and a variable mobility once per cell (k_split_mu writes it beside mu;
// 8 B per cell written and 8 read in FP64 in place of 6 more divisions per
// cell in the update: both kernels issue-bound).  The fused LDS
// tile kernel (k_stencil) evaluates it on a halo-extended box (18 x 10 x 10 for
// a 16 x 8 x 8 tile): for the Wertheim models' FP64 logarithms and square
// roots that recomputation, not memory, bounds the fused kernel.  The price is
// a write and a read of mu (40 B per cell and step in place of 16).  The
// per-cell arithmetic is ch_device.h's, so the fields are bit-identical to the
// fused kernel's.  Free-energy mobility (X^valence per cell) stays on the fused
// kernel, where each cell's mobility is evaluated once.
#include "adjoint_math.h"
#include "ch_internal.h"

namespace chd {

// A thread owns a column of Z cells along z at (x, y): the Z + 2 centre loads
// of a field give both the cells and their z neighbours, and every offset is
// 32-bit (N^3 <= 2^30).  Blocks of 256 threads cover 256 consecutive (x, y)
// of one z chunk, (x, y) fastest, in XCD-aware order, so an XCD's blocks form
// a contiguous slab of planes.
template <int Z>
struct ZCol {
    uint32_t pc, z0, x, y, lg, mask;
    __device__ __forceinline__ ZCol(const Params &p) {
        lg = (uint32_t)p.log2N;
        mask = (uint32_t)p.mask;
        const uint32_t per = 1u << (2 * lg - 8);  // blocks per z chunk
        const uint32_t b = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
        pc = (b % per) * 256 + threadIdx.x;
        z0 = (b / per) * Z;
        x = pc & mask;
        y = pc >> lg;
    }
    // offset of plane z0 + j (j = -1 .. Z), wrapping
    __device__ __forceinline__ uint32_t plane(int j) const { return ((z0 + (uint32_t)j) & mask) << (2 * lg); }
    // in-plane offsets of the -x, +x, -y, +y neighbours
    __device__ __forceinline__ void inplane(uint32_t (&o)[4]) const {
        const uint32_t row = y << lg;
        o[0] = row + ((x - 1) & mask);
        o[1] = row + ((x + 1) & mask);
        o[2] = (((y - 1) & mask) << lg) + x;
        o[3] = (((y + 1) & mask) << lg) + x;
    }
};

// a field's column: c[j + 1] = the cell of plane z0 + j (j = -1 .. Z), and the
// in-plane neighbours n[j][0..3] (-x, +x, -y, +y) of the Z cells
template <typename T, int Z>
__device__ __forceinline__ void zcol_load(const T *__restrict__ f, const ZCol<Z> &g, T (&c)[Z + 2], T (&n)[Z][4]) {
    uint32_t o[4];
    g.inplane(o);
#pragma unroll
    for(int j = -1; j <= Z; j++) c[j + 1] = f[g.plane(j) + g.pc];
#pragma unroll
    for(int j = 0; j < Z; j++) {
        const uint32_t pl = g.plane(j);
#pragma unroll
        for(int d = 0; d < 4; d++) n[j][d] = f[pl + o[d]];
    }
}

// the neighbours of cell j in the order -x, +x, -y, +y, -z, +z as m[d] / q[d]
template <typename T, int Z>
__device__ __forceinline__ void zcol_nb(const T (&c)[Z + 2], const T (&n)[Z][4], int j, T (&m)[3], T (&q)[3]) {
    m[0] = n[j][0];
    q[0] = n[j][1];
    m[1] = n[j][2];
    q[1] = n[j][3];
    m[2] = c[j];
    q[2] = c[j + 2];
}

// mu of every cell; with a variable mobility (MOB != CH_MOB_CONSTANT, the
// euler_mobility integrator) also the cell's mobility into `mob`, so that the
// update evaluates it once per cell rather than at each of 7 stencil points
template <typename T, int MODEL, int MOB, int Z>
__global__ __launch_bounds__(256) void k_split_mu(const T *__restrict__ psi, T *__restrict__ mu, T *__restrict__ mob,
                                                  Params p) {
    const ZCol<Z> g(p);
    T c[Z + 2], n[Z][4];
    zcol_load<T, Z>(psi, g, c, n);
#pragma unroll
    for(int j = 0; j < Z; j++) {
        T m[3], q[3];
        zcol_nb<T, Z>(c, n, j, m, q);
        const uint32_t at = g.plane(j) + g.pc;
        mu[at] = chem_pot<T, 3, MODEL>(p, c[j + 1], m, q);
        if constexpr(MOB != CH_MOB_CONSTANT) mob[at] = mobility<T, MOB>(p, c[j + 1]);
    }
}

template <typename T, int MOB, bool EM, int Z>
__global__ __launch_bounds__(256) void k_split_update(const T *__restrict__ psi, const T *__restrict__ mu,
                                                      const T *__restrict__ mob, T *__restrict__ out, Params p) {
    const ZCol<Z> g(p);
    T c[Z + 2], n[Z][4];
    zcol_load<T, Z>(mu, g, c, n);
    constexpr bool VARM = EM && MOB != CH_MOB_CONSTANT;
    T mc[VARM ? Z + 2 : 1], mn[VARM ? Z : 1][4];
    if constexpr(VARM) zcol_load<T, Z>(mob, g, mc, mn);
#pragma unroll
    for(int j = 0; j < Z; j++) {
        T m[3], q[3];
        zcol_nb<T, Z>(c, n, j, m, q);
        const uint32_t at = g.plane(j) + g.pc;
        const T ps = psi[at];
        if constexpr(!EM) {
            out[at] = euler_update<T, 3>(p, ps, c[j + 1], m, q);
        }
        else if constexpr(MOB == CH_MOB_CONSTANT) {
            const T Mc = (T)p.M;
            const T Mm[3] = {Mc, Mc, Mc}, Mq[3] = {Mc, Mc, Mc};
            out[at] = mobility_update<T, 3>(p, ps, c[j + 1], m, q, Mc, Mm, Mq);
        }
        else {
            T Mm[3], Mq[3];
            zcol_nb<T, Z>(mc, mn, j, Mm, Mq);
            out[at] = mobility_update<T, 3>(p, ps, c[j + 1], m, q, mc[j + 1], Mm, Mq);
        }
    }
}

// cells per thread along z (N = 128, FP64, on MI355X: the Wertheim mu pass
// 13.1 us at 2, 13.8 at 4, 15.7 at 8; the linear-mobility update 15.8 at 2
// and 4, 14.7 at 8)
constexpr int kSplitZMu = 2, kSplitZUpd = 8;

template <typename T, int MODEL, int MOB>
static hipError_t split_mob(const Params &p, int integrator, const T *a, T *b, T *w, T *mw, hipStream_t s) {
    constexpr int ZM = kSplitZMu, ZU = kSplitZUpd;
    const int64_t bins = (int64_t)1 << (3 * p.log2N);
    const int nbm = (int)(bins / (256 * ZM)), nbu = (int)(bins / (256 * ZU));
    const bool em = integrator == CH_EULER_MOBILITY;
    if(em && MOB != CH_MOB_CONSTANT) {
        hipLaunchKernelGGL((k_split_mu<T, MODEL, MOB, ZM>), dim3(nbm), dim3(256), 0, s, a, w, mw, p);
        hipLaunchKernelGGL((k_split_update<T, MOB, true, ZU>), dim3(nbu), dim3(256), 0, s, a, w, mw, b, p);
    }
    else {
        hipLaunchKernelGGL((k_split_mu<T, MODEL, CH_MOB_CONSTANT, ZM>), dim3(nbm), dim3(256), 0, s, a, w, nullptr, p);
        if(em)
            hipLaunchKernelGGL((k_split_update<T, CH_MOB_CONSTANT, true, ZU>), dim3(nbu), dim3(256), 0, s, a, w, nullptr, b,
                               p);
        else
            hipLaunchKernelGGL((k_split_update<T, CH_MOB_CONSTANT, false, ZU>), dim3(nbu), dim3(256), 0, s, a, w, nullptr,
                               b, p);
    }
    return hipGetLastError();
}

template <typename T, int MODEL>
static hipError_t split_t(const Params &p, int integrator, int mob, const void *in, void *out, void *work,
                          void *mob_work, hipStream_t s) {
    const T *a = (const T *)in;
    T *b = (T *)out, *w = (T *)work, *mw = (T *)mob_work;
    if(integrator == CH_EULER) return split_mob<T, MODEL, CH_MOB_CONSTANT>(p, integrator, a, b, w, mw, s);
    switch(mob) {
    case CH_MOB_CONSTANT: return split_mob<T, MODEL, CH_MOB_CONSTANT>(p, integrator, a, b, w, mw, s);
    case CH_MOB_REGULARISED: return split_mob<T, MODEL, CH_MOB_REGULARISED>(p, integrator, a, b, w, mw, s);
    case CH_MOB_LINEAR: return split_mob<T, MODEL, CH_MOB_LINEAR>(p, integrator, a, b, w, mw, s);
    default: return hipErrorInvalidValue;
    }
}

bool split3d_ok(const Params &p, int integrator, int mob) {
    return p.dims == 3 && p.N >= 16 && p.N % kSplitZUpd == 0 && p.N % kSplitZMu == 0 &&
           (integrator == CH_EULER || mob != CH_MOB_FREE_ENERGY);
}

bool split3d_needs_mobility(int integrator, int mob) {
    return integrator == CH_EULER_MOBILITY && mob != CH_MOB_CONSTANT;
}

hipError_t stencil_split3d(const Params &p, int precision, int integrator, int model, int mob, const void *in,
                           void *out, void *work, void *mob_work, hipStream_t s) {
    if(!split3d_ok(p, integrator, mob)) return hipErrorInvalidValue;
    if(split3d_needs_mobility(integrator, mob) && !mob_work) return hipErrorInvalidValue;
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return split_t<double, CH_LANDAU>(p, integrator, mob, in, out, work, mob_work, s);
        return split_t<double, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, out, work, mob_work, s);
    }
    if(model == CH_LANDAU) return split_t<float, CH_LANDAU>(p, integrator, mob, in, out, work, mob_work, s);
    return split_t<float, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, out, work, mob_work, s);
}

// ---- the euler_mobility reverse step in three launches (the reverse of the
// two-launch forward; the LDS tile form is k_adjoint_em in adjoint.hip):
//   k_split_mu       mu_n and the cells' mobilities from psi_n
//   k_adj_split_nu   nu = D(Mob, lambda) of every cell
//   k_adj_split_upd  eta from mu_n and lambda, then lambda_n from lambda, nu
//                    (with its Laplacian), eta and psi_n, and the block sums of
//                    the parameter gradients
// Each cell's f', mobility, f'' and mobility derivatives are evaluated once,
// where the tile evaluates f' and the mobility on its halo-extended box; the
// per-cell arithmetic is the tile's (adjoint_math.h), so lambda has the same
// bits, and only the order of the FP64 gradient sums differs.
template <typename T, int MOB, int Z>
__global__ __launch_bounds__(256) void k_adj_split_nu(const T *__restrict__ lam, const T *__restrict__ mob,
                                                      T *__restrict__ nu, Params p) {
    const ZCol<Z> g(p);
    constexpr bool VARM = MOB != CH_MOB_CONSTANT;
    T lc[Z + 2], ln[Z][4];
    T mc[VARM ? Z + 2 : 1], mn[VARM ? Z : 1][4];
    zcol_load<T, Z>(lam, g, lc, ln);
    if constexpr(VARM) zcol_load<T, Z>(mob, g, mc, mn);
    const T inv_dx2 = (T)p.inv_dx2, Mc = (T)p.M;
#pragma unroll
    for(int j = 0; j < Z; j++) {
        T lm[3], lq[3], Mm[3] = {Mc, Mc, Mc}, Mq[3] = {Mc, Mc, Mc};
        zcol_nb<T, Z>(lc, ln, j, lm, lq);
        if constexpr(VARM) zcol_nb<T, Z>(mc, mn, j, Mm, Mq);
        nu[g.plane(j) + g.pc] = adj_nu<T, 3>(lc[j + 1], VARM ? mc[j + 1] : Mc, lm, lq, Mm, Mq, inv_dx2);
    }
}

template <typename T, int MODEL, int MOB, int Z>
__global__ __launch_bounds__(256) void k_adj_split_upd(const T *__restrict__ lam, const T *__restrict__ nu,
                                                       const T *__restrict__ mu, const T *__restrict__ psi,
                                                       T *__restrict__ lam_out, Params p, double *__restrict__ gacc) {
    __shared__ double red[kGradSlots][4];
    const ZCol<Z> g(p);
    T nc[Z + 2], nn[Z][4], pc[Z + 2], pn[Z][4], lc[Z + 2], ln[Z][4], uc[Z + 2], un[Z][4];
    zcol_load<T, Z>(nu, g, nc, nn);
    zcol_load<T, Z>(psi, g, pc, pn);
    zcol_load<T, Z>(lam, g, lc, ln);
    zcol_load<T, Z>(mu, g, uc, un);
    const T inv_dx2 = (T)p.inv_dx2;
    double ge = 0, gk = 0, gm = 0, gd = 0;
#pragma unroll
    for(int j = 0; j < Z; j++) {
        T nm[3], nq[3], pm[3], pq[3], lm[3], lq[3], um[3], uq[3];
        zcol_nb<T, Z>(nc, nn, j, nm, nq);
        zcol_nb<T, Z>(pc, pn, j, pm, pq);
        zcol_nb<T, Z>(lc, ln, j, lm, lq);
        zcol_nb<T, Z>(uc, un, j, um, uq);
        const uint32_t at = g.plane(j) + g.pc;
        const T nv = nc[j + 1], et = adj_eta<T, 3>(uc[j + 1], lc[j + 1], um, uq, lm, lq, inv_dx2);
        const T lnu = laplacian<T, 3>(nv, nm, nq, inv_dx2);
        const double ps = (double)pc[j + 1];
        const MobDeriv<MOB> md = mob_deriv<MOB>(p, ps);
        lam_out[at] = adj_em_step<T>(p, lc[j + 1], (T)fsecond<MODEL>(p, ps), nv, lnu, (T)md.drho, et);
        const double lps = (double)laplacian<T, 3>(pc[j + 1], pm, pq, inv_dx2);
        ge += (double)nv * ps;
        gk += (double)nv * lps;
        gm += (double)et * md.hat;
        if constexpr(MODEL == CH_SIMPLE_WERTHEIM) gd += (double)nv * dfprime_ddelta<MODEL>(p, ps) + (double)et * md.ddelta;
    }
    const double gs[kGradSlots] = {ge, gk, gm, gd};
    block_accumulate<256>(gs, red, gacc + kGradSlots * blockIdx.x);
}

// cells per thread along z of the reverse passes (the update holds two fields'
// columns and the FP64 derivatives: 2 keeps it under 128 VGPRs)
constexpr int kAdjZNu = 8, kAdjZUpd = 2;

template <typename T, int MODEL, int MOB>
static hipError_t adj_split_t(const Params &p, const void *lam_in, void *lam_out, const void *psi, void *const w[4],
                              bool have_mu, double *gacc, hipStream_t s) {
    constexpr int ZM = kSplitZMu, ZN = kAdjZNu, ZU = kAdjZUpd;
    const int64_t bins = (int64_t)1 << (3 * p.log2N);
    const int nbm = (int)(bins / (256 * ZM)), nbn = (int)(bins / (256 * ZN)), nbu = (int)(bins / (256 * ZU));
    const T *ps = (const T *)psi;
    T *mu = (T *)w[0], *mob = (T *)w[1], *nu = (T *)w[2];
    if(!have_mu) hipLaunchKernelGGL((k_split_mu<T, MODEL, MOB, ZM>), dim3(nbm), dim3(256), 0, s, ps, mu, mob, p);
    hipLaunchKernelGGL((k_adj_split_nu<T, MOB, ZN>), dim3(nbn), dim3(256), 0, s, (const T *)lam_in, mob, nu, p);
    hipLaunchKernelGGL((k_adj_split_upd<T, MODEL, MOB, ZU>), dim3(nbu), dim3(256), 0, s, (const T *)lam_in, nu, mu,
                       ps, (T *)lam_out, p, gacc);
    return hipGetLastError();
}

template <typename T, int MODEL>
static hipError_t adj_split_mob(const Params &p, int mob, const void *lam_in, void *lam_out, const void *psi,
                                void *const w[4], bool have_mu, double *gacc, hipStream_t s) {
    switch(mob) {
    case CH_MOB_CONSTANT:
        return adj_split_t<T, MODEL, CH_MOB_CONSTANT>(p, lam_in, lam_out, psi, w, have_mu, gacc, s);
    case CH_MOB_REGULARISED:
        return adj_split_t<T, MODEL, CH_MOB_REGULARISED>(p, lam_in, lam_out, psi, w, have_mu, gacc, s);
    case CH_MOB_LINEAR: return adj_split_t<T, MODEL, CH_MOB_LINEAR>(p, lam_in, lam_out, psi, w, have_mu, gacc, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int MODEL>
static hipError_t mu_only_t(const Params &p, int integrator, int mob, const void *in, void *mu, void *mob_out,
                            hipStream_t s) {
    constexpr int ZM = kSplitZMu;
    const int nbm = (int)(((int64_t)1 << (3 * p.log2N)) / (256 * ZM));
    const T *a = (const T *)in;
    T *w = (T *)mu, *mw = (T *)mob_out;
    const bool varm = split3d_needs_mobility(integrator, mob);
    if(varm && mob == CH_MOB_REGULARISED)
        hipLaunchKernelGGL((k_split_mu<T, MODEL, CH_MOB_REGULARISED, ZM>), dim3(nbm), dim3(256), 0, s, a, w, mw, p);
    else if(varm && mob == CH_MOB_LINEAR)
        hipLaunchKernelGGL((k_split_mu<T, MODEL, CH_MOB_LINEAR, ZM>), dim3(nbm), dim3(256), 0, s, a, w, mw, p);
    else if(!varm)
        hipLaunchKernelGGL((k_split_mu<T, MODEL, CH_MOB_CONSTANT, ZM>), dim3(nbm), dim3(256), 0, s, a, w, nullptr, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t split3d_mu(const Params &p, int precision, int integrator, int model, int mob, const void *in, void *mu,
                      void *mob_out, hipStream_t s) {
    if(!split3d_ok(p, integrator, mob)) return hipErrorInvalidValue;
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return mu_only_t<double, CH_LANDAU>(p, integrator, mob, in, mu, mob_out, s);
        return mu_only_t<double, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, mu, mob_out, s);
    }
    if(model == CH_LANDAU) return mu_only_t<float, CH_LANDAU>(p, integrator, mob, in, mu, mob_out, s);
    return mu_only_t<float, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, mu, mob_out, s);
}

bool adjoint_split3d_ok(const Params &p, int mob) {
    return p.tile3d == 3 && split3d_ok(p, CH_EULER_MOBILITY, mob) && p.N % kAdjZNu == 0;
}

int adjoint_split3d_blocks(const Params &p) { return (int)(((int64_t)1 << (3 * p.log2N)) / (256 * kAdjZUpd)); }

hipError_t adjoint_split3d(const Params &p, int precision, int model, int mob, const void *lam_in, void *lam_out,
                           const void *psi, void *const work[4], bool have_mu, double *gacc, hipStream_t s) {
    if(!adjoint_split3d_ok(p, mob)) return hipErrorInvalidValue;
    if(precision == CH_FP64) {
        if(model == CH_LANDAU)
            return adj_split_mob<double, CH_LANDAU>(p, mob, lam_in, lam_out, psi, work, have_mu, gacc, s);
        return adj_split_mob<double, CH_SIMPLE_WERTHEIM>(p, mob, lam_in, lam_out, psi, work, have_mu, gacc, s);
    }
    if(model == CH_LANDAU) return adj_split_mob<float, CH_LANDAU>(p, mob, lam_in, lam_out, psi, work, have_mu, gacc, s);
    return adj_split_mob<float, CH_SIMPLE_WERTHEIM>(p, mob, lam_in, lam_out, psi, work, have_mu, gacc, s);
}

}  // namespace chd
