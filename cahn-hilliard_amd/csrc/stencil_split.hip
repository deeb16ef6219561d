// stencil_split.hip -- the 3D euler / euler_mobility step as two launches per
// step (Params::tile3d == 3):
//   k_split_mu      mu = f'(psi) - 2k lap(psi) of every cell -> work field
//   k_split_update  psi + dt M lap(mu), or psi + dt div J with the face fluxes
//                   of EulerMobilityCPU.cpp:44-82, the mobility recomputed
//                   from psi at the cell and its neighbours
// One thread per cell, neighbours through the caches (blocks in XCD-aware
// order, so an XCD's cells form a contiguous slab whose z-neighbour planes stay
// in its L2).  Each cell's f' is evaluated once per step, where the fused LDS
// tile kernel (k_stencil) evaluates it on a halo-extended box (18 x 10 x 10 for
// a 16 x 8 x 8 tile): for the Wertheim models' FP64 logarithms and square
// roots that recomputation, not memory, bounds the fused kernel.  The price is
// a write and a read of mu (40 B per cell and step in place of 16).  The
// per-cell arithmetic is ch_device.h's, so the fields are bit-identical to the
// fused kernel's.  Free-energy mobility (X^valence per cell) stays on the fused
// kernel, where each cell's mobility is evaluated once.
#include "ch_internal.h"

namespace chd {

__device__ __forceinline__ void cell3(int64_t i, int lg, int mask, int &x, int &y, int &z) {
    x = (int)(i & mask);
    y = (int)((i >> lg) & mask);
    z = (int)(i >> (2 * lg));
}

// the 6 neighbours of cell (x, y, z) of field f, order -x, +x, -y, +y, -z, +z
template <typename T>
__device__ __forceinline__ void nb6(const T *__restrict__ f, int x, int y, int z, int lg, int mask, T (&m)[3],
                                    T (&q)[3]) {
    const int64_t row = ((int64_t)z << (2 * lg)) + ((int64_t)y << lg);
    m[0] = f[row + ((x - 1) & mask)];
    q[0] = f[row + ((x + 1) & mask)];
    const int64_t pl = (int64_t)z << (2 * lg);
    m[1] = f[pl + ((int64_t)((y - 1) & mask) << lg) + x];
    q[1] = f[pl + ((int64_t)((y + 1) & mask) << lg) + x];
    const int64_t ry = ((int64_t)y << lg) + x;
    m[2] = f[((int64_t)((z - 1) & mask) << (2 * lg)) + ry];
    q[2] = f[((int64_t)((z + 1) & mask) << (2 * lg)) + ry];
}

template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_split_mu(const T *__restrict__ psi, T *__restrict__ mu, Params p) {
    const int64_t i = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    int x, y, z;
    cell3(i, p.log2N, p.mask, x, y, z);
    T m[3], q[3];
    nb6(psi, x, y, z, p.log2N, p.mask, m, q);
    mu[i] = chem_pot<T, 3, MODEL>(p, psi[i], m, q);
}

template <typename T, int MOB, bool EM>
__global__ __launch_bounds__(256) void k_split_update(const T *__restrict__ psi, const T *__restrict__ mu,
                                                      T *__restrict__ out, Params p) {
    const int64_t i = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    int x, y, z;
    cell3(i, p.log2N, p.mask, x, y, z);
    T m[3], q[3];
    nb6(mu, x, y, z, p.log2N, p.mask, m, q);
    const T c = psi[i];
    if constexpr(!EM) {
        out[i] = euler_update<T, 3>(p, c, mu[i], m, q);
    }
    else if constexpr(MOB == CH_MOB_CONSTANT) {
        const T Mc = (T)p.M;
        const T Mm[3] = {Mc, Mc, Mc}, Mq[3] = {Mc, Mc, Mc};
        out[i] = mobility_update<T, 3>(p, c, mu[i], m, q, Mc, Mm, Mq);
    }
    else {
        T rm[3], rq[3], Mm[3], Mq[3];
        nb6(psi, x, y, z, p.log2N, p.mask, rm, rq);
#pragma unroll
        for(int d = 0; d < 3; d++) {
            Mm[d] = mobility<T, MOB>(p, rm[d]);
            Mq[d] = mobility<T, MOB>(p, rq[d]);
        }
        out[i] = mobility_update<T, 3>(p, c, mu[i], m, q, mobility<T, MOB>(p, c), Mm, Mq);
    }
}

template <typename T, int MODEL>
static hipError_t split_t(const Params &p, int integrator, int mob, const void *in, void *out, void *work,
                          hipStream_t s) {
    const int64_t bins = (int64_t)1 << (3 * p.log2N);
    const int nb = (int)(bins / 256);
    const T *a = (const T *)in;
    T *b = (T *)out, *w = (T *)work;
    hipLaunchKernelGGL((k_split_mu<T, MODEL>), dim3(nb), dim3(256), 0, s, a, w, p);
    if(integrator == CH_EULER) {
        hipLaunchKernelGGL((k_split_update<T, CH_MOB_CONSTANT, false>), dim3(nb), dim3(256), 0, s, a, w, b, p);
        return hipGetLastError();
    }
    switch(mob) {
    case CH_MOB_CONSTANT:
        hipLaunchKernelGGL((k_split_update<T, CH_MOB_CONSTANT, true>), dim3(nb), dim3(256), 0, s, a, w, b, p);
        break;
    case CH_MOB_REGULARISED:
        hipLaunchKernelGGL((k_split_update<T, CH_MOB_REGULARISED, true>), dim3(nb), dim3(256), 0, s, a, w, b, p);
        break;
    case CH_MOB_LINEAR:
        hipLaunchKernelGGL((k_split_update<T, CH_MOB_LINEAR, true>), dim3(nb), dim3(256), 0, s, a, w, b, p);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool split3d_ok(const Params &p, int integrator, int mob) {
    return p.dims == 3 && p.N >= 16 && (integrator == CH_EULER || mob != CH_MOB_FREE_ENERGY);
}

hipError_t stencil_split3d(const Params &p, int precision, int integrator, int model, int mob, const void *in,
                           void *out, void *work, hipStream_t s) {
    if(!split3d_ok(p, integrator, mob)) return hipErrorInvalidValue;
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return split_t<double, CH_LANDAU>(p, integrator, mob, in, out, work, s);
        return split_t<double, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, out, work, s);
    }
    if(model == CH_LANDAU) return split_t<float, CH_LANDAU>(p, integrator, mob, in, out, work, s);
    return split_t<float, CH_SIMPLE_WERTHEIM>(p, integrator, mob, in, out, work, s);
}

}  // namespace chd
