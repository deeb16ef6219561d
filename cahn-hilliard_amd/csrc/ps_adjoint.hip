// ps_adjoint.hip -- reverse step of the pseudospectral integrator.
//
// On the real field one step (src/integrators/PseudospectralCPU.cpp:112-136) is
//     rho_{n+1} = F^-1 ( A F rho_n + B F f'(rho_n) ),
//     A = (1 + dt M S q^2) / D,  B = -dt M q^2 deal / D,  D = 1 + dt M (S q^2 + 2 k q^4),
// with F the unnormalised r2c and F^-1 = c2r / N^d.  A and B are real and even
// in q, so both operators are symmetric and the transpose is
//     lambda_n = a + f''(rho_n) b,   a = F^-1 (A F lambda_{n+1}),  b = F^-1 (B F lambda_{n+1}).
// Parameter gradients, summed over steps:
//     dL/dk  += <lambda', F^-1(-2 dt M q^4 rho_hat' / D)>
//     dL/dM  += <lambda', F^-1((-dt q^2 (deal f_hat - S rho_hat) - dt (S q^2 + 2 k q^4) rho_hat') / D)>
//     dL/deps = -<b, rho_n> (Landau),  dL/dB2 = 2 <b, rho_n>,  dL/ddelta = <b, d f'/d delta> (SimpleWertheim)
// where rho_hat' = A rho_hat + B f_hat is the step's new spectrum and
// <x, F^-1 X> = (1/N^d) sum over the full spectrum of conj(x_hat) X (weight 2
// on the half spectrum off the kx = 0 and kx = N/2 planes).  One step costs
// three r2c (lambda', rho_n, f'(rho_n)), one element-wise pass over the
// spectrum, two c2r and one element-wise pass over the field.  No reference
// counterpart; pinned by the oracle's FP64 adjoint and by finite differences.
#include "adjoint_math.h"
#include "ch_internal.h"

namespace chd {

// accumulates the block's sums of two per-thread values into slots s0, s1
__device__ __forceinline__ void add_block_pair(double a, double b, double *__restrict__ slots, int s0, int s1) {
    __shared__ double red[2][4];
    a = wave_sum64(a);
    b = wave_sum64(b);
    const int w = threadIdx.x >> 6;
    if((threadIdx.x & 63) == 0) {
        red[0][w] = a;
        red[1][w] = b;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        slots[s0] += (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        slots[s1] += (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// L = lambda_hat', R = rho_hat_n -> A lambda_hat' / N^d, Fh = f_hat_n -> B lambda_hat' / N^d
template <typename T>
__global__ __launch_bounds__(256) void k_psadj_modes(const typename Complex<T>::type *__restrict__ L,
                                                     typename Complex<T>::type *__restrict__ R,
                                                     typename Complex<T>::type *__restrict__ Fh, Params p, int64_t hat,
                                                     double *__restrict__ gacc) {
    using T2 = typename Complex<T>::type;
    const double dt = p.dt, M = p.M, S = p.S, inv_bins = p.inv_bins;
    double gk = 0, gm = 0;
    for(int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < hat; q += (int64_t)gridDim.x * 256) {
        int ix, iy, iz;
        mode_index(p, q, ix, iy, iz);
        double k2, deal;
        mode_q2(p, ix, iy, iz, k2, deal);
        if(!p.use_dealias) deal = 1.0;
        const double k4 = k2 * k2;
        const double D = 1.0 + dt * M * (S * k2 + 2.0 * p.k * k4);
        const double A = (1.0 + dt * M * S * k2) / D, B = -dt * M * k2 * deal / D;
        const T2 r = R[q], f = Fh[q], l = L[q];
        const double nr = A * r.x + B * f.x, ni = A * r.y + B * f.y;  // rho_hat_{n+1}
        const double w = ((ix == 0 || ix == p.N / 2) ? 1.0 : 2.0) * inv_bins;
        const double dk = -2.0 * dt * M * k4 / D;
        gk += w * dk * ((double)l.x * nr + (double)l.y * ni);
        const double c = dt * (S * k2 + 2.0 * p.k * k4);
        const double xr = (-dt * k2 * (deal * f.x - S * r.x) - c * nr) / D;
        const double xi = (-dt * k2 * (deal * f.y - S * r.y) - c * ni) / D;
        gm += w * ((double)l.x * xr + (double)l.y * xi);
        R[q] = T2{(T)(A * inv_bins * l.x), (T)(A * inv_bins * l.y)};
        Fh[q] = T2{(T)(B * inv_bins * l.x), (T)(B * inv_bins * l.y)};
    }
    add_block_pair(gk, gm, gacc + kGradSlots * blockIdx.x, 1, 2);
}

// lambda_n = a + f''(rho_n) b; <b, rho_n> and <b, d f'/d delta>
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_psadj_real(const T *__restrict__ a, const T *__restrict__ b,
                                                    const T *__restrict__ rho, T *__restrict__ lam_out, Params p,
                                                    int64_t bins, double *__restrict__ gacc) {
    double ge = 0, gd = 0;
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < bins; i += (int64_t)gridDim.x * 256) {
        const T bi = b[i];
        const double r = (double)rho[i];
        lam_out[i] = a[i] + (T)fsecond<MODEL>(p, r) * bi;
        ge += (double)bi * r;
        if constexpr(MODEL == CH_SIMPLE_WERTHEIM) gd += (double)bi * dfprime_ddelta<MODEL>(p, r);
    }
    add_block_pair(ge, gd, gacc + kGradSlots * blockIdx.x, 0, 3);
}

hipError_t ps_adjoint_step(const SpectralArgs &a, const void *lam_in, void *lam_out, const void *rho, void *const spec[3],
                           void *const real[2], double *gacc, hipStream_t s) {
    const Params &p = a.p;
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int64_t hat = bins / p.N * (p.N / 2 + 1);
    hipError_t e = spec_r2c(a, lam_in, spec[0], false, s);
    if(e == hipSuccess) e = spec_r2c(a, rho, spec[1], false, s);
    if(e == hipSuccess) e = spec_r2c(a, rho, spec[2], true, s);
    if(e != hipSuccess) return e;
    auto go = [&](auto tag) -> hipError_t {
        using T = decltype(tag);
        using T2 = typename Complex<T>::type;
        hipLaunchKernelGGL((k_psadj_modes<T>), dim3(kPsGradBlocks), dim3(256), 0, s, (const T2 *)spec[0], (T2 *)spec[1],
                           (T2 *)spec[2], p, hat, gacc);
        hipError_t e2 = hipGetLastError();
        if(e2 == hipSuccess) e2 = spec_c2r(a, spec[1], real[0], s);
        if(e2 == hipSuccess) e2 = spec_c2r(a, spec[2], real[1], s);
        if(e2 != hipSuccess) return e2;
        if(a.model == CH_LANDAU)
            hipLaunchKernelGGL((k_psadj_real<T, CH_LANDAU>), dim3(kPsGradBlocks), dim3(256), 0, s, (const T *)real[0],
                               (const T *)real[1], (const T *)rho, (T *)lam_out, p, bins, gacc);
        else
            hipLaunchKernelGGL((k_psadj_real<T, CH_SIMPLE_WERTHEIM>), dim3(kPsGradBlocks), dim3(256), 0, s,
                               (const T *)real[0], (const T *)real[1], (const T *)rho, (T *)lam_out, p, bins, gacc);
        return hipGetLastError();
    };
    if(a.precision == CH_FP64) return go(double{});
    return go(float{});
}

}  // namespace chd
