// psm.hip -- the pseudospectral_mobility integrator
// (src/integrators/PseudospectralMobilityCPU.cpp).
//
// d rho / dt = div(M(rho) grad mu), mu = f'(rho) - 2k lap rho, on the half
// spectrum of the library's own FFT (row-major (z, y, kx), kx fastest).
//
// Splitting scheme (_evolve_simple, :234-301), one step:
//   M_i = mobility(rho_i), M0 = max(0, max_i M_i)            k_mob_max, k_m0_final
//   f_hat = r2c f'(rho);  mu_hat = (f_hat + 2k q^2 rho_hat) deal
//   div_hat = sum_d (i k_d) deal r2c((M - M0) c2r((i k_d) deal mu_hat) / N^d)
//   rho_hat = (rho_hat + dt div_hat - dt M0 q^2 (f_hat - S rho_hat))
//             / (1 + dt M0 (S q^2 + 2k q^4));   rho = c2r(rho_hat) / N^d
// Everything stays on the device (M0 included), so a chain of steps is
// captured into a graph like the other integrators.
//
// Full scheme (_evolve_full, :303-404): A x = b with
//   A x = x - dt div(M grad((S - 2k lap) x)),  b = rho + dt div(M grad(f' - S rho)),
// solved by restarted GMRES left-preconditioned with the FFT-diagonal
// P = 1 + dt M0 (S q^2 + 2k q^4) (:431-663): modified Gram-Schmidt and Givens
// rotations as the reference.  Inner products are FP64 block partials summed
// in a fixed order (deterministic); the Gram-Schmidt coefficient is summed on
// the device by the kernel that applies it, so an Arnoldi step costs one
// host round trip (the new Hessenberg column, needed by the Givens
// rotations on the host).
//
// i k_d uses the reference's signed wave numbers: the Nyquist index of a full
// axis is -N/2.  The c2r keeps only the real part of the DC and Nyquist
// entries of each x line, as FFTW's does: the half-axis derivative on the
// Nyquist plane then vanishes, as in the reference.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "ch_internal.h"
#include "psm_ops.h"

namespace chd {

namespace {

constexpr int kBlocks = kPsmParts;  // reduction partials and grid of the element-wise kernels

template <typename T>
using C2 = typename Complex<T>::type;

// ---- mobility field and M0 = max(0, max M)
template <typename T, int MOB>
__global__ __launch_bounds__(256) void k_mob_max(const T *__restrict__ rho, T *__restrict__ mob, Params p, int64_t bins,
                                                 double *__restrict__ partials) {
    __shared__ double red[256];
    double m = 0.0;
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < bins; i += (int64_t)gridDim.x * 256) {
        const T v = mobility<T, MOB>(p, rho[i]);
        mob[i] = v;
        m = (double)v > m ? (double)v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1) {
        if(threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x + s] > red[threadIdx.x] ? red[threadIdx.x + s] : red[threadIdx.x];
        __syncthreads();
    }
    if(threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_m0_final(const double *__restrict__ partials, int n, double *__restrict__ M0) {
    __shared__ double red[256];
    double m = 0.0;
    for(int i = threadIdx.x; i < n; i += 256) m = partials[i] > m ? partials[i] : m;
    red[threadIdx.x] = m;
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1) {
        if(threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x + s] > red[threadIdx.x] ? red[threadIdx.x + s] : red[threadIdx.x];
        __syncthreads();
    }
    if(threadIdx.x == 0) *M0 = red[0];
}

// ---- spectral element-wise kernels (one thread per half-spectrum entry)
#define PSM_MODES                                                                                       \
    for(int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < hat; q += (int64_t)gridDim.x * 256) { \
        int ix, iy, iz;                                                                                 \
        mode_index(p, q, ix, iy, iz);                                                                   \
        double k2, deal;                                                                                \
        mode_q2(p, ix, iy, iz, k2, deal);                                                               \
        if(!p.use_dealias) deal = 1.0;

// mu_hat = (f_hat + 2k q^2 rho_hat) deal (:243-249); with `gmres`, the operator
// of A: (S + 2k q^2) x_hat deal (:479-486)
template <typename T>
__global__ __launch_bounds__(256) void k_psm_mu(const C2<T> *__restrict__ fh, const C2<T> *__restrict__ rh,
                                                C2<T> *__restrict__ out, Params p, int64_t hat, int op_A) {
#pragma clang fp contract(off)
    PSM_MODES
        const T d = (T)deal;
        C2<T> v;
        if(op_A) {
            const T c = (T)(p.S + 2.0 * p.k * k2);
            const C2<T> x = rh[q];
            v = C2<T>{c * x.x * d, c * x.y * d};
        }
        else {
            const T c = (T)(2.0 * p.k * k2);
            const C2<T> f = fh[q], r = rh[q];
            v = C2<T>{(f.x + c * r.x) * d, (f.y + c * r.y) * d};
        }
        out[q] = v;
    }
}

// out = (i k_axis) src deal (:252-258)
template <typename T>
__global__ __launch_bounds__(256) void k_psm_grad(const C2<T> *__restrict__ src, C2<T> *__restrict__ out, Params p,
                                                  int64_t hat, int axis) {
    for(int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < hat; q += (int64_t)gridDim.x * 256) {
        int ix, iy, iz;
        mode_index(p, q, ix, iy, iz);
        out[q] = psm_grad_term<T>(p, ix, iy, iz, axis, src[q]);
    }
}

// div_hat (=|+=) (i k_axis) F deal (:278-287)
template <typename T>
__global__ __launch_bounds__(256) void k_psm_divacc(const C2<T> *__restrict__ F, C2<T> *__restrict__ div, Params p,
                                                    int64_t hat, int axis) {
    for(int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < hat; q += (int64_t)gridDim.x * 256) {
        int ix, iy, iz;
        mode_index(p, q, ix, iy, iz);
        div[q] = psm_div_term<T>(p, ix, iy, iz, axis, F[q], axis == 0 ? C2<T>{0, 0} : div[q]);
    }
}

// the splitting update (:290-297); copy = rho_hat / N^d for the c2r
template <typename T>
__global__ __launch_bounds__(256) void k_psm_update(C2<T> *__restrict__ rh, const C2<T> *__restrict__ fh,
                                                    const C2<T> *__restrict__ div, C2<T> *__restrict__ copy, Params p,
                                                    int64_t hat, const double *__restrict__ M0p) {
#pragma clang fp contract(off)
    const double M0 = *M0p;
    PSM_MODES
        (void)deal;
        const double k4 = k2 * k2;
        const T den = (T)(1.0 + p.dt * M0 * (p.S * k2 + 2.0 * p.k * k4));
        const T dt = (T)p.dt, a = (T)(p.dt * M0 * k2), S = (T)p.S, ib = (T)p.inv_bins;
        const C2<T> r = rh[q], f = fh[q], dv = div[q];
        const T re = (r.x + dt * dv.x - a * (f.x - S * r.x)) / den;
        const T im = (r.y + dt * dv.y - a * (f.y - S * r.y)) / den;
        rh[q] = C2<T>{re, im};
        copy[q] = C2<T>{re * ib, im * ib};
    }
}

// the preconditioner's diagonal: x_hat /= 1 + dt M0 (S q^2 + 2k q^4) (:446-454)
template <typename T>
__global__ __launch_bounds__(256) void k_psm_pinv(C2<T> *__restrict__ x, Params p, int64_t hat,
                                                  const double *__restrict__ M0p) {
#pragma clang fp contract(off)
    const double M0 = *M0p;
    PSM_MODES
        (void)deal;
        const T den = (T)(1.0 + p.dt * M0 * (p.S * k2 + 2.0 * p.k * (k2 * k2)));
        x[q] = C2<T>{x[q].x / den, x[q].y / den};
    }
}
#undef PSM_MODES

// ---- real-space element-wise kernels
#define PSM_CELLS for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)

// g = (M - M0) (g / N^d) (splitting) or M (g / N^d) (full), :264-275
template <typename T>
__global__ __launch_bounds__(256) void k_psm_flux(T *__restrict__ g, const T *__restrict__ mob, Params p, int64_t n,
                                                  const double *__restrict__ M0p, int sub) {
#pragma clang fp contract(off)
    const T M0 = sub ? (T)*M0p : (T)0, ib = (T)p.inv_bins;
    PSM_CELLS g[i] = psm_flux<T>(g[i], mob[i], M0, ib);
}

// out = f'(rho) - S rho (:309-314)
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_psm_g(const T *__restrict__ rho, T *__restrict__ out, Params p, int64_t n) {
#pragma clang fp contract(off)
    PSM_CELLS out[i] = fprime<T, MODEL>(p, rho[i]) - (T)p.S * rho[i];
}

// out = x + sign dt (t / N^d): b = rho + dt div (:366-372), A x = x - dt div (:519-525)
template <typename T>
__global__ __launch_bounds__(256) void k_psm_axpy_div(const T *__restrict__ x, const T *__restrict__ t,
                                                      T *__restrict__ out, Params p, int64_t n, T sign) {
#pragma clang fp contract(off)
    const T dt = (T)p.dt, ib = (T)p.inv_bins;
    PSM_CELLS out[i] = x[i] + sign * dt * (t[i] * ib);
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale(const T *__restrict__ x, T *__restrict__ out, T a, int64_t n) {
#pragma clang fp contract(off)
    PSM_CELLS out[i] = x[i] * a;
}

// out = a - b
template <typename T>
__global__ __launch_bounds__(256) void k_sub(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ out,
                                             int64_t n) {
#pragma clang fp contract(off)
    PSM_CELLS out[i] = a[i] - b[i];
}

// FP64 block partials of <a, b>
template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T *__restrict__ a, const T *__restrict__ b, int64_t n,
                                             double *__restrict__ partials) {
    __shared__ double red[256];
    double s = 0.0;
    PSM_CELLS s += (double)a[i] * (double)b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for(int k = 128; k > 0; k >>= 1) {
        if(threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if(threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// the sum of the partials, the same fixed order in every block
__device__ __forceinline__ double sum_partials(const double *__restrict__ partials, int nparts) {
    __shared__ double red[256];
    double s = 0.0;
    for(int i = threadIdx.x; i < nparts; i += 256) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for(int k = 128; k > 0; k >>= 1) {
        if(threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    const double v = red[0];
    __syncthreads();
    return v;
}

// h = <w, v> from the partials; H[slot] = h; w -= h v (:580-584)
template <typename T>
__global__ __launch_bounds__(256) void k_mgs(T *__restrict__ w, const T *__restrict__ v, int64_t n,
                                             const double *__restrict__ partials, int nparts, double *__restrict__ Hslot) {
#pragma clang fp contract(off)
    const double h = sum_partials(partials, nparts);
    if(blockIdx.x == 0 && threadIdx.x == 0) *Hslot = h;
    const T ht = (T)h;
    PSM_CELLS w[i] = w[i] - ht * v[i];
}

// h = |w| from the partials; H[slot] = h; v = w / h when h > 1e-30 (:585-589)
template <typename T>
__global__ __launch_bounds__(256) void k_normalize(const T *__restrict__ w, T *__restrict__ v, int64_t n,
                                                   const double *__restrict__ partials, int nparts,
                                                   double *__restrict__ Hslot) {
#pragma clang fp contract(off)
    const double h = sqrt(sum_partials(partials, nparts));
    if(blockIdx.x == 0 && threadIdx.x == 0) *Hslot = h;
    if(!(h > 1e-30)) return;
    const T ht = (T)h;
    PSM_CELLS v[i] = w[i] / ht;
}

// x += sum_i y_i V_i, i = 0 .. m-1 in order (:635-638)
template <typename T>
__global__ __launch_bounds__(256) void k_update_x(T *__restrict__ x, const T *__restrict__ V, const double *__restrict__ y,
                                                  int m, int64_t n) {
#pragma clang fp contract(off)
    PSM_CELLS {
        T acc = x[i];
        for(int k = 0; k < m; k++) acc = acc + V[(size_t)k * n + i] * (T)y[k];
        x[i] = acc;
    }
}
#undef PSM_CELLS

template <typename T>
struct Ops {
    const PsmArgs &B;
    hipStream_t s;
    int64_t bins, hat;
    int grid_f, grid_h;

    Ops(const PsmArgs &b, hipStream_t st) : B(b), s(st) {
        const Params &p = B.a.p;
        bins = (int64_t)1 << (p.log2N * p.dims);
        hat = bins / p.N * (p.N / 2 + 1);
        grid_f = (int)std::min<int64_t>(kBlocks, (bins + 255) / 256);
        grid_h = (int)std::min<int64_t>(kBlocks, (hat + 255) / 256);
    }
    const Params &p() const { return B.a.p; }
    T *f(void *ptr) const { return (T *)ptr; }
    C2<T> *c(void *ptr) const { return (C2<T> *)ptr; }

    hipError_t r2c(const void *in, void *out, bool fprime = false) { return spec_r2c(B.a, in, out, fprime, s); }
    hipError_t c2r(void *in, void *out) { return spec_c2r(B.a, in, out, s); }

    // M = mobility(rho) into B.mob and M0 into B.scal[0]
    hipError_t mobility_and_m0() {
        auto go = [&](auto tag) {
            constexpr int MOB = decltype(tag)::value;
            hipLaunchKernelGGL((k_mob_max<T, MOB>), dim3(grid_f), dim3(256), 0, s, (const T *)B.a.rho, f(B.mob), p(),
                               bins, B.partials);
        };
        switch(B.mob_rule) {
        case CH_MOB_CONSTANT: go(std::integral_constant<int, CH_MOB_CONSTANT>{}); break;
        case CH_MOB_REGULARISED: go(std::integral_constant<int, CH_MOB_REGULARISED>{}); break;
        case CH_MOB_LINEAR: go(std::integral_constant<int, CH_MOB_LINEAR>{}); break;
        default: go(std::integral_constant<int, CH_MOB_FREE_ENERGY>{}); break;
        }
        hipLaunchKernelGGL(k_m0_final, dim3(1), dim3(256), 0, s, B.partials, grid_f, B.scal);
        return hipGetLastError();
    }

    // B.divh = sum_d (i k_d) deal r2c(w c2r((i k_d) deal src) / N^d), w = M - M0 (sub) or M
    hipError_t div_w_grad(const void *src, int sub) {
        if(B.fuse) {
            const hipError_t ef = psm_div_fused(B.a, src, B.gh, B.divh, B.mob, B.scal, sub, s);
            if(ef != hipErrorNotSupported) return ef;
        }
        hipError_t e = hipSuccess;
        for(int d = 0; d < p().dims && e == hipSuccess; d++) {
            hipLaunchKernelGGL((k_psm_grad<T>), dim3(grid_h), dim3(256), 0, s, (const C2<T> *)src, c(B.gh), p(), hat, d);
            e = c2r(B.gh, B.g);
            if(e != hipSuccess) break;
            hipLaunchKernelGGL((k_psm_flux<T>), dim3(grid_f), dim3(256), 0, s, f(B.g), (const T *)B.mob, p(), bins,
                               B.scal, sub);
            e = r2c(B.g, B.gh);
            if(e != hipSuccess) break;
            hipLaunchKernelGGL((k_psm_divacc<T>), dim3(grid_h), dim3(256), 0, s, (const C2<T> *)B.gh, c(B.divh), p(),
                               hat, d);
            e = hipGetLastError();
        }
        return e;
    }

    hipError_t step_simple() {
        hipError_t e = mobility_and_m0();
        if(e == hipSuccess) e = r2c(B.a.rho, B.fh, true);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_mu<T>), dim3(grid_h), dim3(256), 0, s, (const C2<T> *)B.fh,
                           (const C2<T> *)B.a.rho_hat, c(B.muh), p(), hat, 0);
        e = div_w_grad(B.muh, 1);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_update<T>), dim3(grid_h), dim3(256), 0, s, c(B.a.rho_hat), (const C2<T> *)B.fh,
                           (const C2<T> *)B.divh, c(B.muh), p(), hat, B.scal);
        e = hipGetLastError();
        if(e == hipSuccess) e = c2r(B.muh, B.a.rho);
        return e;
    }

    // Ax = A x
    hipError_t apply_A(const void *x, void *Ax) {
        hipError_t e = r2c(x, B.fh);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_mu<T>), dim3(grid_h), dim3(256), 0, s, (const C2<T> *)nullptr, (const C2<T> *)B.fh,
                           c(B.muh), p(), hat, 1);
        e = div_w_grad(B.muh, 0);
        if(e == hipSuccess) e = c2r(B.divh, B.tr);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_axpy_div<T>), dim3(grid_f), dim3(256), 0, s, (const T *)x, (const T *)B.tr, f(Ax), p(),
                           bins, (T)-1);
        return hipGetLastError();
    }

    // z = P^-1 r
    hipError_t apply_Pinv(const void *r, void *z) {
        hipError_t e = r2c(r, B.fh);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_pinv<T>), dim3(grid_h), dim3(256), 0, s, c(B.fh), p(), hat, B.scal);
        e = c2r(B.fh, B.tr);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_scale<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.tr, f(z), (T)p().inv_bins, bins);
        return hipGetLastError();
    }

    hipError_t apply_M(const void *v, void *out) {
        hipError_t e = apply_A(v, B.av);
        return e == hipSuccess ? apply_Pinv(B.av, out) : e;
    }

    // |a|^2 into B.scal[1] (synchronous), via the partials
    hipError_t dot_host(const void *a, const void *b, double &out) {
        hipLaunchKernelGGL((k_dot<T>), dim3(grid_f), dim3(256), 0, s, (const T *)a, (const T *)b, bins, B.partials);
        double *hp = B.host_partials;
        hipError_t e = hipMemcpyAsync(hp, B.partials, sizeof(double) * grid_f, hipMemcpyDeviceToHost, s);
        if(e == hipSuccess) e = hipStreamSynchronize(s);
        // the device kernels sum partials as a 256-thread tree; on the host the
        // same tree, so the same value
        double red[256];
        for(int t = 0; t < 256; t++) {
            double v = 0.0;
            for(int i = t; i < grid_f; i += 256) v += hp[i];
            red[t] = v;
        }
        for(int k = 128; k > 0; k >>= 1)
            for(int t = 0; t < k; t++) red[t] += red[t + k];
        out = red[0];
        return e;
    }

    T *vec(int k) const { return (T *)B.V + (size_t)k * bins; }

    hipError_t gmres(int &iters) {
        const int m = std::max(2, B.restart);
        const double tol = B.tol;
        iters = 0;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
        auto HH = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
        hipError_t e = apply_Pinv(B.b, B.rhs);
        if(e == hipSuccess) e = apply_M(B.x, B.w);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sub<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.rhs, (const T *)B.w, f(B.r), bins);
        double rn2, beta2;
        if((e = dot_host(B.rhs, B.rhs, rn2)) != hipSuccess) return e;
        const double rhs_norm = std::max(std::sqrt(rn2), 1e-30);
        if((e = dot_host(B.r, B.r, beta2)) != hipSuccess) return e;
        double beta = std::sqrt(beta2), rel = beta / rhs_norm;
        if(rel < tol) return hipSuccess;
        while(iters < B.max_iter) {
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(cs.begin(), cs.end(), 0.0);
            std::fill(sn.begin(), sn.end(), 0.0);
            std::fill(g.begin(), g.end(), 0.0);
            hipLaunchKernelGGL((k_scale<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.r, vec(0), (T)(1.0 / beta),
                               bins);
            g[0] = beta;
            int j_end = 0;
            for(int j = 0; j < m && iters < B.max_iter; j++, iters++) {
                if((e = apply_M(vec(j), B.w)) != hipSuccess) return e;
                for(int i = 0; i <= j; i++) {
                    hipLaunchKernelGGL((k_dot<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.w, (const T *)vec(i),
                                       bins, B.partials);
                    hipLaunchKernelGGL((k_mgs<T>), dim3(grid_f), dim3(256), 0, s, f(B.w), (const T *)vec(i), bins,
                                       B.partials, grid_f, B.Hcol + i);
                }
                hipLaunchKernelGGL((k_dot<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.w, (const T *)B.w, bins,
                                   B.partials);
                hipLaunchKernelGGL((k_normalize<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.w, vec(j + 1), bins,
                                   B.partials, grid_f, B.Hcol + j + 1);
                if((e = hipGetLastError()) != hipSuccess) return e;
                if((e = hipMemcpyAsync(B.host_partials, B.Hcol, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, s)) !=
                   hipSuccess)
                    return e;
                if((e = hipStreamSynchronize(s)) != hipSuccess) return e;
                for(int i = 0; i <= j + 1; i++) HH(i, j) = B.host_partials[i];
                if(!(HH(j + 1, j) > 1e-30)) {
                    j_end = j + 1;
                    break;
                }
                for(int i = 0; i < j; i++) {
                    const double t1 = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                    const double t2 = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                    HH(i, j) = t1;
                    HH(i + 1, j) = t2;
                }
                const double a = HH(j, j), b2 = HH(j + 1, j);
                const double den = std::sqrt(a * a + b2 * b2);
                cs[j] = den > 0 ? a / den : 1.0;
                sn[j] = den > 0 ? b2 / den : 0.0;
                HH(j, j) = cs[j] * a + sn[j] * b2;
                HH(j + 1, j) = 0.0;
                const double gj = g[j];
                g[j] = cs[j] * gj;
                g[j + 1] = -sn[j] * gj;
                rel = std::fabs(g[j + 1]) / rhs_norm;
                j_end = j + 1;
                if(rel < tol) break;
            }
            for(int i = j_end - 1; i >= 0; i--) {
                double sum = g[i];
                for(int k = i + 1; k < j_end; k++) sum -= HH(i, k) * y[k];
                y[i] = sum / HH(i, i);
            }
            if(j_end > 0) {
                if((e = hipMemcpyAsync(B.ydev, y.data(), sizeof(double) * j_end, hipMemcpyHostToDevice, s)) !=
                   hipSuccess)
                    return e;
                hipLaunchKernelGGL((k_update_x<T>), dim3(grid_f), dim3(256), 0, s, f(B.x), (const T *)B.V, B.ydev,
                                   j_end, bins);
                // the copy reads y from host memory: wait before y changes
                if((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            }
            if((e = apply_M(B.x, B.w)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_sub<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.rhs, (const T *)B.w, f(B.r), bins);
            if((e = dot_host(B.r, B.r, beta2)) != hipSuccess) return e;
            beta = std::sqrt(beta2);
            rel = beta / rhs_norm;
            if(rel < tol) break;
        }
        return hipSuccess;
    }

    hipError_t step_full(int &iters) {
        hipError_t e = mobility_and_m0();
        if(e != hipSuccess) return e;
        if(B.a.model == CH_LANDAU)
            hipLaunchKernelGGL((k_psm_g<T, CH_LANDAU>), dim3(grid_f), dim3(256), 0, s, (const T *)B.a.rho, f(B.tr), p(),
                               bins);
        else
            hipLaunchKernelGGL((k_psm_g<T, CH_SIMPLE_WERTHEIM>), dim3(grid_f), dim3(256), 0, s, (const T *)B.a.rho,
                               f(B.tr), p(), bins);
        if((e = r2c(B.tr, B.fh)) != hipSuccess) return e;
        if((e = div_w_grad(B.fh, 0)) != hipSuccess) return e;
        if((e = c2r(B.divh, B.tr)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_psm_axpy_div<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.a.rho, (const T *)B.tr,
                           f(B.b), p(), bins, (T)1);
        hipLaunchKernelGGL((k_scale<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.a.rho, f(B.x), (T)1, bins);
        if((e = gmres(iters)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_scale<T>), dim3(grid_f), dim3(256), 0, s, (const T *)B.x, f(B.a.rho), (T)1, bins);
        return r2c(B.a.rho, B.a.rho_hat);
    }
};

}  // namespace

hipError_t psm_step_simple(const PsmArgs &B, hipStream_t s) {
    if(B.a.precision == CH_FP64) return Ops<double>(B, s).step_simple();
    return Ops<float>(B, s).step_simple();
}

hipError_t psm_step_full(const PsmArgs &B, hipStream_t s, int &iters) {
    if(B.a.precision == CH_FP64) return Ops<double>(B, s).step_full(iters);
    return Ops<float>(B, s).step_full(iters);
}

}  // namespace chd
