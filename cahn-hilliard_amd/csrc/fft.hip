// fft.hip -- hand-written FFT pipeline of the pseudospectral integrator.
//
// The reference step (src/integrators/PseudospectralCPU.cpp:112-136) is
//   f' = f'(rho); f_hat = r2c(f'); rho_hat <- (rho_hat - dt M q^2 (f_hat - S rho_hat))
//        / (1 + dt M (S q^2 + 2 k q^4)); rho = c2r(rho_hat / N^d)
// with FFTW conventions (unnormalised, e^{-i} forward, half spectrum along x).
// Here it runs as one pass per axis over the half spectrum W:
//   x pass   rho -> f' -> r2c (length-N/2 complex FFT + split)      -> W
//   y pass   W -> FFT_y                                             -> W   (3D only)
//   last     W -> FFT -> spectral update of rho_hat -> /N^d -> IFFT  -> W   (fused)
//   y pass   W -> IFFT_y                                            -> W   (3D only)
//   x pass   W -> c2r (merge + length-N/2 IFFT)                     -> rho
// 1D uses x pass, an element-wise update kernel and the inverse x pass.
// Every line FFT is a Stockham radix-4 (plus one radix-2 pass when log2 n is
// odd) in LDS; twiddles come from an FP64 table W_N^m = e^{-2 pi i m / N}
// computed on the host from exact integer indices.
#include "ch_internal.h"

namespace chd {

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2>
__device__ __forceinline__ T2 conj_(T2 a) { return T2{a.x, -a.y}; }

// In-place FFT of `E / n` lines of length n stored line-major in LDS.
// NT threads, EPT = E / NT elements per thread.  INV: e^{+i}, unnormalised.
template <typename T2, bool INV, int EPT, int NT>
__device__ void lds_fft(T2 *buf, int n, int log2n, const T2 *__restrict__ tw, int twstride) {
    const int tid = threadIdx.x;
    int Ns = 1;
    if(log2n & 1) {
        T2 v0[EPT / 2], v1[EPT / 2];
        const int h = n >> 1;
#pragma unroll
        for(int i = 0; i < EPT / 2; i++) {
            const int q = tid + i * NT;
            const int line = q / h, j = q - line * h;
            v0[i] = buf[line * n + j];
            v1[i] = buf[line * n + j + h];
        }
        __syncthreads();
#pragma unroll
        for(int i = 0; i < EPT / 2; i++) {
            const int q = tid + i * NT;
            const int line = q / h, j = q - line * h;
            buf[line * n + 2 * j] = cadd(v0[i], v1[i]);
            buf[line * n + 2 * j + 1] = csub(v0[i], v1[i]);
        }
        __syncthreads();
        Ns = 2;
    }
    const int qn = n >> 2;
    for(; Ns < n; Ns <<= 2) {
        T2 v[EPT / 4][4];
        const int tscale = (n / (4 * Ns)) * twstride;
#pragma unroll
        for(int i = 0; i < EPT / 4; i++) {
            const int q = tid + i * NT;
            const int line = q / qn, j = q - line * qn;
            const int jm = j & (Ns - 1);
#pragma unroll
            for(int r = 0; r < 4; r++) v[i][r] = buf[line * n + j + r * qn];
#pragma unroll
            for(int r = 1; r < 4; r++) {
                T2 w = tw[r * jm * tscale];
                if(INV) w.y = -w.y;
                v[i][r] = cmul(v[i][r], w);
            }
        }
        __syncthreads();
#pragma unroll
        for(int i = 0; i < EPT / 4; i++) {
            const int q = tid + i * NT;
            const int line = q / qn, j = q - line * qn;
            const int jm = j & (Ns - 1);
            const T2 a0 = cadd(v[i][0], v[i][2]), a1 = csub(v[i][0], v[i][2]);
            const T2 a2 = cadd(v[i][1], v[i][3]);
            const T2 d = csub(v[i][1], v[i][3]);
            const T2 a3 = INV ? T2{-d.y, d.x} : T2{d.y, -d.x};  // (+-i) (v1 - v3)
            const int o = line * n + (j - jm) * 4 + jm;
            buf[o] = cadd(a0, a2);
            buf[o + Ns] = cadd(a1, a3);
            buf[o + 2 * Ns] = csub(a0, a2);
            buf[o + 3 * Ns] = csub(a1, a3);
        }
        __syncthreads();
    }
}

// ---- spectral update (PseudospectralCPU.cpp:119-133), mode -> q^2 exactly as
// the reference builds its tables (PseudospectralCPU.cpp:34-83)
__device__ __forceinline__ void mode_q2(const Params &p, int ix, int iy, int iz, double &k2, double &deal) {
    const double PI = 3.14159265358979323846;
    const int N = p.N;
    const double Ndx = N * p.dx;
    const double k_cut = N * PI / (N * p.dx) * 2.0 / 3.0;
    if(p.dims == 1) {
        double k = 2.0 * PI * ix / Ndx;
        deal = (k <= k_cut) ? 1.0 : 0.0;
        k2 = k * k;
    }
    else if(p.dims == 2) {
        int kx = (iy < N / 2) ? iy : N - iy;
        double k = 2.0 * PI * sqrt((double)(kx * kx + ix * ix)) / Ndx;
        deal = (k < k_cut) ? 1.0 : 0.0;
        k2 = k * k;
    }
    else {
        int kx_i = (iz < N / 2) ? iz : iz - N;
        int ky_i = (iy < N / 2) ? iy : iy - N;
        double kx = 2.0 * PI * kx_i / Ndx, ky = 2.0 * PI * ky_i / Ndx, kz = 2.0 * PI * ix / Ndx;
        k2 = kx * kx + ky * ky + kz * kz;
        double k = sqrt(k2);
        deal = (k <= k_cut) ? 1.0 : 0.0;
    }
}

template <typename T, typename T2>
__device__ __forceinline__ T2 spectral_update(const Params &p, int ix, int iy, int iz, T2 fh, T2 *rho_hat_elem,
                                              T inv_bins) {
    double k2d, deal;
    mode_q2(p, ix, iy, iz, k2d, deal);
    if(p.use_dealias) { fh.x *= (T)deal; fh.y *= (T)deal; }
    const T k2 = (T)k2d, k4 = k2 * k2;
    const T dt = (T)p.dt, M = (T)p.M, S = (T)p.S;
    const T denom = (T)1.0 + dt * M * (S * k2 + (T)2.0 * (T)p.k * k4);
    const T2 rh = *rho_hat_elem;
    const T c = dt * M * k2;
    T2 nw;
    nw.x = (rh.x - c * (fh.x - S * rh.x)) / denom;
    nw.y = (rh.y - c * (fh.y - S * rh.y)) / denom;
    *rho_hat_elem = nw;
    return T2{nw.x * inv_bins, nw.y * inv_bins};
}

// ---- x pass, forward: rows of N reals -> N/2+1 complex (r2c)
template <typename T, int MODEL, bool FPRIME, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_xfwd(const T *__restrict__ in, typename Complex<T>::type *__restrict__ out,
                                            const typename Complex<T>::type *__restrict__ tw, Params p, int rows) {
    using T2 = typename Complex<T>::type;
    constexpr int E = EPT * NT;
    __shared__ T2 buf[E];
    const int N = p.N, n = N >> 1, nl = E / n;
    const int row0 = blockIdx.x * nl;
    T *bufT = reinterpret_cast<T *>(buf);
    for(int e = threadIdx.x; e < nl * N; e += NT) {
        const int line = e / N, x = e - line * N;
        const int row = row0 + line;
        T v = 0;
        if(row < rows) {
            v = in[(size_t)row * N + x];
            if constexpr(FPRIME) v = fprime<T, MODEL>(p, v);
        }
        bufT[line * N + x] = v;
    }
    __syncthreads();
    lds_fft<T2, false, EPT, NT>(buf, n, p.log2N - 1, tw, 2);
    const int Nh = n + 1;
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e - line * Nh;
        const int row = row0 + line;
        if(row >= rows) continue;
        const T2 zk = buf[line * n + (k == n ? 0 : k)];
        const T2 znk = buf[line * n + (k == 0 ? 0 : n - k)];
        // E = (Z_k + conj Z_{n-k}) / 2, O = (Z_k - conj Z_{n-k}) / (2i), X_k = E + W_N^k O
        const T2 Ev = T2{(zk.x + znk.x) * (T)0.5, (zk.y - znk.y) * (T)0.5};
        const T2 dd = T2{zk.x - znk.x, zk.y + znk.y};
        const T2 Ov = T2{dd.y * (T)0.5, -dd.x * (T)0.5};
        const T2 w = tw[k];
        out[(size_t)row * Nh + k] = cadd(Ev, cmul(w, Ov));
    }
}

// ---- x pass, inverse: N/2+1 complex -> N reals (c2r, unnormalised).  The
// imaginary parts of X_0 and X_{N/2} are dropped (Hermitian completion).
template <typename T, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_xinv(const typename Complex<T>::type *__restrict__ in, T *__restrict__ out,
                                            const typename Complex<T>::type *__restrict__ tw, Params p, int rows) {
    using T2 = typename Complex<T>::type;
    constexpr int E = EPT * NT;
    __shared__ T2 buf[E + E / 2];  // nl (n + 1) <= E + E/2 since n >= 2
    const int N = p.N, n = N >> 1, nl = E / n, Nh = n + 1;
    const int row0 = blockIdx.x * nl;
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e - line * Nh;
        const int row = row0 + line;
        buf[e] = (row < rows) ? in[(size_t)row * Nh + k] : T2{0, 0};
    }
    __syncthreads();
    T2 z[EPT];
#pragma unroll
    for(int i = 0; i < EPT; i++) {
        const int q = threadIdx.x + i * NT;
        const int line = q / n, k = q - line * n;
        T2 xk = buf[line * Nh + k];
        T2 xnk = buf[line * Nh + n - k];
        if(k == 0) { xk.y = 0; xnk.y = 0; }
        // E = X_k + conj X_{n-k}; O = (X_k - conj X_{n-k}) conj(W_N^k); Z = E + i O
        const T2 Ev = T2{xk.x + xnk.x, xk.y - xnk.y};
        const T2 Ov = cmul(T2{xk.x - xnk.x, xk.y + xnk.y}, conj_(tw[k]));
        z[i] = T2{Ev.x - Ov.y, Ev.y + Ov.x};
    }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < EPT; i++) buf[threadIdx.x + i * NT] = z[i];
    __syncthreads();
    lds_fft<T2, true, EPT, NT>(buf, n, p.log2N - 1, tw, 2);
    const T *bufT = reinterpret_cast<const T *>(buf);
    for(int e = threadIdx.x; e < nl * N; e += NT) {
        const int line = e / N, x = e - line * N;
        const int row = row0 + line;
        if(row < rows) out[(size_t)row * N + x] = bufT[line * N + x];
    }
}

// ---- strided axis pass over the half spectrum.  Line l starts at
// (l / Nh) * P + (l % Nh) and steps by `stride`.  MODE 0: forward, 1: inverse,
// 2: forward + spectral update of rho_hat + /N^d + inverse (the last axis).
template <typename T, int MODE, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_col(const typename Complex<T>::type *__restrict__ src,
                                           typename Complex<T>::type *__restrict__ dst,
                                           typename Complex<T>::type *__restrict__ rho_hat,
                                           const typename Complex<T>::type *__restrict__ tw, Params p, int lines,
                                           long P, long stride, int axis) {
    using T2 = typename Complex<T>::type;
    constexpr int E = EPT * NT;
    __shared__ T2 buf[E];
    const int N = p.N, n = N, nl = E / n, Nh = N / 2 + 1;
    const int l0 = blockIdx.x * nl;
    for(int e = threadIdx.x; e < nl * n; e += NT) {
        const int j = e / nl, lb = e - j * nl;
        const int l = l0 + lb;
        T2 v{0, 0};
        if(l < lines) v = src[(size_t)(l / Nh) * P + (l % Nh) + (size_t)j * stride];
        buf[lb * n + j] = v;
    }
    __syncthreads();
    if constexpr(MODE == 1) {
        lds_fft<T2, true, EPT, NT>(buf, n, p.log2N, tw, 1);
    }
    else {
        lds_fft<T2, false, EPT, NT>(buf, n, p.log2N, tw, 1);
    }
    if constexpr(MODE == 2) {
        const T inv_bins = (T)p.inv_bins;
#pragma unroll
        for(int i = 0; i < EPT; i++) {
            const int q = threadIdx.x + i * NT;
            const int lb = q / n, j = q - lb * n;
            const int l = l0 + lb;
            if(l < lines) {
                const size_t gi = (size_t)(l / Nh) * P + (l % Nh) + (size_t)j * stride;
                const int ix = l % Nh;
                int iy, iz;
                if(axis == 1) { iy = j; iz = l / Nh; }   // y pass (2D last axis)
                else { iy = l / Nh; iz = j; }             // z pass (3D last axis)
                buf[q] = spectral_update<T, T2>(p, ix, iy, iz, buf[q], rho_hat + gi, inv_bins);
            }
        }
        __syncthreads();
        lds_fft<T2, true, EPT, NT>(buf, n, p.log2N, tw, 1);
    }
    for(int e = threadIdx.x; e < nl * n; e += NT) {
        const int j = e / nl, lb = e - j * nl;
        const int l = l0 + lb;
        if(l < lines) dst[(size_t)(l / Nh) * P + (l % Nh) + (size_t)j * stride] = buf[lb * n + j];
    }
}

// ---- element-wise spectral update (1D)
template <typename T>
__global__ void k_update1d(typename Complex<T>::type *__restrict__ W, typename Complex<T>::type *__restrict__ rho_hat,
                           Params p, int nh) {
    using T2 = typename Complex<T>::type;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nh) return;
    W[i] = spectral_update<T, T2>(p, i, 0, 0, W[i], rho_hat + i, (T)p.inv_bins);
}

// ------------------------------------------------------------------ host side
template <typename T, int EPT>
static hipError_t ps_run_t(const SpectralArgs &a, hipStream_t s, bool init) {
    using T2 = typename Complex<T>::type;
    constexpr int NT = 256, E = EPT * NT;
    const Params &p = a.p;
    const int N = p.N, dims = p.dims, Nh = N / 2 + 1;
    const int rows = dims == 1 ? 1 : (dims == 2 ? N : N * N);
    const T2 *tw2 = (const T2 *)a.twiddles;
    T *rho = (T *)a.rho;
    T2 *W = (T2 *)a.work, *RH = (T2 *)a.rho_hat;
    const int nlx = E / (N / 2);
    const int bx = (rows + nlx - 1) / nlx;
    const int nlc = E / N;
    // forward x pass
    if(init) {
        hipLaunchKernelGGL((k_xfwd<T, CH_LANDAU, false, EPT, NT>), dim3(bx), dim3(NT), 0, s, rho,
                           dims == 1 ? RH : W, tw2, p, rows);
    }
    else if(a.model == CH_LANDAU) {
        hipLaunchKernelGGL((k_xfwd<T, CH_LANDAU, true, EPT, NT>), dim3(bx), dim3(NT), 0, s, rho, W, tw2, p, rows);
    }
    else {
        hipLaunchKernelGGL((k_xfwd<T, CH_SIMPLE_WERTHEIM, true, EPT, NT>), dim3(bx), dim3(NT), 0, s, rho, W, tw2, p,
                           rows);
    }
    if(dims == 1) {
        if(init) return hipGetLastError();
        hipLaunchKernelGGL((k_update1d<T>), dim3((Nh + 255) / 256), dim3(256), 0, s, W, RH, p, Nh);
    }
    else if(dims == 2) {
        const int lines = Nh;
        const int bc = (lines + nlc - 1) / nlc;
        if(init) {
            hipLaunchKernelGGL((k_col<T, 0, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, RH, RH, tw2, p, lines, (long)N * Nh,
                               (long)Nh, 1);
            return hipGetLastError();
        }
        hipLaunchKernelGGL((k_col<T, 2, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, W, RH, tw2, p, lines, (long)N * Nh,
                           (long)Nh, 1);
    }
    else {
        const int lines = N * Nh;
        const int bc = (lines + nlc - 1) / nlc;
        // y forward (line l: plane l / Nh, column l % Nh), then the fused z pass
        hipLaunchKernelGGL((k_col<T, 0, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, W, RH, tw2, p, lines, (long)N * Nh,
                           (long)Nh, 1);
        if(init) {
            hipLaunchKernelGGL((k_col<T, 0, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, RH, RH, tw2, p, lines, (long)Nh,
                               (long)N * Nh, 2);
            return hipGetLastError();
        }
        hipLaunchKernelGGL((k_col<T, 2, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, W, RH, tw2, p, lines, (long)Nh,
                           (long)N * Nh, 2);
        hipLaunchKernelGGL((k_col<T, 1, EPT, NT>), dim3(bc), dim3(NT), 0, s, W, W, RH, tw2, p, lines, (long)N * Nh,
                           (long)Nh, 1);
    }
    hipLaunchKernelGGL((k_xinv<T, EPT, NT>), dim3(bx), dim3(NT), 0, s, W, rho, tw2, p, rows);
    return hipGetLastError();
}

// E = EPT * 256 complex elements per block must hold one line: n <= E.
hipError_t ps_run(const SpectralArgs &a, hipStream_t s, bool init) {
    const int N = a.p.N;
    if(a.precision == CH_FP64) {
        if(N <= 2048) return ps_run_t<double, 8>(a, s, init);
        return ps_run_t<double, 16>(a, s, init);
    }
    if(N <= 2048) return ps_run_t<float, 8>(a, s, init);
    return ps_run_t<float, 16>(a, s, init);
}

}  // namespace chd
