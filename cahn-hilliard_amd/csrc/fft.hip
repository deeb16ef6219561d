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
// 1D uses the x pass, an element-wise update kernel and the inverse x pass.
//
// Every line FFT is a Stockham radix-4 (plus one radix-2 pass when log2 n is
// odd) in LDS with the length a template parameter: each thread owns one
// radix-4 butterfly per pass (4 elements), all of its twiddles are loaded
// into registers before the first pass, and lines of n <= 256 are packed
// several per block.  Twiddles come from an FP64 table W_N^m = e^{-2 pi i m/N}
// built on the host from exact integer indices (rounded once for FP32).
#include "ch_internal.h"
#include "psm_ops.h"
#include "wave_fft.h"

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

// streaming store of one complex entry (written once, read by the next pass)
template <typename T2>
__device__ __forceinline__ void nt_store(T2 *p, T2 v) {
    using T = decltype(T2{}.x);
    typedef T V2 __attribute__((ext_vector_type(2)));
    V2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<V2 *>(p));
}

// LDS index with one pad element every 32: the Stockham writes of the first
// passes (stride 4 elements) then spread over every bank group.
__device__ __forceinline__ constexpr int pidx(int i) { return i + (i >> 5); }
constexpr int padded(int n) { return n + (n >> 5) + 1; }

// Block geometry of a length-2^LOG2N line FFT: E elements, NT = E / 4 threads.
template <int LOG2N, int NLINES>
struct FftGeom {
    static constexpr int n = 1 << LOG2N;
    static constexpr int NL = NLINES;
    static constexpr int E = n * NL;
    static constexpr int NT = E / 4;
    static constexpr bool R2 = (LOG2N & 1) != 0;
    static constexpr int NP4 = LOG2N / 2;  // radix-4 passes
};

// lines of n <= 256 are packed four-per-1024 elements (NT = 256)
template <int LOG2N>
struct PackedLines { static constexpr int value = LOG2N >= 8 ? 1 : (1 << (10 - LOG2N)); };

template <typename T2, int LOG2N, int NLINES>
struct LineFFT {
    using G = FftGeom<LOG2N, NLINES>;
    T2 tw[G::NP4 > 0 ? G::NP4 : 1][3];

    // table: W_Ntab^m; twstride = Ntab / n
    __device__ __forceinline__ void prefetch(const T2 *__restrict__ table, int twstride) {
        constexpr int n = G::n, qn = n >> 2;
        const int j = (int)threadIdx.x % (qn > 0 ? qn : 1);
#pragma unroll
        for(int ps = 0; ps < G::NP4; ps++) {
            const int Ns = (G::R2 ? 2 : 1) << (2 * ps);
            const int tscale = (n / (4 * Ns)) * twstride;
            const int jm = j & (Ns - 1);
#pragma unroll
            for(int r = 1; r < 4; r++) tw[ps][r - 1] = table[r * jm * tscale];
        }
    }

    // in-place transform of NL lines stored line-major in buf; INV: e^{+i}
    template <bool INV>
    __device__ __forceinline__ void run(T2 *buf) const {
        constexpr int n = G::n, NT = G::NT;
        const int tid = threadIdx.x;
        if constexpr(G::R2) {
            constexpr int h = n >> 1;
            T2 v0[2], v1[2];
#pragma unroll
            for(int i = 0; i < 2; i++) {
                const int q = tid + i * NT;
                const int line = q / h, j = q % h;
                v0[i] = buf[pidx(line * n + j)];
                v1[i] = buf[pidx(line * n + j + h)];
            }
            __syncthreads();
#pragma unroll
            for(int i = 0; i < 2; i++) {
                const int q = tid + i * NT;
                const int line = q / h, j = q % h;
                buf[pidx(line * n + 2 * j)] = cadd(v0[i], v1[i]);
                buf[pidx(line * n + 2 * j + 1)] = csub(v0[i], v1[i]);
            }
            __syncthreads();
        }
        constexpr int qn = n >> 2;
        const int line = tid / (qn > 0 ? qn : 1), j = tid % (qn > 0 ? qn : 1);
#pragma unroll
        for(int ps = 0; ps < G::NP4; ps++) {
            const int Ns = (G::R2 ? 2 : 1) << (2 * ps);
            const int jm = j & (Ns - 1);
            T2 v[4];
#pragma unroll
            for(int r = 0; r < 4; r++) v[r] = buf[pidx(line * n + j + r * qn)];
#pragma unroll
            for(int r = 1; r < 4; r++) v[r] = cmul(v[r], INV ? conj_(tw[ps][r - 1]) : tw[ps][r - 1]);
            __syncthreads();
            const T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
            const T2 a2 = cadd(v[1], v[3]);
            const T2 d = csub(v[1], v[3]);
            const T2 a3 = INV ? T2{-d.y, d.x} : T2{d.y, -d.x};  // (+-i)(v1 - v3)
            const int o = line * n + (j - jm) * 4 + jm;
            buf[pidx(o)] = cadd(a0, a2);
            buf[pidx(o + Ns)] = cadd(a1, a3);
            buf[pidx(o + 2 * Ns)] = csub(a0, a2);
            buf[pidx(o + 3 * Ns)] = csub(a1, a3);
            __syncthreads();
        }
    }
};

// ---- spectral update (PseudospectralCPU.cpp:119-133); mode_q2 (ch_device.h)
// gives q^2 exactly as the reference builds its tables
template <typename T, typename T2>
__device__ __forceinline__ T2 spectral_update(const Params &p, int ix, int iy, int iz, T2 fh, T2 &rh, T inv_bins) {
    double k2d, deal;
    mode_q2(p, ix, iy, iz, k2d, deal);
    if(p.use_dealias) { fh.x *= (T)deal; fh.y *= (T)deal; }
    const T k2 = (T)k2d, k4 = k2 * k2;
    const T dt = (T)p.dt, M = (T)p.M, S = (T)p.S;
    const T inv_denom = (T)1.0 / ((T)1.0 + dt * M * (S * k2 + (T)2.0 * (T)p.k * k4));
    const T c = dt * M * k2;
    T2 nw;
    nw.x = (rh.x - c * (fh.x - S * rh.x)) * inv_denom;
    nw.y = (rh.y - c * (fh.y - S * rh.y)) * inv_denom;
    rh = nw;
    return T2{nw.x * inv_bins, nw.y * inv_bins};
}

// ---- x pass, forward: rows of N = 2n reals -> n + 1 complex (r2c).
// LOG2N is log2 of the complex length n = N / 2.
template <typename T, int MODEL, bool FPRIME, int LOG2N>
__global__ __launch_bounds__((FftGeom<LOG2N, PackedLines<LOG2N>::value>::NT)) void k_xfwd(
    const T *__restrict__ in, typename Complex<T>::type *__restrict__ out, const typename Complex<T>::type *__restrict__ tw,
    Params p, int rows, long rs, long ks) {
    using T2 = typename Complex<T>::type;
    using G = FftGeom<LOG2N, PackedLines<LOG2N>::value>;
    constexpr int n = G::n, N = 2 * n, nl = G::NL, NT = G::NT;
    __shared__ T2 buf[padded(G::E)];
    LineFFT<T2, LOG2N, nl> fft;
    fft.prefetch(tw, 2);
    const int row0 = xcd_tile(blockIdx.x, gridDim.x) * nl;
    T *bufT = reinterpret_cast<T *>(buf);
#pragma unroll
    for(int i = 0; i < 8; i++) {
        const int e = threadIdx.x + i * NT;  // nl * N = 8 NT reals
        const int line = e / N, x = e % N;
        const int row = row0 + line;
        T v = 0;
        if(row < rows) {
            v = in[(size_t)row * N + x];
            if constexpr(FPRIME) v = fprime<T, MODEL>(p, v);
        }
        bufT[2 * pidx(e >> 1) + (e & 1)] = v;
    }
    __syncthreads();
    fft.template run<false>(buf);
    constexpr int Nh = n + 1;
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e % Nh;
        const int row = row0 + line;
        if(row >= rows) continue;
        const T2 zk = buf[pidx(line * n + (k == n ? 0 : k))];
        const T2 znk = buf[pidx(line * n + (k == 0 ? 0 : n - k))];
        // E = (Z_k + conj Z_{n-k}) / 2, O = (Z_k - conj Z_{n-k}) / (2i), X_k = E + W_N^k O
        const T2 Ev = T2{(zk.x + znk.x) * (T)0.5, (zk.y - znk.y) * (T)0.5};
        const T2 dd = T2{zk.x - znk.x, zk.y + znk.y};
        const T2 Ov = T2{dd.y * (T)0.5, -dd.x * (T)0.5};
        out[(size_t)row * rs + (size_t)k * ks] = cadd(Ev, cmul(tw[k], Ov));
    }
}

// ---- x pass, inverse: n + 1 complex -> N = 2n reals (c2r, unnormalised).  The
// imaginary parts of X_0 and X_{N/2} are dropped (Hermitian completion).
template <typename T, int LOG2N>
__global__ __launch_bounds__((FftGeom<LOG2N, PackedLines<LOG2N>::value>::NT)) void k_xinv(
    const typename Complex<T>::type *__restrict__ in, T *__restrict__ out, const typename Complex<T>::type *__restrict__ tw,
    Params p, int rows, long rs, long ks) {
    using T2 = typename Complex<T>::type;
    using G = FftGeom<LOG2N, PackedLines<LOG2N>::value>;
    constexpr int n = G::n, N = 2 * n, nl = G::NL, NT = G::NT, Nh = n + 1;
    __shared__ T2 buf[padded(G::E + nl)];
    LineFFT<T2, LOG2N, nl> fft;
    fft.prefetch(tw, 2);
    const int row0 = xcd_tile(blockIdx.x, gridDim.x) * nl;
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e % Nh;
        const int row = row0 + line;
        buf[pidx(e)] = (row < rows) ? in[(size_t)row * rs + (size_t)k * ks] : T2{0, 0};
    }
    __syncthreads();
    T2 z[4];
#pragma unroll
    for(int i = 0; i < 4; i++) {
        const int q = threadIdx.x + i * NT;
        const int line = q / n, k = q % n;
        T2 xk = buf[pidx(line * Nh + k)];
        T2 xnk = buf[pidx(line * Nh + n - k)];
        if(k == 0) { xk.y = 0; xnk.y = 0; }
        // E = X_k + conj X_{n-k}; O = (X_k - conj X_{n-k}) conj(W_N^k); Z = E + i O
        const T2 Ev = T2{xk.x + xnk.x, xk.y - xnk.y};
        const T2 Ov = cmul(T2{xk.x - xnk.x, xk.y + xnk.y}, conj_(tw[k]));
        z[i] = T2{Ev.x - Ov.y, Ev.y + Ov.x};
    }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < 4; i++) buf[pidx(threadIdx.x + i * NT)] = z[i];
    __syncthreads();
    fft.template run<true>(buf);
    const T *bufT = reinterpret_cast<const T *>(buf);
#pragma unroll
    for(int i = 0; i < 8; i++) {
        const int e = threadIdx.x + i * NT;
        const int line = e / N, x = e % N;
        const int row = row0 + line;
        if(row < rows) out[(size_t)row * N + x] = bufT[2 * pidx(e >> 1) + (e & 1)];
    }
}

// ---- x pass between two time steps: c2r of step n's half spectrum (as
// k_xinv), then f' and r2c of step n + 1 (as k_xfwd), in place on W.  The real
// row stays in LDS as T values, so the result has the same bits as k_xinv
// followed by k_xfwd, with one read and one write of W instead of two each.
// rho_out (may be null): the c2r output rows are also stored there (the
// recompute of a reverse sweep keeps every state)
template <typename T, int MODEL, int LOG2N>
__global__ __launch_bounds__((FftGeom<LOG2N, PackedLines<LOG2N>::value>::NT)) void k_xrow(
    typename Complex<T>::type *__restrict__ W, const typename Complex<T>::type *__restrict__ tw, Params p, int rows,
    long rs, long ks, T *__restrict__ rho_out) {
    using T2 = typename Complex<T>::type;
    using G = FftGeom<LOG2N, PackedLines<LOG2N>::value>;
    constexpr int n = G::n, N = 2 * n, nl = G::NL, NT = G::NT, Nh = n + 1;
    __shared__ T2 buf[padded(G::E + nl)];
    LineFFT<T2, LOG2N, nl> fft;
    fft.prefetch(tw, 2);
    const int row0 = xcd_tile(blockIdx.x, gridDim.x) * nl;
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e % Nh;
        const int row = row0 + line;
        buf[pidx(e)] = (row < rows) ? W[(size_t)row * rs + (size_t)k * ks] : T2{0, 0};
    }
    __syncthreads();
    T2 z[4];
#pragma unroll
    for(int i = 0; i < 4; i++) {
        const int q = threadIdx.x + i * NT;
        const int line = q / n, k = q % n;
        T2 xk = buf[pidx(line * Nh + k)];
        T2 xnk = buf[pidx(line * Nh + n - k)];
        if(k == 0) { xk.y = 0; xnk.y = 0; }
        const T2 Ev = T2{xk.x + xnk.x, xk.y - xnk.y};
        const T2 Ov = cmul(T2{xk.x - xnk.x, xk.y + xnk.y}, conj_(tw[k]));
        z[i] = T2{Ev.x - Ov.y, Ev.y + Ov.x};
    }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < 4; i++) buf[pidx(threadIdx.x + i * NT)] = z[i];
    __syncthreads();
    fft.template run<true>(buf);
    T *bufT = reinterpret_cast<T *>(buf);
    // rho (the c2r output, as T) -> f'(rho); each thread rewrites its own reals
#pragma unroll
    for(int i = 0; i < 8; i++) {
        const int e = threadIdx.x + i * NT;
        const int row = row0 + e / N;
        const int a = 2 * pidx(e >> 1) + (e & 1);
        if(rho_out && row < rows) rho_out[(size_t)row * N + e % N] = bufT[a];
        bufT[a] = (row < rows) ? fprime<T, MODEL>(p, bufT[a]) : (T)0;
    }
    __syncthreads();
    fft.template run<false>(buf);
    for(int e = threadIdx.x; e < nl * Nh; e += NT) {
        const int line = e / Nh, k = e % Nh;
        const int row = row0 + line;
        if(row >= rows) continue;
        const T2 zk = buf[pidx(line * n + (k == n ? 0 : k))];
        const T2 znk = buf[pidx(line * n + (k == 0 ? 0 : n - k))];
        const T2 Ev = T2{(zk.x + znk.x) * (T)0.5, (zk.y - znk.y) * (T)0.5};
        const T2 dd = T2{zk.x - znk.x, zk.y + znk.y};
        const T2 Ov = T2{dd.y * (T)0.5, -dd.x * (T)0.5};
        W[(size_t)row * rs + (size_t)k * ks] = cadd(Ev, cmul(tw[k], Ov));
    }
}

// ---- axis pass over the half spectrum, NLC lines per block.  Line l starts
// at (l / Nh) * P + (l % Nh) * ls and steps by `stride`.  MODE 0: forward,
// 1: inverse, 2: forward + spectral update of rho_hat + /N^d + inverse.
// Strided lines (stride > 1): element e of a block is (row j = e / NLC, line
// l0 + e % NLC), so a wave touches NLC adjacent lines of 32 rows; the lines of
// one 128-byte row segment belong to consecutive blocks, which the XCD-aware
// block order keeps on one XCD (one L2).  Contiguous lines (CONTIG, stride 1):
// element e is (j = e % n, line l0 + e / n), a wave reads 64 consecutive
// entries.  rho_hat is fetched at kernel start, ahead of the transform.
template <typename T, int MODE, int LOG2N, int NLC, bool CONTIG = false>
__global__ __launch_bounds__((FftGeom<LOG2N, NLC>::NT)) void k_col(
    const typename Complex<T>::type *__restrict__ src, typename Complex<T>::type *__restrict__ dst,
    typename Complex<T>::type *__restrict__ rho_hat, const typename Complex<T>::type *__restrict__ tw, Params p,
    int lines, long P, long ls, long stride, int axis) {
    using T2 = typename Complex<T>::type;
    using G = FftGeom<LOG2N, NLC>;
    constexpr int n = G::n, NT = G::NT, nl = NLC;
    __shared__ T2 buf[padded(G::E)];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int l0 = tile * nl;
    if(l0 >= lines) return;  // grid padded to a multiple of 8 blocks
    LineFFT<T2, LOG2N, nl> fft;
    fft.prefetch(tw, 1);
    const int Nh = p.N / 2 + 1;
    size_t gi[4];
    T2 rh[4];
#pragma unroll
    for(int i = 0; i < 4; i++) {
        const int e = threadIdx.x + i * NT;
        const int j = CONTIG ? e % n : e / nl, lb = CONTIG ? e / n : e % nl;
        const int l = l0 + lb;
        gi[i] = (size_t)(l / Nh) * P + (size_t)(l % Nh) * ls + (size_t)j * stride;
        T2 v{0, 0};
        if(l < lines) {
            v = src[gi[i]];
            if constexpr(MODE == 2) rh[i] = rho_hat[gi[i]];
        }
        buf[pidx(lb * n + j)] = v;
    }
    __syncthreads();
    fft.template run<MODE == 1>(buf);
    if constexpr(MODE == 2) {
        const T inv_bins = (T)p.inv_bins;
#pragma unroll
        for(int i = 0; i < 4; i++) {
            const int e = threadIdx.x + i * NT;
            const int j = CONTIG ? e % n : e / nl, lb = CONTIG ? e / n : e % nl;
            const int l = l0 + lb;
            if(l < lines) {
                const int ix = l % Nh;
                int iy, iz;
                if(axis == 1) { iy = j; iz = l / Nh; }  // y pass (2D last axis)
                else { iy = l / Nh; iz = j; }            // z pass (3D last axis)
                const int q = pidx(lb * n + j);
                buf[q] = spectral_update<T, T2>(p, ix, iy, iz, buf[q], rh[i], inv_bins);
            }
        }
        __syncthreads();
        fft.template run<true>(buf);
        // after the transform's barriers (each barrier waits for earlier global stores)
#pragma unroll
        for(int i = 0; i < 4; i++) {
            const int e = threadIdx.x + i * NT;
            const int lb = CONTIG ? e / n : e % nl;
            if(l0 + lb < lines) rho_hat[gi[i]] = rh[i];
        }
    }
#pragma unroll
    for(int i = 0; i < 4; i++) {
        const int e = threadIdx.x + i * NT;
        const int j = CONTIG ? e % n : e / nl, lb = CONTIG ? e / n : e % nl;
        if(l0 + lb < lines) dst[gi[i]] = buf[pidx(lb * n + j)];
    }
}

// ---- element-wise spectral update (1D)
template <typename T>
__global__ void k_update1d(typename Complex<T>::type *__restrict__ W, typename Complex<T>::type *__restrict__ rho_hat,
                           Params p, int nh) {
    using T2 = typename Complex<T>::type;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nh) return;
    T2 rh = rho_hat[i];
    W[i] = spectral_update<T, T2>(p, i, 0, 0, W[i], rh, (T)p.inv_bins);
    rho_hat[i] = rh;
}

// ---- one wave per line (wave_fft.h), for lines of 512, 1024 or 2048 points.
// Their barriers order LDS only (lds_barrier): a line's global loads stay in
// flight across them and its stores need not complete before the next one.
// Every global load of these kernels is consumed before the first barrier or
// rewritten later by its own lane only, so no other lane can overwrite it early.
// Fewer LDS passes than LineFFT (two exchanges per transform, wave-local) and
// the line's entries stay in registers between the forward transform, the
// spectral update and the inverse transform.

// The first barrier orders the block's earlier LDS accesses (the staged
// twiddle table, another wave's reads of this buffer) before the transform;
// the buffer is the wave's own from there on, so the passes within the
// transform are ordered by wave_lds_sync alone.  BLOCK_FIRST = false: the
// caller has ordered the block's accesses already and the buffer is the
// wave's own, so the first sync is the wave's too.
template <typename T2, int LOG2N, bool INV, bool BLOCK_FIRST = true, int E>
__device__ __forceinline__ void wave_transform(int lane, T2 (&v)[E], T2 *lds, const T2 *tw, int tstride) {
    static_assert(E == WaveFFT<T2, LOG2N>::E, "entries per lane");
    using F = WaveFFT<T2, LOG2N>;
    if constexpr(BLOCK_FIRST) lds_barrier();
    else wave_lds_sync();
    F::template pass1_write<INV>(lane, v, lds);
    wave_lds_sync();
    F::pass_read(lane, v, lds);
    wave_lds_sync();
    F::template pass2_write<INV>(lane, v, lds, tw, tstride);
    wave_lds_sync();
    F::pass_read(lane, v, lds);
    F::template pass3_final<INV>(lane, v, tw, tstride);
}

// The system's twiddle table (N entries, W_N^m) staged in LDS when N <= 1024:
// one wave runs per SIMD at most (a line per wave), so nothing hides a global
// load, and every transform pass after the first barrier would otherwise wait
// on its twiddles.  load() fetches a thread's share of the table into
// registers and store() writes it to LDS: a kernel issues load(), then the
// line's own loads, then store(), so that the two sets of loads are in flight
// together (a store straight after each load would wait out the table's
// latency before the line's loads issue).
template <typename T2, int TN>
struct TwiddleStage {
    static constexpr bool on = TN <= 1024;
    T2 t[on ? TN : 1];
    // entries per thread of NT threads (the register array of load / store)
    template <int NT>
    static constexpr int per() { return on ? TN / NT : 1; }
    template <int NT>
    __device__ __forceinline__ void load(const T2 *__restrict__ tw, int tid, T2 (&r)[per<NT>()]) const {
        if constexpr(on) {
#pragma unroll
            for(int i = 0; i < TN / NT; i++) r[i] = tw[tid + NT * i];
        }
        else {
            (void)tw;
            (void)tid;
            (void)r;
        }
    }
    // the table to read from: the LDS copy, visible after the next barrier
    template <int NT>
    __device__ __forceinline__ const T2 *store(const T2 *tw, int tid, const T2 (&r)[per<NT>()]) {
        if constexpr(on) {
#pragma unroll
            for(int i = 0; i < TN / NT; i++) t[tid + NT * i] = r[i];
            return t;
        }
        else {
            (void)tid;
            (void)r;
            return tw;
        }
    }
};

// axis pass (as k_col), one line per 64-thread block
template <typename T, int MODE, int LOG2N>
__global__ __launch_bounds__(64) void k_colw(const typename Complex<T>::type *__restrict__ src,
                                             typename Complex<T>::type *__restrict__ dst,
                                             typename Complex<T>::type *__restrict__ rho_hat,
                                             const typename Complex<T>::type *tw, Params p, int lines,
                                             long P, long ls, long stride, int axis) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E;
    __shared__ T2 lds[F::LDS];
    using Stage = TwiddleStage<T2, (1 << LOG2N)>;  // table length N = line length
    __shared__ Stage stage;
    const int l = xcd_tile(blockIdx.x, gridDim.x);
    if(l >= lines) return;
    const int lane = threadIdx.x;
    T2 twr[Stage::template per<64>()];
    stage.template load<64>(tw, lane, twr);
    const int Nh = p.N / 2 + 1;
    const size_t base = (size_t)(l / Nh) * P + (size_t)(l % Nh) * ls;
    const int tstride = p.N >> LOG2N;
    T2 v[E], rh[MODE == 2 ? E : 1];
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = src[base + (size_t)(lane + 64 * m) * stride];
    if constexpr(MODE == 2) {
#pragma unroll
        for(int m = 0; m < E; m++) rh[m] = rho_hat[base + (size_t)(lane + 64 * m) * stride];
    }
    tw = stage.template store<64>(tw, lane, twr);
    wave_transform<T2, LOG2N, MODE == 1>(lane, v, lds, tw, tstride);
    if constexpr(MODE == 2) {
        const T inv_bins = (T)p.inv_bins;
        const int ix = l % Nh;
#pragma unroll
        for(int m = 0; m < E; m++) {
            const int j = lane + 64 * m;
            int iy, iz;
            if(axis == 1) { iy = j; iz = l / Nh; }  // y pass (2D last axis)
            else { iy = l / Nh; iz = j; }            // z pass (3D last axis)
            v[m] = spectral_update<T, T2>(p, ix, iy, iz, v[m], rh[m], inv_bins);
        }
        wave_transform<T2, LOG2N, true>(lane, v, lds, tw, tstride);
        // after the transform's barriers (each barrier waits for earlier global stores)
#pragma unroll
        for(int m = 0; m < E; m++) rho_hat[base + (size_t)(lane + 64 * m) * stride] = rh[m];
    }
#pragma unroll
    for(int m = 0; m < E; m++) dst[base + (size_t)(lane + 64 * m) * stride] = v[m];
}

// ---- two waves per line of n = 2 n2 points (n2 = 64 E): the axis pass for
// n = 1024 and 2048, where one wave's chain of transform passes, not memory,
// bounds the single-wave kernel (a 2D N = 1024 grid has only 513 lines).
// Forward, decimation in time: wave h transforms the entries 2 i + h
// (i = lane + 64 m) with the n2-point wave FFT; X_k = E_k + W_n^k O_k and
// X_{k + n2} = E_k - W_n^k O_k, so after one exchange wave h holds X_{k + h n2}.
// Inverse, decimation in frequency: e_k = X_k + X_{k + n2},
// o_k = (X_k - X_{k + n2}) W_n^{-k}; wave 0 transforms e into the outputs 2 j,
// wave 1 transforms o into the outputs 2 j + 1.  Each exchange goes through the
// waves' own transform buffers (free between transforms).
// MODE 0 forward, 1 inverse, 2 forward + spectral update + inverse (as k_col);
// for the pseudospectral_mobility divergence term (gaxis = gradient axis):
// 3 the gradient factor (i k_gaxis) deal on load, then inverse (src -> dst);
// 4 forward, then rho_hat (the divergence) =|+= (i k_gaxis) deal (transform),
// no dst.  Modes 3 and 4 use the per-entry operations of psm_ops.h.
// rhl (modes 0 and 2): the spectral state is stored line-major, entry j of line
// l at l N + j, so that its loads and stores are whole 64-entry rows of the
// line; mode 2 reads and writes rho_hat so, mode 0 writes its output so (the
// initial transform into rho_hat).
template <typename T, int MODE, int LOG2N>
__global__ __launch_bounds__(128) void k_colw2(const typename Complex<T>::type *__restrict__ src,
                                              typename Complex<T>::type *__restrict__ dst,
                                              typename Complex<T>::type *__restrict__ rho_hat,
                                              const typename Complex<T>::type *tw, Params p, int lines, long P,
                                              long ls, long stride, int axis, int gaxis = 0, int rhl = 0) {
    constexpr bool INV_IN = MODE == 1 || MODE == 3;  // input entries k + h n2 (else 2 i + h)
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N - 1>;
    constexpr int E = F::E, n2 = F::n;
    __shared__ T2 lds[2][F::LDS];
    using Stage = TwiddleStage<T2, 2 * n2>;  // table length N = line length
    __shared__ Stage stage;
    const int l = xcd_tile(blockIdx.x, gridDim.x);
    if(l >= lines) return;
    const int lane = threadIdx.x & 63;
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int TWR = Stage::on ? 2 * n2 / 128 : 1;
    T twx[TWR], twy[TWR];  // this thread's share of the table, stored to LDS after the line's loads issue
    if constexpr(Stage::on) {
#pragma unroll
        for(int i = 0; i < TWR; i++) {
            const T2 w = tw[threadIdx.x + 128 * i];
            twx[i] = w.x;
            twy[i] = w.y;
        }
    }
    const int Nh = p.N / 2 + 1;
    const size_t base = (size_t)(l / Nh) * P + (size_t)(l % Nh) * ls;
    const int ts = 2 * (p.N >> LOG2N);  // the n2-point transforms use W_n2^m = W_n^(2m)
    T2 *own = lds[h], *other = lds[h ^ 1];
    T2 v[E], rh[MODE == 2 ? E : 1];
    // forward input: entries 2 i + h; inverse input: entries k + h n2
#pragma unroll
    for(int m = 0; m < E; m++) {
        const int i = lane + 64 * m;
        v[m] = src[base + (size_t)(INV_IN ? i + h * n2 : 2 * i + h) * stride];
    }
    // the line's mode indices: ix = l % Nh; entry j along the pass axis
    auto modes = [&](int j, int &ix, int &iy, int &iz) {
        ix = l % Nh;
        if(axis == 1) { iy = j; iz = l / Nh; }
        else { iy = l / Nh; iz = j; }
    };
    if constexpr(MODE == 3) {
#pragma unroll
        for(int m = 0; m < E; m++) {
            int ix, iy, iz;
            modes(lane + 64 * m + h * n2, ix, iy, iz);
            v[m] = psm_grad_term<T>(p, ix, iy, iz, gaxis, v[m]);
        }
    }
    // entry j of the line in the spectral state
    auto rh_at = [&](int j) -> size_t { return rhl ? (size_t)l * p.N + j : base + (size_t)j * stride; };
    if constexpr(MODE == 2) {
#pragma unroll
        for(int m = 0; m < E; m++) rh[m] = rho_hat[rh_at(lane + 64 * m + h * n2)];
    }
    if constexpr(Stage::on) {
#pragma unroll
        for(int i = 0; i < TWR; i++) stage.t[threadIdx.x + 128 * i] = T2{twx[i], twy[i]};
        tw = stage.t;  // visible after the first barrier
    }
    auto forward = [&]() {
        wave_transform<T2, LOG2N - 1, false>(lane, v, own, tw, ts);
        // wave 1 turns O_k into W_n^k O_k; both publish, then combine
#pragma unroll
        for(int m = 0; m < E; m++) {
            const int k = lane + 64 * m;
            if(h == 1) v[m] = cmul(v[m], tw[k * (ts / 2)]);
        }
        wave_lds_sync();  // the other wave has not read this buffer since the transform's barrier
#pragma unroll
        for(int m = 0; m < E; m++) own[lane + 64 * m] = v[m];
        lds_barrier();
#pragma unroll
        for(int m = 0; m < E; m++) {
            const T2 o = other[lane + 64 * m];
            v[m] = h == 0 ? cadd(v[m], o) : csub(o, v[m]);  // E + WO | E - WO
        }
    };
    auto inverse = [&]() {
        lds_barrier();
#pragma unroll
        for(int m = 0; m < E; m++) own[lane + 64 * m] = v[m];
        lds_barrier();
#pragma unroll
        for(int m = 0; m < E; m++) {
            const int k = lane + 64 * m;
            const T2 o = other[lane + 64 * m];
            if(h == 0) v[m] = cadd(v[m], o);  // X_k + X_{k + n2}
            else v[m] = cmul(csub(o, v[m]), conj_(tw[k * (ts / 2)]));  // (X_k - X_{k + n2}) W^-k
        }
        wave_transform<T2, LOG2N - 1, true>(lane, v, own, tw, ts);
    };
    if constexpr(INV_IN) inverse();
    else forward();
    if constexpr(MODE == 2) {
        const T inv_bins = (T)p.inv_bins;
        const int ix = l % Nh;
#pragma unroll
        for(int m = 0; m < E; m++) {
            const int j = lane + 64 * m + h * n2;
            int iy, iz;
            if(axis == 1) { iy = j; iz = l / Nh; }  // y pass (2D last axis)
            else { iy = l / Nh; iz = j; }            // z pass (3D last axis)
            v[m] = spectral_update<T, T2>(p, ix, iy, iz, v[m], rh[m], inv_bins);
        }
        inverse();
        // the new state after the inverse transform's barriers: a barrier's
        // fence waits for every global store issued before it
#pragma unroll
        for(int m = 0; m < E; m++) rho_hat[rh_at(lane + 64 * m + h * n2)] = rh[m];
    }
    if constexpr(MODE == 4) {
#pragma unroll
        for(int m = 0; m < E; m++) {
            const int j = lane + 64 * m + h * n2;
            int ix, iy, iz;
            modes(j, ix, iy, iz);
            const size_t q = base + (size_t)j * stride;
            rho_hat[q] = psm_div_term<T>(p, ix, iy, iz, gaxis, v[m], gaxis == 0 ? T2{0, 0} : rho_hat[q]);
        }
        return;
    }
    // forward output: entries k + h n2; inverse output: entries 2 j + h
#pragma unroll
    for(int m = 0; m < E; m++) {
        const int i = lane + 64 * m;
        if(MODE == 0 && rhl) dst[rh_at(i + h * n2)] = v[m];
        else if constexpr(MODE == 2) nt_store(&dst[base + (size_t)(2 * i + h) * stride], v[m]);
        else dst[base + (size_t)(MODE == 0 ? i + h * n2 : 2 * i + h) * stride] = v[m];
    }
}

// ---- four waves per line of n = 1024 points (the fused column pass, MODE 2,
// of the 2D N = 1024 chain with the spectral state line-major).  Forward,
// decimation in time by four: wave h transforms the entries 4 i + h
// (i = lane + 64 m) with the 256-point WaveFFT4, G_h(k) = W_n^{h k} F_h(k),
// and one exchange gives wave q the four G_h(k) of k = lane + 64 q, whose
// 4-point DFT is X_{k + 256 q'} (q' = 0..3): wave q holds the entries
// j = lane + 64 q + 256 q'.  Inverse, decimation in frequency: the inverse
// 4-point DFT of X_{k + 256 q'} over q', times W_n^{-h k}, is the input y_h(k)
// of wave h's 256-point inverse transform, whose outputs are the entries
// 4 i + h.  Twice the waves of k_colw2 for the same lines, each with half the
// transform chain: in FP32 the pass takes 10.0 µs against 12.3 µs (used for
// FP32 fields; in FP64 the step measured 2 % slower than with k_colw2).
template <typename T>
__global__ __launch_bounds__(256) void k_colw4(typename Complex<T>::type *__restrict__ W,
                                              typename Complex<T>::type *__restrict__ rho_hat,
                                              const typename Complex<T>::type *tw, Params p, int lines, long P,
                                              long ls, long stride) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT4<T2>;
    constexpr int n = 1024, E = 4;
    __shared__ T2 buf[4][F::LDS];
    using Stage = TwiddleStage<T2, n>;  // table length N = line length
    __shared__ Stage stage;
    const int l = xcd_tile(blockIdx.x, gridDim.x);
    if(l >= lines) return;
    const int lane = threadIdx.x & 63;
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T2 twr[Stage::template per<256>()];
    stage.template load<256>(tw, threadIdx.x, twr);
    const int Nh = p.N / 2 + 1;
    const size_t base = (size_t)(l / Nh) * P + (size_t)(l % Nh) * ls;
    T2 v[E], rh[E];
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = W[base + (size_t)(4 * (lane + 64 * m) + h) * stride];
    // entry j of the line in the line-major spectral state; wave h's entries
    // after the forward transform are j = lane + 64 h + 256 m
#pragma unroll
    for(int m = 0; m < E; m++) rh[m] = rho_hat[(size_t)l * p.N + lane + 64 * h + 256 * m];
    const T2 *t = stage.template store<256>(tw, threadIdx.x, twr);
    T2 *own = buf[h];
    lds_barrier();  // the staged table
    F::template run<false>(lane, v, own, t, 4);
#pragma unroll
    for(int m = 0; m < E; m++)
        if(h > 0) v[m] = cmul(v[m], t[h * (lane + 64 * m)]);  // G_h(k) = W^{h k} F_h(k)
    wave_lds_sync();
#pragma unroll
    for(int m = 0; m < E; m++) own[F::pad(lane + 64 * m)] = v[m];
    lds_barrier();
    // wave h now combines k = lane + 64 h: slice h of every buffer, which no
    // other wave reads or writes until the next barrier
#pragma unroll
    for(int g = 0; g < 4; g++) v[g] = buf[g][F::pad(lane + 64 * h)];
    dft<4, false>(v);
    const T inv_bins = (T)p.inv_bins;
    const int ix = l % Nh;
#pragma unroll
    for(int m = 0; m < E; m++)
        v[m] = spectral_update<T, T2>(p, ix, lane + 64 * h + 256 * m, l / Nh, v[m], rh[m], inv_bins);
    dft<4, true>(v);
    {
        const int k = lane + 64 * h;
#pragma unroll
        for(int g = 0; g < 4; g++) buf[g][F::pad(k)] = g == 0 ? v[0] : cmul(v[g], conj_(t[g * k]));
    }
    lds_barrier();
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = own[F::pad(lane + 64 * m)];
    wave_lds_sync();
    F::template run<true>(lane, v, own, t, 4);
    // the new state and the line, after the last barrier
#pragma unroll
    for(int m = 0; m < E; m++) rho_hat[(size_t)l * p.N + lane + 64 * h + 256 * m] = rh[m];
#pragma unroll
    for(int m = 0; m < E; m++) nt_store(&W[base + (size_t)(4 * (lane + 64 * m) + h) * stride], v[m]);
}

// entry k of the c2r pre-processing: the packed n-point inverse transform
// input Z_k from X_k and X_{n-k} of a row of n + 1 half-spectrum entries
// (table W_N^m, N = 2 n)
template <typename T2>
__device__ __forceinline__ T2 c2r_entry(const T2 *__restrict__ in, long ks, int k, int n, const T2 *__restrict__ tw) {
    T2 xk = in[(size_t)k * ks];
    T2 xnk = in[(size_t)(n - k) * ks];
    if(k == 0) { xk.y = 0; xnk.y = 0; }
    // E = X_k + conj X_{n-k}; O = (X_k - conj X_{n-k}) conj(W_N^k); Z = E + i O
    const T2 Ev = T2{xk.x + xnk.x, xk.y - xnk.y};
    const T2 Ov = cmul(T2{xk.x - xnk.x, xk.y + xnk.y}, conj_(tw[k]));
    return T2{Ev.x - Ov.y, Ev.y + Ov.x};
}

// entry k of the r2c post-processing: X_k from Z_k and Z_{n-k} and w = W_N^k
template <typename T, typename T2>
__device__ __forceinline__ T2 r2c_entry(T2 zk, T2 znk, T2 w) {
    // E = (Z_k + conj Z_{n-k}) / 2, O = (Z_k - conj Z_{n-k}) / (2i), X_k = E + W_N^k O
    const T2 Ev = T2{(zk.x + znk.x) * (T)0.5, (zk.y - znk.y) * (T)0.5};
    const T2 dd = T2{zk.x - znk.x, zk.y + znk.y};
    const T2 Ov = T2{dd.y * (T)0.5, -dd.x * (T)0.5};
    return cadd(Ev, cmul(w, Ov));
}

// c2r pre-processing of a row of n + 1 half-spectrum entries into the packed
// n-point inverse transform input, in distribution D (k = lane + 64 m)
template <typename T, int E>
__device__ __forceinline__ void c2r_pre(const typename Complex<T>::type *__restrict__ in, long ks, int lane,
                                        const typename Complex<T>::type *__restrict__ tw,
                                        typename Complex<T>::type (&z)[E]) {
    constexpr int n = 64 * E;
#pragma unroll
    for(int m = 0; m < E; m++) z[m] = c2r_entry(in, ks, lane + 64 * m, n, tw);
}

// r2c post-processing: Z (distribution D) -> n + 1 half-spectrum entries;
// NT: streaming stores (the pseudospectral chain's work field)
template <typename T, int LOG2N, bool NT = false, typename T2>
__device__ __forceinline__ void r2c_post(const T2 *z, T2 *lds, int lane, const T2 *__restrict__ tw,
                                         T2 *__restrict__ out, long ks) {
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E, n = F::n;
    wave_lds_sync();  // one-wave blocks: the buffer is the wave's own
#pragma unroll
    for(int m = 0; m < E; m++) lds[F::pad(lane + 64 * m)] = z[m];
    wave_lds_sync();
    auto emit = [&](int k, T2 zk, T2 znk) {
        if constexpr(NT) nt_store(&out[(size_t)k * ks], r2c_entry<T>(zk, znk, tw[k]));
        else out[(size_t)k * ks] = r2c_entry<T>(zk, znk, tw[k]);
    };
#pragma unroll
    for(int m = 0; m < E; m++) {
        const int k = lane + 64 * m;
        emit(k, z[m], lds[F::pad(k == 0 ? 0 : n - k)]);
    }
    if(lane == 0) emit(n, z[0], z[0]);
}

// x pass, forward (as k_xfwd), one row per 64-thread block; n = 2^LOG2N = N / 2
template <typename T, int MODEL, bool FPRIME, int LOG2N>
__global__ __launch_bounds__(64) void k_xfwdw(const T *__restrict__ in, typename Complex<T>::type *__restrict__ out,
                                              const typename Complex<T>::type *tw, Params p, int rows,
                                              long rs, long ks) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E, n = F::n;
    __shared__ T2 lds[F::LDS];
    using Stage = TwiddleStage<T2, 2 * n>;  // table length N = 2 n
    __shared__ Stage stage;
    const int row = xcd_tile(blockIdx.x, gridDim.x);
    if(row >= rows) return;
    const int lane = threadIdx.x;
    T2 twr[Stage::template per<64>()];
    stage.template load<64>(tw, lane, twr);
    const T2 *rin = reinterpret_cast<const T2 *>(in + (size_t)row * 2 * n);
    T2 v[E];
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = rin[lane + 64 * m];
    tw = stage.template store<64>(tw, lane, twr);
    if constexpr(FPRIME) {
#pragma unroll
        for(int m = 0; m < E; m++) v[m] = T2{fprime<T, MODEL>(p, v[m].x), fprime<T, MODEL>(p, v[m].y)};
    }
    wave_transform<T2, LOG2N, false>(lane, v, lds, tw, 2);
    r2c_post<T, LOG2N>(v, lds, lane, tw, out + (size_t)row * rs, ks);
}

// x pass, inverse (as k_xinv)
template <typename T, int LOG2N>
__global__ __launch_bounds__(64) void k_xinvw(const typename Complex<T>::type *__restrict__ in, T *__restrict__ out,
                                              const typename Complex<T>::type *tw, Params p, int rows,
                                              long rs, long ks) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E, n = F::n;
    __shared__ T2 lds[F::LDS];
    using Stage = TwiddleStage<T2, 2 * n>;
    __shared__ Stage stage;
    const int row = xcd_tile(blockIdx.x, gridDim.x);
    if(row >= rows) return;
    const int lane = threadIdx.x;
    T2 twr[Stage::template per<64>()];
    stage.template load<64>(tw, lane, twr);
    T2 v[E];
    c2r_pre<T, E>(in + (size_t)row * rs, ks, lane, tw, v);  // before the first barrier: global table
    const T2 *twl = stage.template store<64>(tw, lane, twr);
    wave_transform<T2, LOG2N, true>(lane, v, lds, twl, 2);
    T2 *rout = reinterpret_cast<T2 *>(out + (size_t)row * 2 * n);
#pragma unroll
    for(int m = 0; m < E; m++) rout[lane + 64 * m] = v[m];
}

// x pass between two steps (as k_xrow): c2r, f', r2c, in place on W.  NR rows
// per block, one per wave, share one staged twiddle table (rows % NR == 0)
template <typename T, int MODEL, int LOG2N, int NR = 1>
__global__ __launch_bounds__(64 * NR) void k_xroww(typename Complex<T>::type *__restrict__ W,
                                                   const typename Complex<T>::type *tw, Params p, int rows,
                                                   long rs, long ks, T *__restrict__ rho_out) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E, n = F::n;
    __shared__ T2 lds[NR][F::LDS];
    using Stage = TwiddleStage<T2, 2 * n>;
    __shared__ Stage stage;
    const int row = xcd_tile(blockIdx.x, gridDim.x) * NR + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if(row - (int)(threadIdx.x >> 6) >= rows) return;  // the whole block lies past the last row
    const int lane = threadIdx.x & 63;
    T2 twr[Stage::template per<64 * NR>()];
    stage.template load<64 * NR>(tw, threadIdx.x, twr);
    T2 *w = W + (size_t)row * rs;
    T2 v[E];
    c2r_pre<T, E>(w, ks, lane, tw, v);  // before the first barrier: global table
    tw = stage.template store<64 * NR>(tw, threadIdx.x, twr);
    T2 *own = lds[threadIdx.x >> 6];
    lds_barrier();  // the staged table; every buffer below is one wave's own
    wave_transform<T2, LOG2N, true, false>(lane, v, own, tw, 2);
    if(rho_out) {
        T2 *ro = reinterpret_cast<T2 *>(rho_out + (size_t)row * 2 * n);
#pragma unroll
        for(int m = 0; m < E; m++) ro[lane + 64 * m] = v[m];
    }
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = T2{fprime<T, MODEL>(p, v[m].x), fprime<T, MODEL>(p, v[m].y)};
    wave_transform<T2, LOG2N, false, false>(lane, v, own, tw, 2);
    r2c_post<T, LOG2N, true>(v, own, lane, tw, w, ks);
}

// x pass of the pseudospectral_mobility flux (as k_xroww with the flux in place
// of f'): c2r of a gradient row, g -> (M - M0) g / N^d (psm_flux), r2c, in place
template <typename T, int LOG2N>
__global__ __launch_bounds__(64) void k_xfluxw(typename Complex<T>::type *__restrict__ W,
                                               const typename Complex<T>::type *tw, Params p, int rows, long rs,
                                               long ks, const T *__restrict__ mob, const double *__restrict__ M0p,
                                               int sub) {
    using T2 = typename Complex<T>::type;
    using F = WaveFFT<T2, LOG2N>;
    constexpr int E = F::E, n = F::n;
    __shared__ T2 lds[F::LDS];
    using Stage = TwiddleStage<T2, 2 * n>;
    __shared__ Stage stage;
    const int row = xcd_tile(blockIdx.x, gridDim.x);
    if(row >= rows) return;
    const int lane = threadIdx.x;
    T2 twr[Stage::template per<64>()];
    stage.template load<64>(tw, lane, twr);
    T2 *w = W + (size_t)row * rs;
    const T2 *mrow = reinterpret_cast<const T2 *>(mob + (size_t)row * 2 * n);
    const T M0 = sub ? (T)*M0p : (T)0, ib = (T)p.inv_bins;
    T2 v[E], mv[E];
#pragma unroll
    for(int m = 0; m < E; m++) mv[m] = mrow[lane + 64 * m];
    c2r_pre<T, E>(w, ks, lane, tw, v);  // before the first barrier: global table
    tw = stage.template store<64>(tw, lane, twr);
    wave_transform<T2, LOG2N, true>(lane, v, lds, tw, 2);
#pragma unroll
    for(int m = 0; m < E; m++) v[m] = T2{psm_flux<T>(v[m].x, mv[m].x, M0, ib), psm_flux<T>(v[m].y, mv[m].y, M0, ib)};
    wave_transform<T2, LOG2N, false>(lane, v, lds, tw, 2);
    r2c_post<T, LOG2N>(v, lds, lane, tw, w, ks);
}

// launch wrappers: the wave kernels where the line length allows, else the LDS ones
template <int LOG2N>
constexpr bool wave_len() { return LOG2N >= 9 && LOG2N <= 11; }

template <typename T, int MODEL, bool FPRIME, int LX>
static void launch_xfwd(const T *in, typename Complex<T>::type *out, const typename Complex<T>::type *tw, const Params &p,
                        int rows, long rs, long ks, hipStream_t s) {
    if constexpr(wave_len<LX>()) {
        hipLaunchKernelGGL((k_xfwdw<T, MODEL, FPRIME, LX>), dim3((rows + 7) / 8 * 8), dim3(64), 0, s, in, out, tw, p, rows,
                           rs, ks);
    }
    else {
        using GX = FftGeom<LX, PackedLines<LX>::value>;
        hipLaunchKernelGGL((k_xfwd<T, MODEL, FPRIME, LX>), dim3((rows + GX::NL - 1) / GX::NL), dim3(GX::NT), 0, s, in, out,
                           tw, p, rows, rs, ks);
    }
}

template <typename T, int LX>
static void launch_xinv(const typename Complex<T>::type *in, T *out, const typename Complex<T>::type *tw, const Params &p,
                        int rows, long rs, long ks, hipStream_t s) {
    if constexpr(wave_len<LX>()) {
        hipLaunchKernelGGL((k_xinvw<T, LX>), dim3((rows + 7) / 8 * 8), dim3(64), 0, s, in, out, tw, p, rows, rs, ks);
    }
    else {
        using GX = FftGeom<LX, PackedLines<LX>::value>;
        hipLaunchKernelGGL((k_xinv<T, LX>), dim3((rows + GX::NL - 1) / GX::NL), dim3(GX::NT), 0, s, in, out, tw, p, rows,
                           rs, ks);
    }
}

template <typename T, int MODEL, int LX>
static void launch_xrow(typename Complex<T>::type *W, const typename Complex<T>::type *tw, const Params &p, int rows,
                        long rs, long ks, T *rho_out, hipStream_t s) {
    if constexpr(wave_len<LX>()) {
        constexpr int NR = 4;  // rows per block sharing one staged table
        if(rows % NR == 0)
            hipLaunchKernelGGL((k_xroww<T, MODEL, LX, NR>), dim3((rows / NR + 7) / 8 * 8), dim3(64 * NR), 0, s, W, tw,
                               p, rows, rs, ks, rho_out);
        else
            hipLaunchKernelGGL((k_xroww<T, MODEL, LX>), dim3((rows + 7) / 8 * 8), dim3(64), 0, s, W, tw, p, rows, rs, ks,
                               rho_out);
    }
    else {
        using GX = FftGeom<LX, PackedLines<LX>::value>;
        hipLaunchKernelGGL((k_xrow<T, MODEL, LX>), dim3((rows + GX::NL - 1) / GX::NL), dim3(GX::NT), 0, s, W, tw, p, rows,
                           rs, ks, rho_out);
    }
}

// ------------------------------------------------------------------ host side
// columns per block of the strided passes: NLC lines of n = N, NT = NLC n / 4
// threads (<= 1024)
template <int LOG2N>
struct ColLines { static constexpr int value = LOG2N >= 11 ? 1 : (LOG2N == 10 ? 2 : (LOG2N >= 8 ? 4 : (1024 >> LOG2N))); };

// k_col grid: the line groups padded to a multiple of the 8 XCDs
static inline int col_blocks(int lines, int nlc) { return ((lines + nlc - 1) / nlc + 7) / 8 * 8; }

// Half-spectrum layout: (z, y, kx) with kx fastest, FFTW's r2c layout
// (PseudospectralCPU.cpp:90-96), in 1D, 2D and 3D.  (A transposed 2D layout,
// whole contiguous lines for the column pass, measured slower overall on
// MI355X: the column pass gained 5 %, the x passes lost 18 %.)
struct SpecLayout {
    long rs, ks;              // x pass: entry (row, kx) at row * rs + kx * ks
    long yP, yls, ystride;    // y pass: line l at (l / Nh) * yP + (l % Nh) * yls, step ystride
    long zP, zls, zstride;    // z pass (3D)
    bool ycontig;
};

static SpecLayout spec_layout(int dims, int N) {
    const long Nh = N / 2 + 1;
    (void)dims;
    return SpecLayout{Nh, 1, (long)N * Nh, 1, Nh, Nh, 1, (long)N * Nh, false};
}

// rhl: the line-major spectral state of the two-wave kernel (N = 1024, 2048)
template <typename T, int MODE, int LOG2N>
static void launch_col(const typename Complex<T>::type *src, typename Complex<T>::type *dst,
                       typename Complex<T>::type *rh, const typename Complex<T>::type *tw, const Params &p, int lines,
                       long P, long ls, long stride, int axis, bool contig, hipStream_t s, bool rhl = false) {
    if constexpr(LOG2N == 10 || LOG2N == 11) {
        (void)contig;
        // FP32 fused column pass at N = 1024: four waves per line (in FP64 the
        // two-wave kernel measured 2 % faster in the chain)
        if constexpr(LOG2N == 10 && MODE == 2 && sizeof(T) == 4)
            if(rhl && axis == 1 && src == dst) {
                hipLaunchKernelGGL((k_colw4<T>), dim3((lines + 7) / 8 * 8), dim3(256), 0, s, dst, rh, tw, p, lines, P,
                                   ls, stride);
                return;
            }
        hipLaunchKernelGGL((k_colw2<T, MODE, LOG2N>), dim3((lines + 7) / 8 * 8), dim3(128), 0, s, src, dst, rh, tw, p,
                           lines, P, ls, stride, axis, 0, rhl ? 1 : 0);
        return;
    }
    (void)rhl;
    if constexpr(wave_len<LOG2N>()) {
        (void)contig;
        hipLaunchKernelGGL((k_colw<T, MODE, LOG2N>), dim3((lines + 7) / 8 * 8), dim3(64), 0, s, src, dst, rh, tw, p, lines,
                           P, ls, stride, axis);
        return;
    }
    constexpr int NLC = ColLines<LOG2N>::value;
    using GC = FftGeom<LOG2N, NLC>;
    const int bc = col_blocks(lines, NLC);
    if(contig)
        hipLaunchKernelGGL((k_col<T, MODE, LOG2N, NLC, true>), dim3(bc), dim3(GC::NT), 0, s, src, dst, rh, tw, p, lines,
                           P, ls, stride, axis);
    else
        hipLaunchKernelGGL((k_col<T, MODE, LOG2N, NLC, false>), dim3(bc), dim3(GC::NT), 0, s, src, dst, rh, tw, p,
                           lines, P, ls, stride, axis);
}

// y pass (2D: the last axis) and z pass (3D) of the whole half spectrum
template <typename T, int MODE, int LOG2N>
static void pass_y(const SpecLayout &L, typename Complex<T>::type *W, typename Complex<T>::type *RH,
                   const typename Complex<T>::type *tw, const Params &p, int lines, hipStream_t s,
                   typename Complex<T>::type *dst = nullptr, bool rhl = false) {
    launch_col<T, MODE, LOG2N>(W, dst ? dst : W, RH, tw, p, lines, L.yP, L.yls, L.ystride, 1, L.ycontig, s, rhl);
}
template <typename T, int MODE, int LOG2N>
static void pass_z(const SpecLayout &L, typename Complex<T>::type *W, typename Complex<T>::type *RH,
                   const typename Complex<T>::type *tw, const Params &p, int lines, hipStream_t s,
                   typename Complex<T>::type *dst = nullptr) {
    launch_col<T, MODE, LOG2N>(W, dst ? dst : W, RH, tw, p, lines, L.zP, L.zls, L.zstride, 2, false, s);
}

// init: rho_hat = r2c(rho).  Otherwise `nsteps` time steps: the x pass of f'
// (rho), then per step the axis passes with the fused spectral update, joined
// to the next step by k_xrow; the last step ends with the c2r into rho.
// tape (may be null): state t + 1 of the chain goes to tape field t + 1 (the
// last one from k_xinv, instead of rho)
template <typename T, int LOG2N>
static hipError_t ps_chain_t(const SpectralArgs &a, hipStream_t s, int64_t nsteps, bool init, T *tape = nullptr) {
    using T2 = typename Complex<T>::type;
    const Params &p = a.p;
    constexpr int N = 1 << LOG2N, Nh = N / 2 + 1;
    constexpr int LX = LOG2N - 1;                     // x pass: complex length N / 2
    const int dims = p.dims;
    const SpecLayout L = spec_layout(dims, N);
    const int rows = dims == 1 ? 1 : (dims == 2 ? N : N * N);
    const int lines = dims == 3 ? N * Nh : Nh;
    const T2 *tw2 = (const T2 *)a.twiddles;
    T *rho = (T *)a.rho;
    T2 *W = (T2 *)a.work, *RH = (T2 *)a.rho_hat;
    const bool landau = a.model == CH_LANDAU;
    const size_t bins = (size_t)rows * N;
    const bool rhl = a.rh_lines && dims == 2 && (LOG2N == 10 || LOG2N == 11);
    if(init) {
        launch_xfwd<T, CH_LANDAU, false, LX>(rho, dims == 1 ? RH : W, tw2, p, rows, L.rs, L.ks, s);
        if(dims == 2) pass_y<T, 0, LOG2N>(L, W, RH, tw2, p, lines, s, RH, rhl);
        if(dims == 3) {
            pass_y<T, 0, LOG2N>(L, W, RH, tw2, p, lines, s);
            pass_z<T, 0, LOG2N>(L, W, RH, tw2, p, lines, s, RH);
        }
        return hipGetLastError();
    }
    if(nsteps < 1) return hipSuccess;
    if(landau)
        launch_xfwd<T, CH_LANDAU, true, LX>(rho, W, tw2, p, rows, L.rs, L.ks, s);
    else
        launch_xfwd<T, CH_SIMPLE_WERTHEIM, true, LX>(rho, W, tw2, p, rows, L.rs, L.ks, s);
    for(int64_t t = 0; t < nsteps; t++) {
        if(dims == 1) {
            hipLaunchKernelGGL((k_update1d<T>), dim3((Nh + 255) / 256), dim3(256), 0, s, W, RH, p, Nh);
        }
        else if(dims == 2) {
            pass_y<T, 2, LOG2N>(L, W, RH, tw2, p, lines, s, nullptr, rhl);
        }
        else {
            // y forward (line l: plane l / Nh, column l % Nh), the fused z pass, y inverse
            pass_y<T, 0, LOG2N>(L, W, RH, tw2, p, lines, s);
            pass_z<T, 2, LOG2N>(L, W, RH, tw2, p, lines, s);
            pass_y<T, 1, LOG2N>(L, W, RH, tw2, p, lines, s);
        }
        if(t + 1 < nsteps) {
            T *ro = tape ? tape + (size_t)(t + 1) * bins : nullptr;
            if(landau)
                launch_xrow<T, CH_LANDAU, LX>(W, tw2, p, rows, L.rs, L.ks, ro, s);
            else
                launch_xrow<T, CH_SIMPLE_WERTHEIM, LX>(W, tw2, p, rows, L.rs, L.ks, ro, s);
        }
    }
    launch_xinv<T, LX>(W, tape ? tape + (size_t)nsteps * bins : rho, tw2, p, rows, L.rs, L.ks, s);
    return hipGetLastError();
}

template <typename T>
static hipError_t ps_chain_p(const SpectralArgs &a, hipStream_t s, int64_t nsteps, bool init, T *tape = nullptr) {
    switch(a.p.log2N) {
    case 2: return ps_chain_t<T, 2>(a, s, nsteps, init, tape);
    case 3: return ps_chain_t<T, 3>(a, s, nsteps, init, tape);
    case 4: return ps_chain_t<T, 4>(a, s, nsteps, init, tape);
    case 5: return ps_chain_t<T, 5>(a, s, nsteps, init, tape);
    case 6: return ps_chain_t<T, 6>(a, s, nsteps, init, tape);
    case 7: return ps_chain_t<T, 7>(a, s, nsteps, init, tape);
    case 8: return ps_chain_t<T, 8>(a, s, nsteps, init, tape);
    case 9: return ps_chain_t<T, 9>(a, s, nsteps, init, tape);
    case 10: return ps_chain_t<T, 10>(a, s, nsteps, init, tape);
    case 11: return ps_chain_t<T, 11>(a, s, nsteps, init, tape);
    case 12: return ps_chain_t<T, 12>(a, s, nsteps, init, tape);
    }
    return hipErrorInvalidValue;
}

hipError_t ps_run(const SpectralArgs &a, hipStream_t s, bool init) {
    if(a.precision == CH_FP64) return ps_chain_p<double>(a, s, 1, init);
    return ps_chain_p<float>(a, s, 1, init);
}

hipError_t ps_steps(const SpectralArgs &a, hipStream_t s, int64_t nsteps) {
    if(a.precision == CH_FP64) return ps_chain_p<double>(a, s, nsteps, false);
    return ps_chain_p<float>(a, s, nsteps, false);
}

hipError_t ps_steps_tape(const SpectralArgs &a, hipStream_t s, int64_t nsteps, void *tape) {
    if(nsteps < 1) return hipSuccess;
    if(a.precision == CH_FP64) return ps_chain_p<double>(a, s, nsteps, false, (double *)tape);
    return ps_chain_p<float>(a, s, nsteps, false, (float *)tape);
}

// ---- whole-field transforms for any dims: r2c (optionally of f'(in)) and
// c2r (unnormalised; the spectrum is overwritten)
template <typename T, int LOG2N>
static hipError_t spec_t(const SpectralArgs &a, const void *real_in, void *spec, void *real_out, bool fprime,
                         bool inverse, hipStream_t s) {
    using T2 = typename Complex<T>::type;
    const Params &p = a.p;
    constexpr int N = 1 << LOG2N, Nh = N / 2 + 1, LX = LOG2N - 1;
    const int dims = p.dims;
    const SpecLayout L = spec_layout(dims, N);
    const int rows = dims == 1 ? 1 : (dims == 2 ? N : N * N);
    const T2 *tw2 = (const T2 *)a.twiddles;
    T2 *W = (T2 *)spec;
    const int lines = dims == 3 ? N * Nh : Nh;
    if(!inverse) {
        if(!fprime)
            launch_xfwd<T, CH_LANDAU, false, LX>((const T *)real_in, W, tw2, p, rows, L.rs, L.ks, s);
        else if(a.model == CH_LANDAU)
            launch_xfwd<T, CH_LANDAU, true, LX>((const T *)real_in, W, tw2, p, rows, L.rs, L.ks, s);
        else
            launch_xfwd<T, CH_SIMPLE_WERTHEIM, true, LX>((const T *)real_in, W, tw2, p, rows, L.rs, L.ks, s);
        if(dims >= 2) pass_y<T, 0, LOG2N>(L, W, W, tw2, p, lines, s);
        if(dims == 3) pass_z<T, 0, LOG2N>(L, W, W, tw2, p, lines, s);
    }
    else {
        if(dims == 3) pass_z<T, 1, LOG2N>(L, W, W, tw2, p, lines, s);
        if(dims >= 2) pass_y<T, 1, LOG2N>(L, W, W, tw2, p, lines, s);
        launch_xinv<T, LX>(W, (T *)real_out, tw2, p, rows, L.rs, L.ks, s);
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t spec_p(const SpectralArgs &a, const void *in, void *spec, void *out, bool fprime, bool inverse,
                         hipStream_t s) {
    switch(a.p.log2N) {
    case 2: return spec_t<T, 2>(a, in, spec, out, fprime, inverse, s);
    case 3: return spec_t<T, 3>(a, in, spec, out, fprime, inverse, s);
    case 4: return spec_t<T, 4>(a, in, spec, out, fprime, inverse, s);
    case 5: return spec_t<T, 5>(a, in, spec, out, fprime, inverse, s);
    case 6: return spec_t<T, 6>(a, in, spec, out, fprime, inverse, s);
    case 7: return spec_t<T, 7>(a, in, spec, out, fprime, inverse, s);
    case 8: return spec_t<T, 8>(a, in, spec, out, fprime, inverse, s);
    case 9: return spec_t<T, 9>(a, in, spec, out, fprime, inverse, s);
    case 10: return spec_t<T, 10>(a, in, spec, out, fprime, inverse, s);
    case 11: return spec_t<T, 11>(a, in, spec, out, fprime, inverse, s);
    case 12: return spec_t<T, 12>(a, in, spec, out, fprime, inverse, s);
    }
    return hipErrorInvalidValue;
}

hipError_t spec_r2c(const SpectralArgs &a, const void *real_in, void *spec_out, bool fprime, hipStream_t s) {
    if(a.precision == CH_FP64) return spec_p<double>(a, real_in, spec_out, nullptr, fprime, false, s);
    return spec_p<float>(a, real_in, spec_out, nullptr, fprime, false, s);
}

hipError_t spec_c2r(const SpectralArgs &a, void *spec, void *real_out, hipStream_t s) {
    if(a.precision == CH_FP64) return spec_p<double>(a, nullptr, spec, real_out, false, true, s);
    return spec_p<float>(a, nullptr, spec, real_out, false, true, s);
}

// ---- pseudospectral_mobility divergence term with the element-wise work
// folded into the transform passes (2D, N = 1024 or 2048): per axis d, the
// inverse column pass with the gradient factor (k_colw2 mode 3, src -> gh),
// the row pass c2r -> flux -> r2c (k_xfluxw, in place on gh) and the forward
// column pass accumulating the divergence (k_colw2 mode 4, gh -> divh): three
// launches in place of seven, the same operations in the same order
template <typename T, int LOG2N>
static hipError_t psm_div_t(const SpectralArgs &a, const void *src, void *gh, void *divh, const void *mob,
                            const double *M0p, int sub, hipStream_t s) {
    using T2 = typename Complex<T>::type;
    const Params &p = a.p;
    constexpr int N = 1 << LOG2N, Nh = N / 2 + 1, LX = LOG2N - 1;
    const SpecLayout L = spec_layout(2, N);
    const T2 *tw = (const T2 *)a.twiddles;
    const int colb = (Nh + 7) / 8 * 8, rowb = (N + 7) / 8 * 8;
    for(int d = 0; d < 2; d++) {
        hipLaunchKernelGGL((k_colw2<T, 3, LOG2N>), dim3(colb), dim3(128), 0, s, (const T2 *)src, (T2 *)gh, nullptr, tw,
                           p, Nh, L.yP, L.yls, L.ystride, 1, d);
        hipLaunchKernelGGL((k_xfluxw<T, LX>), dim3(rowb), dim3(64), 0, s, (T2 *)gh, tw, p, N, L.rs, L.ks, (const T *)mob,
                           M0p, sub);
        hipLaunchKernelGGL((k_colw2<T, 4, LOG2N>), dim3(colb), dim3(128), 0, s, (const T2 *)gh, nullptr, (T2 *)divh, tw,
                           p, Nh, L.yP, L.yls, L.ystride, 1, d);
    }
    return hipGetLastError();
}

hipError_t psm_div_fused(const SpectralArgs &a, const void *src, void *gh, void *divh, const void *mob,
                         const double *M0p, int sub, hipStream_t s) {
    if(a.p.dims != 2 || (a.p.log2N != 10 && a.p.log2N != 11)) return hipErrorNotSupported;
    if(a.precision == CH_FP64) {
        if(a.p.log2N == 10) return psm_div_t<double, 10>(a, src, gh, divh, mob, M0p, sub, s);
        return psm_div_t<double, 11>(a, src, gh, divh, mob, M0p, sub, s);
    }
    if(a.p.log2N == 10) return psm_div_t<float, 10>(a, src, gh, divh, mob, M0p, sub, s);
    return psm_div_t<float, 11>(a, src, gh, divh, mob, M0p, sub, s);
}

// ---- 2D inverse transform (column IFFT + c2r) of a.work into `out`
template <typename T, int LOG2N>
static hipError_t fft2_inverse_t(const SpectralArgs &a, void *out, hipStream_t s) {
    using T2 = typename Complex<T>::type;
    constexpr int N = 1 << LOG2N, Nh = N / 2 + 1, LX = LOG2N - 1;
    const SpecLayout L = spec_layout(2, N);
    const T2 *tw2 = (const T2 *)a.twiddles;
    T2 *W = (T2 *)a.work;
    pass_y<T, 1, LOG2N>(L, W, W, tw2, a.p, Nh, s);
    launch_xinv<T, LX>(W, (T *)out, tw2, a.p, N, L.rs, L.ks, s);
    return hipGetLastError();
}

template <typename T>
static hipError_t fft2_inverse_p(const SpectralArgs &a, void *out, hipStream_t s) {
    switch(a.p.log2N) {
    case 2: return fft2_inverse_t<T, 2>(a, out, s);
    case 3: return fft2_inverse_t<T, 3>(a, out, s);
    case 4: return fft2_inverse_t<T, 4>(a, out, s);
    case 5: return fft2_inverse_t<T, 5>(a, out, s);
    case 6: return fft2_inverse_t<T, 6>(a, out, s);
    case 7: return fft2_inverse_t<T, 7>(a, out, s);
    case 8: return fft2_inverse_t<T, 8>(a, out, s);
    case 9: return fft2_inverse_t<T, 9>(a, out, s);
    case 10: return fft2_inverse_t<T, 10>(a, out, s);
    case 11: return fft2_inverse_t<T, 11>(a, out, s);
    case 12: return fft2_inverse_t<T, 12>(a, out, s);
    }
    return hipErrorInvalidValue;
}

hipError_t fft2_forward(const SpectralArgs &a, hipStream_t s) {
    SpectralArgs b = a;
    b.rho_hat = a.work;  // the forward r2c of ps_run's set-up, written into work
    return ps_run(b, s, true);
}

hipError_t fft2_inverse(const SpectralArgs &a, void *out, hipStream_t s) {
    if(a.precision == CH_FP64) return fft2_inverse_p<double>(a, out, s);
    return fft2_inverse_p<float>(a, out, s);
}

// ---- circularly averaged structure factor (utils/sq_from_traj.py:72-103)
// Shell of the mode (x-half index kx, row y): rint(sqrt(kx^2 + iy^2)) with iy the
// signed frequency of y; every half-spectrum entry with 0 < kx < N/2 stands for
// two modes of the full spectrum.  One block per shell visits the candidate
// modes of its annulus in a fixed order (deterministic sums).
__device__ __forceinline__ int sq_shell(int kx, int iy) { return (int)rint(sqrt((double)(kx * kx + iy * iy))); }

template <typename T>
__global__ __launch_bounds__(256) void k_sq_shells(const typename Complex<T>::type *__restrict__ W, int N,
                                                   double *__restrict__ sq, double *__restrict__ cnt) {
    const int s = blockIdx.x, Nh = N / 2 + 1;
    const int r = min(s + 1, N / 2);
    const int nkx = r + 1, niy = 2 * r + 1;  // kx in [0, r], iy in [-r, r]
    double acc = 0, c = 0;
    const double inv_n = 1.0 / ((double)N * N);
    for(int q = threadIdx.x; q < nkx * niy; q += 256) {
        const int kx = q / niy, iy = q % niy - r;
        if(iy >= N / 2 || iy < -N / 2 || sq_shell(kx, iy) != s) continue;
        const int y = iy < 0 ? iy + N : iy;
        const auto f = W[(size_t)y * Nh + kx];
        const double w = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
        const double a2 = (kx == 0 && iy == 0) ? 0.0 : ((double)f.x * f.x + (double)f.y * f.y);  // mean removed
        acc += w * a2 * inv_n;
        c += w;
    }
    __shared__ double sa[256], sc[256];
    sa[threadIdx.x] = acc;
    sc[threadIdx.x] = c;
    __syncthreads();
    for(int h = 128; h > 0; h >>= 1) {
        if(threadIdx.x < h) {
            sa[threadIdx.x] += sa[threadIdx.x + h];
            sc[threadIdx.x] += sc[threadIdx.x + h];
        }
        __syncthreads();
    }
    if(threadIdx.x == 0) {
        sq[s] = sc[0] > 0 ? sa[0] / sc[0] : 0.0;
        cnt[s] = sc[0];
    }
}

// W_k <- 2 G_shell W_k with G_s = 2 (S_s - T_s) / (count_s N^2), DC set to 0:
// the c2r of the result is dLoss/dpsi for Loss = sum_s (S_s - T_s)^2
template <typename T>
__global__ void k_sq_scale(typename Complex<T>::type *__restrict__ W, int N, const double *__restrict__ sq,
                           const double *__restrict__ target, const double *__restrict__ cnt) {
    const int Nh = N / 2 + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= (int64_t)N * Nh) return;
    const int y = (int)(i / Nh), kx = (int)(i % Nh);
    const int iy = y < N / 2 ? y : y - N;
    const int s = sq_shell(kx, iy);
    double g = 0.0;
    if(!(kx == 0 && iy == 0)) g = 2.0 * 2.0 * (sq[s] - target[s]) / (cnt[s] * (double)N * N);
    auto f = W[i];
    f.x = (T)(g * (double)f.x);
    f.y = (T)(g * (double)f.y);
    W[i] = f;
}

hipError_t sq_shells(int precision, const void *W, int N, int nshells, double *sq, double *cnt, hipStream_t s) {
    if(precision == CH_FP64) hipLaunchKernelGGL((k_sq_shells<double>), dim3(nshells), dim3(256), 0, s, (const double2 *)W, N, sq, cnt);
    else hipLaunchKernelGGL((k_sq_shells<float>), dim3(nshells), dim3(256), 0, s, (const float2 *)W, N, sq, cnt);
    return hipGetLastError();
}

hipError_t sq_scale(int precision, void *W, int N, const double *sq, const double *target, const double *cnt,
                    hipStream_t s) {
    const int64_t n = (int64_t)N * (N / 2 + 1);
    const int g = (int)((n + 255) / 256);
    if(precision == CH_FP64) hipLaunchKernelGGL((k_sq_scale<double>), dim3(g), dim3(256), 0, s, (double2 *)W, N, sq, target, cnt);
    else hipLaunchKernelGGL((k_sq_scale<float>), dim3(g), dim3(256), 0, s, (float2 *)W, N, sq, target, cnt);
    return hipGetLastError();
}

}  // namespace chd
