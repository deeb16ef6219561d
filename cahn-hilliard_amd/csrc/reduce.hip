// reduce.hip -- FP64 observables and pointwise model kernels.
//
// average_mass and average_free_energy (src/CahnHilliard.cpp:285-319) are
// evaluated on the device in FP64 whatever the field precision: each thread
// accumulates a grid-stride slice, waves combine with 64-lane shuffles, waves
// of a block through LDS, and a single-block pass sums the per-block partials
// in a fixed order (deterministic, no atomics).  Only two doubles return to
// the host, instead of the reference's blocking copy of the whole field and a
// serial host loop (CUDAIntegrator.cu:55-70).
#include <mutex>

#include "ch_internal.h"

namespace chd {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// adds cell i's density and free-energy density (bulk + k sum_d (grad_d+)^2,
// forward differences as CahnHilliard.cpp:256-276, 300-319) to m and f
template <typename T, int MODEL, int DIMS>
__device__ __forceinline__ void obs_cell(const T *__restrict__ rho, const Params &p, int64_t i, double &m, double &f) {
    const int mask = p.mask, lg = p.log2N;
    const double r = (double)rho[i];
    m += r;
    const int x = (int)(i & mask);
    double inter = 0.;
    {
        const int64_t ip = (i - x) + ((x + 1) & mask);
        const double g = ((double)rho[ip] - r) * p.inv_dx;
        inter += p.k * g * g;
    }
    if constexpr(DIMS >= 2) {
        const int y = (int)((i >> lg) & mask);
        const int64_t ip = i + ((int64_t)(((y + 1) & mask) - y) << lg);
        const double g = ((double)rho[ip] - r) * p.inv_dx;
        inter += p.k * g * g;
    }
    if constexpr(DIMS == 3) {
        const int z = (int)(i >> (2 * lg));
        const int64_t ip = i + ((int64_t)(((z + 1) & mask) - z) << (2 * lg));
        const double g = ((double)rho[ip] - r) * p.inv_dx;
        inter += p.k * g * g;
    }
    f += bulk<MODEL>(p, r) + inter;
}

// the sums of n (mass, f) pairs, one block, in a fixed order.  Each thread
// adds pairs t, t + 256, ... in that order, four per round with all their
// loads issued together (a missing pair adds +0.0, which changes no sum).
__device__ __forceinline__ void sum_pairs_block(const double *__restrict__ partials, int n, double &m, double &f) {
    m = 0;
    f = 0;
    for(int i0 = threadIdx.x; i0 < n; i0 += 4 * 256) {
        double pm[4], pf[4];
#pragma unroll
        for(int u = 0; u < 4; u++) {
            const int i = i0 + 256 * u;
            const bool in = i < n;
            pm[u] = in ? partials[2 * i] : 0.0;
            pf[u] = in ? partials[2 * i + 1] : 0.0;
        }
#pragma unroll
        for(int u = 0; u < 4; u++) {
            m += pm[u];
            f += pf[u];
        }
    }
    __shared__ double sm[4], sf[4];
    m = wave_sum(m);
    f = wave_sum(f);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if(lane == 0) { sm[w] = m; sf[w] = f; }
    __syncthreads();
    m = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    f = (sf[0] + sf[1]) + (sf[2] + sf[3]);
}

// a grid-stride loop over the cells, four strides per round so that the
// loads of all four cells issue before the first is used; the cells are
// added in the plain grid-stride order
template <typename T, int MODEL, int DIMS>
__global__ __launch_bounds__(256) void k_observables(const T *__restrict__ rho, Params p, double *__restrict__ partials) {
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int64_t stride = (int64_t)gridDim.x * 256;
    double m = 0, f = 0;
    for(int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x; i0 < bins; i0 += 4 * stride) {
        if(i0 + 3 * stride < bins) {
#pragma unroll
            for(int u = 0; u < 4; u++) obs_cell<T, MODEL, DIMS>(rho, p, i0 + u * stride, m, f);
        }
        else {
            for(int64_t i = i0; i < bins; i += stride) obs_cell<T, MODEL, DIMS>(rho, p, i, m, f);
        }
    }
    __shared__ double sm[4], sf[4];
    m = wave_sum(m);
    f = wave_sum(f);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if(lane == 0) { sm[w] = m; sf[w] = f; }
    __syncthreads();
    if(threadIdx.x == 0) {
        partials[2 * blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
        partials[2 * blockIdx.x + 1] = (sf[0] + sf[1]) + (sf[2] + sf[3]);
    }
}

__global__ __launch_bounds__(256) void k_finalize(const double *__restrict__ partials, int nparts,
                                                  double *__restrict__ result) {
    double m, f;
    sum_pairs_block(partials, nparts, m, f);
    if(threadIdx.x == 0) {
        result[0] = m;
        result[1] = f;
    }
}

// the same sums into slot (*count % CH_OBS_QUEUE) of a ring of pairs (host
// memory the device writes), then ++*count: a captured graph that queues
// observables needs no argument of its own per replay, and the host, which
// counts the queued pairs in the same order, knows each pair's slot
__global__ __launch_bounds__(256) void k_finalize_ring(const double *__restrict__ partials, int nparts,
                                                       double *__restrict__ ring, unsigned *__restrict__ count) {
    double m, f;
    sum_pairs_block(partials, nparts, m, f);
    if(threadIdx.x == 0) {
        const unsigned c = *count;
        const unsigned slot = c % CH_OBS_QUEUE;
        ring[2 * slot] = m;
        ring[2 * slot + 1] = f;
        *count = c + 1;
    }
}

template <typename T, int MODEL>
static void obs_launch(const Params &p, const void *rho, double *partials, int nb, hipStream_t s) {
    const T *r = (const T *)rho;
    if(p.dims == 1) hipLaunchKernelGGL((k_observables<T, MODEL, 1>), dim3(nb), dim3(256), 0, s, r, p, partials);
    else if(p.dims == 2) hipLaunchKernelGGL((k_observables<T, MODEL, 2>), dim3(nb), dim3(256), 0, s, r, p, partials);
    else hipLaunchKernelGGL((k_observables<T, MODEL, 3>), dim3(nb), dim3(256), 0, s, r, p, partials);
}

static hipError_t obs_partials(const Params &p, int precision, int model, const void *rho, double *partials, int nb,
                               hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) obs_launch<double, CH_LANDAU>(p, rho, partials, nb, s);
        else obs_launch<double, CH_SIMPLE_WERTHEIM>(p, rho, partials, nb, s);
    }
    else {
        if(model == CH_LANDAU) obs_launch<float, CH_LANDAU>(p, rho, partials, nb, s);
        else obs_launch<float, CH_SIMPLE_WERTHEIM>(p, rho, partials, nb, s);
    }
    return hipGetLastError();
}

hipError_t observables(const Params &p, int precision, int model, const void *rho, double *partials, int nparts,
                       double *result, hipStream_t s) {
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int nb = (int)std::min<int64_t>(nparts, (bins + 255) / 256);
    hipError_t e = obs_partials(p, precision, model, rho, partials, nb, s);
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, nb, result);
    return hipGetLastError();
}

hipError_t observables_ring(const Params &p, int precision, int model, const void *rho, double *partials, int nparts,
                            double *ring, unsigned *count, hipStream_t s) {
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    const int nb = (int)std::min<int64_t>(nparts, (bins + 255) / 256);
    hipError_t e = obs_partials(p, precision, model, rho, partials, nb, s);
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize_ring, dim3(1), dim3(256), 0, s, partials, nb, ring, count);
    return hipGetLastError();
}

hipError_t finalize_pairs_ring(double *partials, int nparts, double *ring, unsigned *count, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize_ring, dim3(1), dim3(256), 0, s, partials, nparts, ring, count);
    return hipGetLastError();
}

// ---- the allocation-free variant of the observables (ch_observables_device_noalloc)
// One fixed array per device holds the kObsParts partial pairs and the final
// pair: 2 kObsParts + 2 doubles (16,400 bytes), kept until the process ends.
// Host code never passes the array's name to a kernel: noalloc_scratch_init
// looks its device address up once per device (with that device current) and
// noalloc_scratch returns that address.  Every call on a device shares the
// array, so calls that overlap on one device race.
static __device__ __align__(256) double scratch[2 * kObsParts + 2];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

hipError_t noalloc_scratch_init(int device) {
    if(device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if(e != hipSuccess) return e;
    if(!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

double *noalloc_scratch(int device) {
    if(device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return static_cast<double *>(g_scratch_ptr[device]);
}

hipError_t finalize_pairs(double *partials, int nparts, double *result, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, nparts, result);
    return hipGetLastError();
}

// one block per replica of a batch
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_batch_obs(const T *__restrict__ buf, const BatchRep *__restrict__ reps,
                                                   double *__restrict__ result) {
    const Params p = reps[blockIdx.x].p;
    const T *rho = buf + reps[blockIdx.x].off;
    const int64_t bins = (int64_t)1 << (p.log2N * p.dims);
    double m = 0, f = 0;
    for(int64_t i = threadIdx.x; i < bins; i += 256) obs_cell<T, MODEL, 2>(rho, p, i, m, f);  // batches are 2D
    __shared__ double sm[4], sf[4];
    m = wave_sum(m);
    f = wave_sum(f);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if(lane == 0) { sm[w] = m; sf[w] = f; }
    __syncthreads();
    if(threadIdx.x == 0) {
        result[2 * blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
        result[2 * blockIdx.x + 1] = (sf[0] + sf[1]) + (sf[2] + sf[3]);
    }
}

hipError_t batch_observables(int precision, int model, const void *buf, const BatchRep *reps, int n, double *result,
                             hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) hipLaunchKernelGGL((k_batch_obs<double, CH_LANDAU>), dim3(n), dim3(256), 0, s, (const double *)buf, reps, result);
        else hipLaunchKernelGGL((k_batch_obs<double, CH_SIMPLE_WERTHEIM>), dim3(n), dim3(256), 0, s, (const double *)buf, reps, result);
    }
    else {
        if(model == CH_LANDAU) hipLaunchKernelGGL((k_batch_obs<float, CH_LANDAU>), dim3(n), dim3(256), 0, s, (const float *)buf, reps, result);
        else hipLaunchKernelGGL((k_batch_obs<float, CH_SIMPLE_WERTHEIM>), dim3(n), dim3(256), 0, s, (const float *)buf, reps, result);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- pointwise
template <typename T, int MODEL>
__global__ void k_fprime(const T *__restrict__ in, T *__restrict__ out, Params p, int64_t n) {
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = fprime<T, MODEL>(p, in[i]);
}

template <typename T, int MOB>
__global__ void k_mobility(const T *__restrict__ in, T *__restrict__ out, Params p, int64_t n) {
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = mobility<T, MOB>(p, in[i]);
}

static int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

hipError_t pointwise_fprime(const Params &p, int precision, int model, const void *in, void *out, int64_t n,
                            hipStream_t s) {
    const int g = grid_for(n);
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) hipLaunchKernelGGL((k_fprime<double, CH_LANDAU>), dim3(g), dim3(256), 0, s, (const double *)in, (double *)out, p, n);
        else hipLaunchKernelGGL((k_fprime<double, CH_SIMPLE_WERTHEIM>), dim3(g), dim3(256), 0, s, (const double *)in, (double *)out, p, n);
    }
    else {
        if(model == CH_LANDAU) hipLaunchKernelGGL((k_fprime<float, CH_LANDAU>), dim3(g), dim3(256), 0, s, (const float *)in, (float *)out, p, n);
        else hipLaunchKernelGGL((k_fprime<float, CH_SIMPLE_WERTHEIM>), dim3(g), dim3(256), 0, s, (const float *)in, (float *)out, p, n);
    }
    return hipGetLastError();
}

template <typename T>
static void launch_mob(const Params &p, int mob, const T *in, T *out, int64_t n, hipStream_t s) {
    const int g = grid_for(n);
    switch(mob) {
    case CH_MOB_CONSTANT: hipLaunchKernelGGL((k_mobility<T, CH_MOB_CONSTANT>), dim3(g), dim3(256), 0, s, in, out, p, n); break;
    case CH_MOB_FREE_ENERGY: hipLaunchKernelGGL((k_mobility<T, CH_MOB_FREE_ENERGY>), dim3(g), dim3(256), 0, s, in, out, p, n); break;
    case CH_MOB_REGULARISED: hipLaunchKernelGGL((k_mobility<T, CH_MOB_REGULARISED>), dim3(g), dim3(256), 0, s, in, out, p, n); break;
    default: hipLaunchKernelGGL((k_mobility<T, CH_MOB_LINEAR>), dim3(g), dim3(256), 0, s, in, out, p, n); break;
    }
}

hipError_t pointwise_mobility(const Params &p, int precision, int mob, const void *in, void *out, int64_t n,
                              hipStream_t s) {
    if(precision == CH_FP64) launch_mob<double>(p, mob, (const double *)in, (double *)out, n, s);
    else launch_mob<float>(p, mob, (const float *)in, (float *)out, n, s);
    return hipGetLastError();
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ in, B *__restrict__ out, int64_t n) {
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (B)in[i];
}

hipError_t convert_d2f(const double *in, float *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t convert_f2d(const float *in, double *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace chd
