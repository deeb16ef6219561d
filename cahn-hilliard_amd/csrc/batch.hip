// batch.hip -- many independent 2D Euler replicas (mixed N and parameters)
// advanced together.
//
// The reference runs one system per process (src/main.cpp:182-200).  Here a
// ch_batch holds every replica's field inside two shared ping-pong buffers and
// a work list of (replica, tile) items; one launch advances every item by K
// steps with the register/DPP wave kernel (stencil_wave.h).  A replica with
// N = 64 is one item covering the whole periodic grid (no halo); larger
// replicas are covered by halo tiles.  Every item is a 64 x 64 region, so the
// blocks of a launch cost the same whatever the replica sizes.  The region is
// split over the waves in one of kBatchShapes ways; the shape and K are picked
// by timing the candidates on the batch itself (ch_batch_autotune).
// Per-replica epsilon, k, dt, M, dx and N travel in a device array of Params.
#include "ch_internal.h"
#include "stencil_wave.h"

namespace chd {

// item = (replica, tx | ty << 16): one 64 x 64 region (whole replica for N = 64)
template <typename T, int MODEL, int K, int RPW, int NW>
__global__ __launch_bounds__(64 * NW) void k_batch(const T *__restrict__ bufA, T *__restrict__ bufB,
                                                  const BatchRep *__restrict__ reps, const int2 *__restrict__ items) {
    static_assert(RPW * NW == kBatchRegion, "batch regions are 64 x 64");
    constexpr int W = kBatchRegion - 4 * K, TH = kBatchRegion - 4 * K;
    using CFG = WaveCfg<1, RPW, NW>;
    __shared__ WaveExchange<T, 2, NW, 1> ex;
    const int2 it = items[blockIdx.x];
    const Params p = reps[it.x].p;
    const int64_t off = reps[it.x].off;
    const T *in = bufA + off;
    T *out = bufB + off;
    if(reps[it.x].whole) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, CFG, true>(in, out, p, 0, 0, ex);
    }
    else {
        const int tx = it.y & 0xFFFF, ty = it.y >> 16;
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, CFG, false>(in, out, p, tx * W - 2 * K, ty * TH - 2 * K,
                                                                          ex);
    }
}

int batch_tile(int K) { return kBatchRegion - 4 * K; }

// shape s: 0 = 16 rows x 4 waves, 1 = 8 x 8, 2 = 4 x 16
template <typename T, int MODEL, int K>
static hipError_t launch_batch_k(int shape, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, hipStream_t s) {
    switch(shape) {
    case 0: hipLaunchKernelGGL((k_batch<T, MODEL, K, 16, 4>), dim3(nitems), dim3(256), 0, s, (const T *)a, (T *)b, reps, items); break;
    case 1: hipLaunchKernelGGL((k_batch<T, MODEL, K, 8, 8>), dim3(nitems), dim3(512), 0, s, (const T *)a, (T *)b, reps, items); break;
    case 2: hipLaunchKernelGGL((k_batch<T, MODEL, K, 4, 16>), dim3(nitems), dim3(1024), 0, s, (const T *)a, (T *)b, reps, items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int MODEL>
static hipError_t launch_batch_t(int shape, int K, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, hipStream_t s) {
    switch(K) {
    case 1: return launch_batch_k<T, MODEL, 1>(shape, a, b, reps, items, nitems, s);
    case 2: return launch_batch_k<T, MODEL, 2>(shape, a, b, reps, items, nitems, s);
    case 4: return launch_batch_k<T, MODEL, 4>(shape, a, b, reps, items, nitems, s);
    case 8: return launch_batch_k<T, MODEL, 8>(shape, a, b, reps, items, nitems, s);
    }
    return hipErrorInvalidValue;
}

hipError_t batch_launch(int precision, int model, int shape, int K, const void *a, void *b, const BatchRep *reps,
                        const int2 *items, int nitems, hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return launch_batch_t<double, CH_LANDAU>(shape, K, a, b, reps, items, nitems, s);
        return launch_batch_t<double, CH_SIMPLE_WERTHEIM>(shape, K, a, b, reps, items, nitems, s);
    }
    if(model == CH_LANDAU) return launch_batch_t<float, CH_LANDAU>(shape, K, a, b, reps, items, nitems, s);
    return launch_batch_t<float, CH_SIMPLE_WERTHEIM>(shape, K, a, b, reps, items, nitems, s);
}

}  // namespace chd
