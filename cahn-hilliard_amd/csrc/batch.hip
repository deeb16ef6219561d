// batch.hip -- many independent 2D Euler replicas (mixed N and parameters)
// advanced together.
//
// The reference runs one system per process (src/main.cpp:182-200).  Here a
// ch_batch holds every replica's field inside two shared ping-pong buffers and
// a work list of (replica, tile) items; one launch advances every item by K
// steps with the register/DPP wave kernel (stencil_wave.h).  A replica with
// N = 64 is one item covering the whole periodic grid (no halo); larger
// replicas are covered by halo tiles.  Every item is a 64 x (NW RPW) region,
// so the blocks of a launch cost the same whatever the replica sizes.
// Per-replica epsilon, k, dt, M, dx and N travel in a device array of Params.
#include "ch_internal.h"
#include "stencil_wave.h"

namespace chd {

template <typename T, int MODEL, int K, int RPW, int NW>
__global__ __launch_bounds__(64 * NW) void k_batch(const T *__restrict__ bufA, T *__restrict__ bufB,
                                                  const BatchRep *__restrict__ reps, const int2 *__restrict__ items) {
    constexpr int W = 64 - 4 * K, TH = RPW * NW - 4 * K;
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
        const int tiles_x = reps[it.x].tiles_x;
        const int tx = it.y % tiles_x, ty = it.y / tiles_x;
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, CFG, false>(in, out, p, tx * W - 2 * K, ty * TH - 2 * K,
                                                                          ex);
    }
}

constexpr int kBatchRPW = 16, kBatchNW = 4;

int batch_tile_w(int K) { return 64 - 4 * K; }
int batch_tile_h(int K) { return kBatchRPW * kBatchNW - 4 * K; }

template <typename T, int MODEL>
static hipError_t launch_batch_t(int K, const void *a, void *b, const BatchRep *reps, const int2 *items, int nitems,
                                 hipStream_t s) {
    switch(K) {
    case 1: hipLaunchKernelGGL((k_batch<T, MODEL, 1, kBatchRPW, kBatchNW>), dim3(nitems), dim3(64 * kBatchNW), 0, s, (const T *)a, (T *)b, reps, items); break;
    case 2: hipLaunchKernelGGL((k_batch<T, MODEL, 2, kBatchRPW, kBatchNW>), dim3(nitems), dim3(64 * kBatchNW), 0, s, (const T *)a, (T *)b, reps, items); break;
    case 4: hipLaunchKernelGGL((k_batch<T, MODEL, 4, kBatchRPW, kBatchNW>), dim3(nitems), dim3(64 * kBatchNW), 0, s, (const T *)a, (T *)b, reps, items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t batch_launch(int precision, int model, int K, const void *a, void *b, const BatchRep *reps, const int2 *items,
                        int nitems, hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU) return launch_batch_t<double, CH_LANDAU>(K, a, b, reps, items, nitems, s);
        return launch_batch_t<double, CH_SIMPLE_WERTHEIM>(K, a, b, reps, items, nitems, s);
    }
    if(model == CH_LANDAU) return launch_batch_t<float, CH_LANDAU>(K, a, b, reps, items, nitems, s);
    return launch_batch_t<float, CH_SIMPLE_WERTHEIM>(K, a, b, reps, items, nitems, s);
}

}  // namespace chd
