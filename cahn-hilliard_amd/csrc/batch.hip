// batch.hip -- many independent 2D Euler replicas (mixed N and parameters)
// advanced together.
//
// The reference runs one system per process (src/main.cpp:182-200).  Here a
// ch_batch holds every replica's field inside two shared ping-pong buffers and
// a work list of (replica, tile) items; one launch advances every item by K
// steps with the register/DPP wave kernel (stencil_wave.h).  A replica with
// N = 64 is one item covering the whole periodic grid (no halo); larger
// replicas are covered by halo tiles.  Halo items are 64 x 64 regions (one
// column per lane) or, for the wide shapes, 128 x 64 or 192 x 64 regions (two
// or three columns per lane: a larger share of each region is interior, and
// one DPP move serves several cells); a whole N = 64 replica is one 64 x 64
// region split over the same number of waves.  The region is split over the waves in one of kBatchShapes
// ways; the shape and K are picked by timing the candidates on the batch itself
// (ch_batch_autotune).
// Per-replica epsilon, k, dt, M, dx and N travel in a device array of Params.
#include "ch_internal.h"
#include "stencil_wave.h"

namespace chd {

// item = (replica, tx | ty << 16): one halo region of C x 64 columns and 64 rows,
// or the whole replica for N = 64 (64 x 64, one column per lane)
// COMP: compensated updates, the lost low-order parts in compA -> compB
template <typename T, int MODEL, int K, int C, int RPW, int NW, bool COMP = false>
__global__ __launch_bounds__(64 * NW) void k_batch(const T *__restrict__ bufA, T *__restrict__ bufB,
                                                  const BatchRep *__restrict__ reps, const int2 *__restrict__ items,
                                                  const T *__restrict__ compA, T *__restrict__ compB) {
    static_assert(RPW * NW == kBatchRegion, "batch regions are 64 rows high");
    using HALO = WaveCfg<C, RPW, NW>;
    using WHOLE = WaveCfg<1, RPW, NW>;
    constexpr int W = HALO::RW - 4 * K, TH = HALO::RH - 4 * K;
    // one exchange buffer, seen with the whole region's (narrower) layout too
    __shared__ WaveExchange<T, 2, NW, C> ex;
    const int2 it = items[blockIdx.x];
    const Params p = reps[it.x].p;
    const int64_t off = reps[it.x].off;
    const T *in = bufA + off;
    T *out = bufB + off;
    const T *cin = COMP ? compA + off : nullptr;
    T *cout = COMP ? compB + off : nullptr;
    if(reps[it.x].whole) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, WHOLE, true, false, COMP>(
            in, out, p, 0, 0, *reinterpret_cast<WaveExchange<T, 2, NW, 1> *>(&ex), cin, cout);
    }
    else {
        const int tx = it.y & 0xFFFF, ty = it.y >> 16;
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, HALO, false, false, COMP>(
            in, out, p, tx * W - 2 * K, ty * TH - 2 * K, ex, cin, cout);
    }
}

int batch_columns_per_lane(int shape) { return shape >= 5 ? 3 : (shape >= 3 ? 2 : 1); }
int batch_tile_x(int shape, int K) { return 64 * batch_columns_per_lane(shape) - 4 * K; }
int batch_tile_y(int K) { return kBatchRegion - 4 * K; }

// shape s (rows per wave x waves; halo regions of 64, 128 or 192 columns):
//   0 = 16 x 4, 1 = 8 x 8, 2 = 4 x 16 (64 columns);  3 = 4 x 16, 4 = 8 x 8 (128 columns);
//   5 = 4 x 16, 6 = 8 x 8 (192 columns)
template <typename T, int MODEL, int K, bool COMP>
static hipError_t launch_batch_k(int shape, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, const void *ca, void *cb, hipStream_t s) {
    const T *A = (const T *)a, *CA = (const T *)ca;
    T *Bf = (T *)b, *CB = (T *)cb;
    switch(shape) {
    case 0: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 16, 4, COMP>), dim3(nitems), dim3(256), 0, s, A, Bf, reps, items, CA, CB); break;
    case 1: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 8, 8, COMP>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    case 2: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 4, 16, COMP>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 3: hipLaunchKernelGGL((k_batch<T, MODEL, K, 2, 4, 16, COMP>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 4: hipLaunchKernelGGL((k_batch<T, MODEL, K, 2, 8, 8, COMP>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    case 5: hipLaunchKernelGGL((k_batch<T, MODEL, K, 3, 4, 16, COMP>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 6: hipLaunchKernelGGL((k_batch<T, MODEL, K, 3, 8, 8, COMP>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int MODEL, bool COMP>
static hipError_t launch_batch_t(int shape, int K, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, const void *ca, void *cb, hipStream_t s) {
    switch(K) {
    case 1: return launch_batch_k<T, MODEL, 1, COMP>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 2: return launch_batch_k<T, MODEL, 2, COMP>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 4: return launch_batch_k<T, MODEL, 4, COMP>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 8: return launch_batch_k<T, MODEL, 8, COMP>(shape, a, b, reps, items, nitems, ca, cb, s);
    }
    return hipErrorInvalidValue;
}

hipError_t batch_launch(int precision, int model, int shape, int K, const void *a, void *b, const BatchRep *reps,
                        const int2 *items, int nitems, const void *ca, void *cb, hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU)
            return launch_batch_t<double, CH_LANDAU, false>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
        return launch_batch_t<double, CH_SIMPLE_WERTHEIM, false>(shape, K, a, b, reps, items, nitems, nullptr,
                                                                 nullptr, s);
    }
    // FP32 fields: compensated updates when the batch carries the low-order parts
    if(ca) {
        if(model == CH_LANDAU) return launch_batch_t<float, CH_LANDAU, true>(shape, K, a, b, reps, items, nitems, ca, cb, s);
        return launch_batch_t<float, CH_SIMPLE_WERTHEIM, true>(shape, K, a, b, reps, items, nitems, ca, cb, s);
    }
    if(model == CH_LANDAU)
        return launch_batch_t<float, CH_LANDAU, false>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
    return launch_batch_t<float, CH_SIMPLE_WERTHEIM, false>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
}

}  // namespace chd
