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
// UPD: how the FP32 updates are rounded (wave_region); UPD_STORED and
// UPD_LOW8 read the low-order parts from compA and write them to compB (T or
// int8_t per element)
template <typename T, int MODEL, int K, int C, int RPW, int NW, int UPD = UPD_PLAIN>
__global__ __launch_bounds__(64 * NW) void k_batch(const T *__restrict__ bufA, T *__restrict__ bufB,
                                                  const BatchRep *__restrict__ reps, const int2 *__restrict__ items,
                                                  const void *__restrict__ compA, void *__restrict__ compB) {
    static_assert(RPW * NW >= kBatchRegion && kBatchRegion % NW == 0, "whole items split 64 rows over the waves");
    using HALO = WaveCfg<C, RPW, NW>;
    using WHOLE = WaveCfg<1, kBatchRegion / NW, NW>;
    constexpr int W = HALO::RW - 4 * K, TH = HALO::RH - 4 * K;
    // one exchange buffer, seen with the whole region's (narrower) layout too
    __shared__ WaveExchange<T, 2, NW, C> ex;
    const int2 it = items[blockIdx.x];
    const Params p = reps[it.x].p;
    const int64_t off = reps[it.x].off;
    const T *in = bufA + off;
    T *out = bufB + off;
    constexpr size_t celt = UPD == UPD_LOW8 ? 1 : sizeof(T);
    constexpr bool HASC = UPD == UPD_STORED || UPD == UPD_LOW8;
    const void *cin = HASC ? (const char *)compA + off * celt : nullptr;
    void *cout = HASC ? (char *)compB + off * celt : nullptr;
    if(reps[it.x].whole) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, WHOLE, true, false, UPD>(
            in, out, p, 0, 0, *reinterpret_cast<WaveExchange<T, 2, NW, 1> *>(&ex), cin, cout);
    }
    else {
        const int tx = it.y & 0xFFFF, ty = it.y >> 16;
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, HALO, false, false, UPD>(
            in, out, p, tx * W - 2 * K, ty * TH - 2 * K, ex, cin, cout);
    }
}

int batch_family(int shape) { return shape == 7 ? 3 : (shape >= 5 ? 2 : (shape >= 3 ? 1 : 0)); }
int batch_family_shape(int family) { return family == 3 ? 7 : (family == 2 ? 5 : (family == 1 ? 3 : 0)); }
static int batch_columns_per_lane(int shape) { return shape >= 5 ? 3 : (shape >= 3 ? 2 : 1); }
int batch_tile_x(int shape, int K) { return 64 * batch_columns_per_lane(shape) - 4 * K; }
int batch_tile_y(int shape, int K) { return (shape == 7 ? 80 : kBatchRegion) - 4 * K; }

// shape s (rows per wave x waves; halo regions of 64, 128 or 192 columns):
//   0 = 16 x 4, 1 = 8 x 8, 2 = 4 x 16 (64 columns);  3 = 4 x 16, 4 = 8 x 8 (128 columns);
//   5 = 4 x 16, 6 = 8 x 8 (192 columns);  7 = 5 x 16 (192 columns, 80 rows: at K = 4
//   the 64-row interiors tile every power-of-two N >= 64 exactly)
template <typename T, int MODEL, int K, int UPD>
static hipError_t launch_batch_k(int shape, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, const void *ca, void *cb, hipStream_t s) {
    const T *A = (const T *)a;
    T *Bf = (T *)b;
    const void *CA = ca;
    void *CB = cb;
    switch(shape) {
    case 0: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 16, 4, UPD>), dim3(nitems), dim3(256), 0, s, A, Bf, reps, items, CA, CB); break;
    case 1: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 8, 8, UPD>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    case 2: hipLaunchKernelGGL((k_batch<T, MODEL, K, 1, 4, 16, UPD>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 3: hipLaunchKernelGGL((k_batch<T, MODEL, K, 2, 4, 16, UPD>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 4: hipLaunchKernelGGL((k_batch<T, MODEL, K, 2, 8, 8, UPD>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    case 5: hipLaunchKernelGGL((k_batch<T, MODEL, K, 3, 4, 16, UPD>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    case 6: hipLaunchKernelGGL((k_batch<T, MODEL, K, 3, 8, 8, UPD>), dim3(nitems), dim3(512), 0, s, A, Bf, reps, items, CA, CB); break;
    case 7: hipLaunchKernelGGL((k_batch<T, MODEL, K, 3, 5, 16, UPD>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int MODEL, int UPD>
static hipError_t launch_batch_t(int shape, int K, const void *a, void *b, const BatchRep *reps, const int2 *items,
                                 int nitems, const void *ca, void *cb, hipStream_t s) {
    switch(K) {
    case 1: return launch_batch_k<T, MODEL, 1, UPD>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 2: return launch_batch_k<T, MODEL, 2, UPD>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 4: return launch_batch_k<T, MODEL, 4, UPD>(shape, a, b, reps, items, nitems, ca, cb, s);
    case 8: return launch_batch_k<T, MODEL, 8, UPD>(shape, a, b, reps, items, nitems, ca, cb, s);
    }
    return hipErrorInvalidValue;
}

template <typename T, int MODEL>
static hipError_t launch_batch_m(int upd, int shape, int K, const void *a, void *b, const BatchRep *reps,
                                 const int2 *items, int nitems, const void *ca, void *cb, hipStream_t s) {
    switch(upd) {
    case UPD_PLAIN: return launch_batch_t<T, MODEL, UPD_PLAIN>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
    case UPD_LAUNCH: return launch_batch_t<T, MODEL, UPD_LAUNCH>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
    case UPD_STORED: return launch_batch_t<T, MODEL, UPD_STORED>(shape, K, a, b, reps, items, nitems, ca, cb, s);
    case UPD_LOW8: return launch_batch_t<T, MODEL, UPD_LOW8>(shape, K, a, b, reps, items, nitems, ca, cb, s);
    }
    return hipErrorInvalidValue;
}

// upd: the compensation of FP32 updates (CH_COMP_NONE / LAUNCH / STORED, the
// wave_region UPD modes); FP64 batches always update plainly
hipError_t batch_launch(int precision, int model, int upd, int shape, int K, const void *a, void *b,
                        const BatchRep *reps, const int2 *items, int nitems, const void *ca, void *cb, hipStream_t s) {
    if(precision == CH_FP64) {
        if(model == CH_LANDAU)
            return launch_batch_t<double, CH_LANDAU, UPD_PLAIN>(shape, K, a, b, reps, items, nitems, nullptr, nullptr, s);
        return launch_batch_t<double, CH_SIMPLE_WERTHEIM, UPD_PLAIN>(shape, K, a, b, reps, items, nitems, nullptr,
                                                                     nullptr, s);
    }
    if(model == CH_LANDAU) return launch_batch_m<float, CH_LANDAU>(upd, shape, K, a, b, reps, items, nitems, ca, cb, s);
    return launch_batch_m<float, CH_SIMPLE_WERTHEIM>(upd, shape, K, a, b, reps, items, nitems, ca, cb, s);
}

}  // namespace chd
