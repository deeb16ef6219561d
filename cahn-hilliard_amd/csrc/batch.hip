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

#include <vector>

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

// Shape 8 fits each replica size with its own region geometry in the one
// launch (16 waves per block throughout, at most 16 cells per lane):
//   N = 64    the whole replica (64 x 64, one column per lane)
//   N = 128   the whole replica (128 x 128, two columns per lane, 8 rows per wave)
//   N = 256   x-whole regions of all 256 columns (four per lane) and 64 rows,
//             periodic in x with no x halo, a 2K halo in y
//   other N   the 192 x 80 halo regions of shape 7
// so the 128 and 256 replicas, which the fixed-size regions tile with a third
// to a half of their work spent on halos and clipped columns, lose most of it.
// FP32 Landau asks for 8 waves per SIMD (two 16-wave blocks per CU, at most 64
// VGPRs): the compensated forms need 65-68 otherwise and fit one block per CU.
// The other instantiations need far more than 64 and keep their registers.
template <typename T, int MODEL>
constexpr int fit_waves_per_eu() { return (sizeof(T) == 4 && MODEL == CH_LANDAU) ? 8 : 1; }
template <typename T, int MODEL, int K, int UPD>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(fit_waves_per_eu<T, MODEL>()))) void k_batch_fit(const T *__restrict__ bufA, T *__restrict__ bufB,
                                                  const BatchRep *__restrict__ reps, const int2 *__restrict__ items,
                                                  const void *__restrict__ compA, void *__restrict__ compB) {
    using W64 = WaveCfg<1, 4, 16>;
    using W128 = WaveCfg<2, 8, 16>;
    using X256 = WaveCfg<4, 4, 16>;
    using HALO = WaveCfg<3, 5, 16>;
    __shared__ WaveExchange<T, 2, 16, 4> ex;  // seen with the narrower layouts too
    const int2 it = items[blockIdx.x];
    const Params p = reps[it.x].p;
    const int64_t off = reps[it.x].off;
    const T *in = bufA + off;
    T *out = bufB + off;
    constexpr size_t celt = UPD == UPD_LOW8 ? 1 : sizeof(T);
    constexpr bool HASC = UPD == UPD_STORED || UPD == UPD_LOW8;
    const void *cin = HASC ? (const char *)compA + off * celt : nullptr;
    void *cout = HASC ? (char *)compB + off * celt : nullptr;
    const int N = p.N, ty = it.y >> 16;
    if(N == 64) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, W64, true, false, UPD>(
            in, out, p, 0, 0, *reinterpret_cast<WaveExchange<T, 2, 16, 1> *>(&ex), cin, cout);
    }
    else if(N == 128) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, W128, true, false, UPD>(
            in, out, p, 0, 0, *reinterpret_cast<WaveExchange<T, 2, 16, 2> *>(&ex), cin, cout);
    }
    else if(N == 256) {
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, X256, false, false, UPD, true>(
            in, out, p, 0, ty * (X256::RH - 4 * K) - 2 * K, ex, cin, cout);
    }
    else {
        const int tx = it.y & 0xFFFF;
        wave_region<T, MODEL, CH_MOB_CONSTANT, false, K, HALO, false, false, UPD>(
            in, out, p, tx * (HALO::RW - 4 * K) - 2 * K, ty * (HALO::RH - 4 * K) - 2 * K,
            *reinterpret_cast<WaveExchange<T, 2, 16, 3> *>(&ex), cin, cout);
    }
}

int batch_family(int shape) {
    return shape == 8 ? 4 : (shape == 7 ? 3 : (shape >= 5 ? 2 : (shape >= 3 ? 1 : 0)));
}
int batch_family_shape(int family) {
    return family == 4 ? 8 : (family == 3 ? 7 : (family == 2 ? 5 : (family == 1 ? 3 : 0)));
}
static int batch_columns_per_lane(int shape) { return shape >= 5 ? 3 : (shape >= 3 ? 2 : 1); }
int batch_tile_x(int shape, int K) { return 64 * batch_columns_per_lane(shape) - 4 * K; }
int batch_tile_y(int shape, int K) { return (shape >= 7 ? 80 : kBatchRegion) - 4 * K; }

// the items of one replica of size N for a K-step launch of `shape`'s family:
// (tx | ty << 16) per region
std::vector<int> batch_items(int shape, int N, int K) {
    std::vector<int> v;
    if(N == kBatchRegion || (shape == 8 && N == 128)) {
        v.push_back(0);  // the whole replica
        return v;
    }
    if(shape == 8 && N == 256) {
        const int Ty = kBatchRegion - 4 * K;  // x-whole regions: all 256 columns, 64 rows
        for(int ty = 0; ty < (N + Ty - 1) / Ty; ty++) v.push_back(ty << 16);
        return v;
    }
    const int Tx = batch_tile_x(shape, K), Ty = batch_tile_y(shape, K);
    for(int ty = 0; ty < (N + Ty - 1) / Ty; ty++)
        for(int tx = 0; tx < (N + Tx - 1) / Tx; tx++) v.push_back(tx | (ty << 16));
    return v;
}

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
    case 8: hipLaunchKernelGGL((k_batch_fit<T, MODEL, K, UPD>), dim3(nitems), dim3(1024), 0, s, A, Bf, reps, items, CA, CB); break;
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
