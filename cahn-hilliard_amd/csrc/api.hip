// api.hip -- the C ABI (include/ch_hip.h): contexts, systems, step loop.
//
// A ch_sys owns its device buffers (the reference's CUDAIntegrator owns
// _d_rho, _d_rho_der and _d_mobility, CUDAIntegrator.cu:13-39).  The step
// loop replays a captured hipGraph of G steps (ping-pong buffers, G even) so
// the host issues one launch per G steps; the remainder runs eagerly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <mutex>

#include "ch_internal.h"

// default temporal-blocking depth (time steps per stencil launch); the 2D wave
// kernel's region shape defaults to chd::wave_default_cfg(N)
static constexpr int kDefaultTB = 8;

struct ch_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
};

struct ch_sys {
    ch_ctx *ctx = nullptr;
    ch_desc d{};
    chd::Params p{};
    int64_t bins = 0, hat = 0;
    size_t elt = 8;
    void *psi[2] = {nullptr, nullptr};
    int cur = 0;
    void *work = nullptr, *rho_hat = nullptr, *twiddles = nullptr;
    void *stage = nullptr;   // bins doubles: host transfers of FP32 fields, mobility
    void *stage2 = nullptr;  // bins elements of T: mobility
    void *split_mob = nullptr;  // bins elements of T: the cells' mobilities of the two-launch 3D step
    // the 3D euler_mobility reverse step of one cell per thread: nu
    void *adj_split[1] = {nullptr};
    double *partials = nullptr, *result = nullptr;
    // reverse mode: checkpoints, the window tape, lambda ping-pong, gradient slots
    void *ckpt = nullptr, *tape = nullptr, *lam[2] = {nullptr, nullptr};
    double *gacc = nullptr;
    int64_t ckpt_cap = 0, tape_cap = 0;  // capacities in fields
    int64_t rec_steps = 0, rec_window = 0;
    bool rec_valid = false, rec_full_tape = false;
    hipStream_t aux = nullptr;  // window recompute, overlapped with the adjoint
    hipEvent_t ev_start = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr};
    // pseudospectral reverse mode: checkpointed spectra, the recompute's spectral
    // state and work spectrum, the reverse step's three spectra and two fields
    void *ckpt_hat = nullptr, *ps_rh = nullptr, *ps_w = nullptr, *ps_spec[3] = {nullptr, nullptr, nullptr};
    void *ps_real[2] = {nullptr, nullptr};
    int64_t ckpt_hat_cap = 0;
    int graph_steps = 0;
    int tb = 1;  // time steps per stencil launch (temporal blocking depth)
    // region shape of the 2D wave kernel's shorter launches (K < tb), which
    // close a step count that tb does not divide: smaller regions give such a
    // launch enough blocks to fill the chip
    int tail_cfg = 0;
    int64_t *noise_step = nullptr;  // euler_mobility noise: index of the next step (device)
    // pseudospectral_mobility: spectra, fields, reductions and the GMRES workspace
    void *psm_spec[4] = {nullptr, nullptr, nullptr, nullptr};
    void *psm_field[3] = {nullptr, nullptr, nullptr};
    void *psm_vec[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // V, w, b, x, rhs, r, av
    double *psm_dev = nullptr;   // kPsmParts partials, 4 scalars, Hcol, y
    double *psm_host = nullptr;  // pinned, kPsmParts doubles
    // pinned: mass and free energy written by the finalize kernel, the pair of
    // the synchronous call first, then the CH_OBS_QUEUE slots of the queue
    double *obs_host = nullptr;
    hipEvent_t obs_ev[CH_OBS_QUEUE] = {};
    int obs_head = 0, obs_count = 0;
    int gmres_iters = 0;
    bool psm_fuse = true;  // CH_PSM_FUSE=0: the element-wise kernels around whole transforms
    // multi-species models: host and device model, X (Generic), mu, mobility,
    // per-species spectra and two scratch spectra
    int S = 1;
    chd::SpModel *spm_host = nullptr, *spm_dev = nullptr;
    double *sp_X = nullptr;
    void *sp_mu = nullptr, *sp_mob = nullptr, *sp_rho_hat = nullptr, *sp_spec[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // stencil integrators: the launches of a count of at most 4 G steps, or of
    // the remainder nsteps % G of a longer run (per starting buffer), captured
    // when the same count recurs
    hipGraphExec_t rexec[2] = {nullptr, nullptr};
    // ch_sys_evolve_observe: per starting buffer, the captured steps + queued
    // reductions of a recurring count (osteps), its end buffer (ocur), and the
    // count seen once (captured on its second occurrence)
    hipGraphExec_t oexec[2] = {nullptr, nullptr};
    int64_t osteps[2] = {0, 0}, oseen[2] = {-1, -1};
    int ocur[2] = {0, 0};
    int64_t rsteps[2] = {0, 0};
    int rcur[2] = {0, 0};       // the buffer holding the field after those steps
    int64_t rseen[2] = {0, 0};  // the last such count run by direct launches
};

static ch_status fail(ch_ctx *ctx, ch_status st, const char *fmt, ...) {
    if(ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return st;
}

#define HIPCHK(ctx, call)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if(e_ != hipSuccess) return fail((ctx), CH_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while(0)

extern "C" int ch_abi_version(void) { return CH_ABI_VERSION; }

extern "C" const char *ch_status_string(ch_status s) {
    switch(s) {
    case CH_OK: return "ok";
    case CH_EINVAL: return "invalid argument";
    case CH_EHIP: return "HIP runtime error";
    case CH_ENAN: return "NaN encountered";
    case CH_EUNSUPPORTED: return "unsupported";
    case CH_ENOMEM: return "out of device memory";
    case CH_ESTATE: return "invalid call order";
    }
    return "unknown";
}

extern "C" ch_status ch_ctx_create(int device, ch_ctx **out) {
    if(!out) return CH_EINVAL;
    *out = nullptr;
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n == 0) return CH_EHIP;
    if(device < 0 || device >= n) return CH_EINVAL;
    ch_ctx *c = new ch_ctx();
    c->device = device;
    // the scratch address of ch_observables_device_noalloc, once per device
    if(hipSetDevice(device) != hipSuccess || chd::noalloc_scratch_init(device) != hipSuccess ||
       hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CH_EHIP;
    }
    *out = c;
    return CH_OK;
}

extern "C" void ch_ctx_destroy(ch_ctx *ctx) {
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *ch_ctx_last_error(const ch_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" ch_status ch_ctx_synchronize(ch_ctx *ctx) {
    if(!ctx) return CH_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CH_OK;
}

extern "C" void *ch_ctx_stream(ch_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ------------------------------------------------------------------ helpers
static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static ch_status validate(ch_ctx *ctx, const ch_desc *d) {
    if(!d) return fail(ctx, CH_EINVAL, "null descriptor");
    if(d->dims < 1 || d->dims > 3) return fail(ctx, CH_EINVAL, "dims must be 1, 2 or 3 (got %d)", d->dims);
    if(!is_pow2(d->N)) return fail(ctx, CH_EINVAL, "N should be a power of 2 (got %d)", d->N);
    if(d->N < 4) return fail(ctx, CH_EUNSUPPORTED, "N must be at least 4 (got %d)", d->N);
    if(d->precision != CH_FP64 && d->precision != CH_FP32) return fail(ctx, CH_EINVAL, "bad precision %d", d->precision);
    if(d->model < CH_LANDAU || d->model > CH_GENERIC_WERTHEIM) return fail(ctx, CH_EINVAL, "bad model %d", d->model);
    if(!(d->density_conversion >= 0.0)) return fail(ctx, CH_EINVAL, "density_conversion must not be negative");
    if(d->integrator < CH_EULER || d->integrator > CH_PSEUDOSPECTRAL_MOBILITY)
        return fail(ctx, CH_EINVAL, "bad integrator %d", d->integrator);
    if(d->mobility < CH_MOB_CONSTANT || d->mobility > CH_MOB_LINEAR)
        return fail(ctx, CH_EINVAL, "bad mobility type %d", d->mobility);
    // Integrator::validate (src/integrators/Integrator.cpp:42-46)
    if(d->integrator != CH_EULER_MOBILITY && d->integrator != CH_PSEUDOSPECTRAL_MOBILITY && d->mobility != CH_MOB_CONSTANT)
        return fail(ctx, CH_EINVAL, "The selected integrator only supports constant mobility");
    // FreeEnergyModel::set_mobility default (src/models/FreeEnergyModel.h:51-53)
    if(d->mobility == CH_MOB_FREE_ENERGY && d->model != CH_SIMPLE_WERTHEIM && d->model != CH_GENERIC_WERTHEIM)
        return fail(ctx, CH_EUNSUPPORTED, "this model does not support mobility coupling");
    if(d->mobility == CH_MOB_LINEAR && d->rho_ref == 0.0) return fail(ctx, CH_EINVAL, "linear mobility needs rho_ref != 0");
    if(d->model == CH_SIMPLE_WERTHEIM && (d->valence <= 0 || !(d->delta > 0.0)))
        return fail(ctx, CH_EINVAL, "simple_wertheim needs valence > 0 and delta > 0");
    if(!(d->dt > 0.0) || !(d->dx > 0.0)) return fail(ctx, CH_EINVAL, "dt and dx must be positive");
    if(d->integrator == CH_PSEUDOSPECTRAL_MOBILITY && d->use_gmres &&
       (d->gmres_restart < 0 || d->gmres_max_iter < 0 || !(d->gmres_tol >= 0.0)))
        return fail(ctx, CH_EINVAL, "gmres_restart, gmres_max_iter and gmres_tol must not be negative");
    if((d->integrator == CH_PSEUDOSPECTRAL || d->integrator == CH_PSEUDOSPECTRAL_MOBILITY) && d->N > 4096)
        return fail(ctx, CH_EUNSUPPORTED, "pseudospectral supports N <= 4096 (got %d)", d->N);
    int64_t bins = 1;
    for(int i = 0; i < d->dims; i++) bins *= d->N;
    if(bins > ((int64_t)1 << 31)) return fail(ctx, CH_EUNSUPPORTED, "grid too large (%lld cells)", (long long)bins);
    if(d->integrator == CH_EULER_MOBILITY && d->with_noise && !std::isfinite(d->noise_rescale_factor))
        return fail(ctx, CH_EINVAL, "noise_rescale_factor must be finite");
    return CH_OK;
}

// the observables' kObsParts partial pairs, then (2 doubles' room) the device
// count of queued observables, which picks each queued pair's ring slot
// (chd::observables_ring); zeroed here, with the host's count of queued pairs
static bool alloc_partials(double **p) {
    const size_t bytes = (chd::kObsParts * 2 + 2) * sizeof(double);
    if(hipMalloc((void **)p, bytes) != hipSuccess) {
        *p = nullptr;
        return false;
    }
    return hipMemset(*p, 0, bytes) == hipSuccess;
}

static bool has_noise(const ch_sys *s) { return s->d.integrator == CH_EULER_MOBILITY && s->d.with_noise; }
static bool is_spectral(int integrator) {
    return integrator == CH_PSEUDOSPECTRAL || integrator == CH_PSEUDOSPECTRAL_MOBILITY;
}

static chd::Params make_params(const ch_desc *d) {
    chd::Params p{};
    p.N = d->N;
    p.mask = d->N - 1;
    p.log2N = 0;
    while((1 << p.log2N) < d->N) p.log2N++;
    p.dims = d->dims;
    p.valence = d->valence;
    p.use_dealias = d->use_dealias;
    p.eps = d->epsilon;
    p.B2 = d->B2;
    p.reg = d->regularisation_delta;
    p.log_reg = std::log(d->regularisation_delta);
    p.tvd = 2.0 * d->valence * d->delta;
    p.M = d->M;
    p.rho_min = d->rho_min;
    p.rho_ref = d->rho_ref;
    p.k = d->k;
    p.two_k = 2.0 * d->k;
    p.dx = d->dx;
    p.inv_dx = 1.0 / d->dx;
    p.inv_dx2 = 1.0 / (d->dx * d->dx);
    p.dt = d->dt;
    p.Mdt = d->M * d->dt;
    p.mu_lap = -2.0 * d->k / (d->dx * d->dx);
    p.upd_lap = d->M * d->dt / (d->dx * d->dx);
    {
        const double q = 2.0 * 3.14159265358979323846 / (d->N * d->dx);
        p.q2_unit = q * q;
        p.k_unit = q;
    p.dcf = d->density_conversion > 0.0 ? d->density_conversion : 1.0;
    }
    p.wave_cfg = chd::wave_default_cfg(d->N);
    p.adj_cfg = chd::adj_default_cfg(d->N);
    // measured on C5 (SimpleWertheim 3D N=128 euler_mobility): 16x8x8 with 512
    // threads and the 16x8x4 adjoint tile beat the other shapes
    p.tile3d = 1;
    // SimpleWertheim 3D: the two-launch step evaluates each cell's FP64 f' once
    // (C5, N=128 linear mobility: 34.7 against 41.1 us per step with the tile);
    // Landau's cheap f' keeps the fused tile ahead (18.3 against 21.7 us)
    if(d->dims == 3 && d->model == CH_SIMPLE_WERTHEIM && chd::split3d_ok(p, d->integrator, d->mobility)) p.tile3d = 3;
    p.adj_tile3d = 1;
    // noise step: 64 x 32 (2D) and 16 x 8 x 8 (3D) tiles of 512 threads measured
    // 10 % and 30 % faster than 64 x 16 and 8 x 8 x 8 with 256
    p.noise_tile = 2;
    p.S = d->S;
    double bins = std::pow((double)d->N, d->dims);
    p.vbin_over_bins = d->dx * d->dx * d->dx / bins;
    p.inv_bins = 1.0 / bins;
    // EulerMobilityCPU.cpp:14
    p.noise = (d->integrator == CH_EULER_MOBILITY && d->with_noise)
                  ? std::sqrt(2.0 / (std::pow(d->dx, d->dims) * d->dt)) * d->noise_rescale_factor
                  : 0.0;
    p.seed = d->seed;
    return p;
}

// e^{-2 pi i m / N} from the integer index: octant reduction, long double
// sin/cos on [0, pi/4]; exact 0 and +-1 at the axes.
static void twiddle(int m, int N, long double &c, long double &s) {
    const long double PI = 3.141592653589793238462643383279502884L;
    m = ((m % N) + N) % N;
    int sign = 1;
    if(m >= N / 2) { m -= N / 2; sign = -1; }
    long double cc, ss;
    if(4 * m <= N / 2) {  // [0, pi/4]
        long double a = 2 * PI * m / N;
        cc = cosl(a); ss = sinl(a);
    }
    else if(4 * m < N) {  // (pi/4, pi/2)
        long double a = 2 * PI * (N / 4.0L - m) / N;
        cc = sinl(a); ss = cosl(a);
    }
    else if(4 * m == N) { cc = 0; ss = 1; }
    else if(4 * m <= 3 * N / 2) {  // (pi/2, 3pi/4]
        long double a = 2 * PI * (m - N / 4.0L) / N;
        cc = -sinl(a); ss = cosl(a);
    }
    else {  // (3pi/4, pi)
        long double a = 2 * PI * (N / 2.0L - m) / N;
        cc = -cosl(a); ss = sinl(a);
    }
    c = sign * cc;
    s = -sign * ss;  // forward sign
}

// W_N^m, m = 0 .. N-1, in the system's precision
static bool upload_twiddles(ch_sys *s) {
    const int N = s->d.N;
    if(hipMalloc(&s->twiddles, (size_t)N * 2 * s->elt) != hipSuccess) return false;
    std::vector<double> twd(2 * N);
    std::vector<float> twf(2 * N);
    for(int m = 0; m < N; m++) {
        long double c, sn;
        twiddle(m, N, c, sn);
        twd[2 * m] = (double)c;
        twd[2 * m + 1] = (double)sn;
        twf[2 * m] = (float)twd[2 * m];
        twf[2 * m + 1] = (float)twd[2 * m + 1];
    }
    const void *src = s->d.precision == CH_FP64 ? (const void *)twd.data() : (const void *)twf.data();
    return hipMemcpy(s->twiddles, src, (size_t)N * 2 * s->elt, hipMemcpyHostToDevice) == hipSuccess;
}

static void free_sys(ch_sys *s) {
    if(s->aux) {
        (void)hipStreamSynchronize(s->aux);
        (void)hipStreamDestroy(s->aux);
    }
    for(hipEvent_t e : {s->ev_start, s->ev_rec[0], s->ev_rec[1], s->ev_adj[0], s->ev_adj[1]})
        if(e) (void)hipEventDestroy(e);
    for(int i = 0; i < 2; i++) {
        if(s->gexec[i]) (void)hipGraphExecDestroy(s->gexec[i]);
        if(s->rexec[i]) (void)hipGraphExecDestroy(s->rexec[i]);
        if(s->oexec[i]) (void)hipGraphExecDestroy(s->oexec[i]);
        if(s->psi[i]) (void)hipFree(s->psi[i]);
    }
    void *bufs[] = {s->noise_step, s->work, s->rho_hat, s->twiddles, s->stage, s->stage2, s->split_mob, s->partials, s->result,
                    s->ckpt, s->tape, s->lam[0], s->lam[1], s->gacc, s->ckpt_hat, s->ps_rh, s->ps_w,
                    s->ps_spec[0], s->ps_spec[1], s->ps_spec[2], s->ps_real[0], s->ps_real[1],
                    s->adj_split[0]};
    for(void *b : bufs)
        if(b) (void)hipFree(b);
    for(void *b : s->psm_spec)
        if(b) (void)hipFree(b);
    for(void *b : s->psm_field)
        if(b) (void)hipFree(b);
    for(void *b : s->psm_vec)
        if(b) (void)hipFree(b);
    if(s->psm_dev) (void)hipFree(s->psm_dev);
    for(void *b : {(void *)s->spm_dev, (void *)s->sp_X, s->sp_mu, s->sp_mob, s->sp_rho_hat, s->sp_spec[0], s->sp_spec[1]})
        if(b) (void)hipFree(b);
    delete s->spm_host;
    if(s->psm_host) (void)hipHostFree(s->psm_host);
    if(s->obs_host) (void)hipHostFree(s->obs_host);
    for(hipEvent_t e : s->obs_ev)
        if(e) (void)hipEventDestroy(e);
    delete s;
}

extern "C" ch_status ch_sys_create(ch_ctx *ctx, const ch_desc *desc, ch_sys **out) {
    if(!ctx || !out) return CH_EINVAL;
    *out = nullptr;
    ch_status st = validate(ctx, desc);
    if(st != CH_OK) return st;
    if(desc->model > CH_SIMPLE_WERTHEIM)
        return fail(ctx, CH_EINVAL, "multi-species model %d: use ch_sys_create_species", desc->model);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ch_sys *s = new ch_sys();
    s->ctx = ctx;
    s->d = *desc;
    s->p = make_params(desc);
    s->bins = 1;
    for(int i = 0; i < desc->dims; i++) s->bins *= desc->N;
    s->elt = desc->precision == CH_FP64 ? 8 : 4;
    const int N = desc->N;
    s->hat = s->bins / N * (N / 2 + 1);
    // temporal blocking depth: kDefaultTB, or CH_TB_STEPS (environment) when
    // set, rounded down to a depth the shape has (1 for the pseudospectral path)
    // (1 with noise: one launch per step)
    const bool noisy = desc->integrator == CH_EULER_MOBILITY && desc->with_noise;
    const int max_k = (is_spectral(desc->integrator) || noisy) ? 1 : chd::stencil_max_k(desc->dims, desc->N);
    int want = kDefaultTB;
    if(const char *tb = std::getenv("CH_TB_STEPS")) want = std::atoi(tb);
    s->tb = 1;
    while(s->tb * 2 <= max_k && s->tb * 2 <= want) s->tb *= 2;
    if(const char *r = std::getenv("CH_WAVE_CFG")) s->p.wave_cfg = std::max(0, std::min(chd::kWaveCfgs - 1, std::atoi(r)));
    // N >= 1024: 128 x 80 regions for the shorter launches (measured at N = 1024
    // with 8-step launches and a 4-step tail: 0.2635 ms per 100 steps against
    // 0.2674 with the 192 x 64 regions of the full launches)
    s->tail_cfg = N >= 1024 ? 7 : s->p.wave_cfg;
    if(const char *r = std::getenv("CH_WAVE_CFG_TAIL")) s->tail_cfg = std::max(0, std::min(chd::kWaveCfgs - 1, std::atoi(r)));
    if(const char *r = std::getenv("CH_TILE3D")) s->p.tile3d = std::max(0, std::min(3, std::atoi(r)));
    if(const char *r = std::getenv("CH_ADJ_TILE3D")) s->p.adj_tile3d = std::max(0, std::min(2, std::atoi(r)));
    if(const char *r = std::getenv("CH_NOISE_TILE")) s->p.noise_tile = std::max(0, std::min(2, std::atoi(r)));
    if(const char *r = std::getenv("CH_PSM_FUSE")) s->psm_fuse = std::atoi(r) != 0;
    if(const char *r = std::getenv("CH_ADJ_CFG")) {
        const int v = std::atoi(r);
        if(v == 0 || v == 3 || v == 4 || v == 5) s->p.adj_cfg = v;
    }
    // graph chunk: a multiple of 2 tb launches' worth of steps (parity restored)
    const char *g = std::getenv("CH_GRAPH_STEPS");
    s->graph_steps = g ? std::atoi(g) : 64;
    if(s->graph_steps < 0) s->graph_steps = 0;
    s->graph_steps -= s->graph_steps % (2 * s->tb);

    auto alloc = [&](void **p, size_t bytes) -> bool { return hipMalloc(p, bytes) == hipSuccess; };
    bool ok = alloc(&s->psi[0], s->bins * s->elt) && alloc(&s->stage, s->bins * 8) &&
              alloc(&s->stage2, s->bins * s->elt) && alloc_partials(&s->partials) &&
              alloc((void **)&s->result, 2 * sizeof(double));
    if(ok && !is_spectral(desc->integrator)) ok = alloc(&s->psi[1], s->bins * s->elt);
    if(ok && s->p.tile3d == 3 && chd::split3d_ok(s->p, desc->integrator, desc->mobility) &&
       chd::split3d_needs_mobility(desc->integrator, desc->mobility))
        ok = alloc(&s->split_mob, s->bins * s->elt);
    if(ok && noisy) ok = alloc((void **)&s->noise_step, sizeof(int64_t)) &&
                         hipMemset(s->noise_step, 0, sizeof(int64_t)) == hipSuccess;
    if(ok && desc->integrator == CH_PSEUDOSPECTRAL_MOBILITY) {
        const int m = std::max(2, desc->gmres_restart > 0 ? desc->gmres_restart : 30);
        for(void *&b : s->psm_spec) ok = ok && alloc(&b, s->hat * 2 * s->elt);
        for(void *&b : s->psm_field) ok = ok && alloc(&b, s->bins * s->elt);
        if(desc->use_gmres) {
            ok = ok && alloc(&s->psm_vec[0], (size_t)(m + 1) * s->bins * s->elt);
            for(int i = 1; i < 7; i++) ok = ok && alloc(&s->psm_vec[i], s->bins * s->elt);
        }
        ok = ok && alloc((void **)&s->psm_dev, sizeof(double) * (chd::kPsmParts + 4 + 2 * (m + 2))) &&
             hipHostMalloc((void **)&s->psm_host, sizeof(double) * std::max(chd::kPsmParts, m + 2)) == hipSuccess;
    }
    if(ok && is_spectral(desc->integrator))
        ok = alloc(&s->work, s->hat * 2 * s->elt) && alloc(&s->rho_hat, s->hat * 2 * s->elt) && upload_twiddles(s);
    if(!ok) {
        free_sys(s);
        return fail(ctx, CH_ENOMEM, "device allocation failed for N=%d dims=%d", desc->N, desc->dims);
    }
    *out = s;
    return CH_OK;
}

extern "C" void ch_sys_destroy(ch_sys *s) {
    if(!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_sys(s);
}

// ---------------------------------------------------------------- species
static chd::SpectralArgs spectral_args(ch_sys *s);
static bool is_species(const ch_sys *s) { return s->spm_host != nullptr; }

static chd::SpArgs sp_args(ch_sys *s) {
    chd::SpArgs a{};
    a.p = s->p;
    a.precision = s->d.precision;
    a.mob_rule = s->d.mobility;
    a.m = s->spm_dev;
    a.m_host = s->spm_host;
    a.rho = s->psi[0];
    a.X = s->sp_X;
    a.mu = s->sp_mu;
    a.mob = s->sp_mob;
    a.rho_hat = s->sp_rho_hat;
    a.spec[0] = s->sp_spec[0];
    a.spec[1] = s->sp_spec[1];
    a.bins = s->bins;
    a.hat = s->hat;
    return a;
}

static ch_status validate_species(ch_ctx *ctx, const ch_desc *d, const ch_species_model *m) {
    if(m->model < CH_SALEH_WERTHEIM || m->model > CH_GENERIC_WERTHEIM)
        return fail(ctx, CH_EINVAL, "bad multi-species model %d", m->model);
    if(d->model != m->model) return fail(ctx, CH_EINVAL, "desc->model (%d) differs from model->model (%d)", d->model, m->model);
    const int want = m->model == CH_SALEH_WERTHEIM ? 3 : (m->model == CH_RICCI_WERTHEIM ? 2 : 0);
    if(m->n_species < 1 || m->n_species > CH_MAX_SPECIES || (want && m->n_species != want))
        return fail(ctx, CH_EINVAL, "bad number of species %d for model %d", m->n_species, m->model);
    if(d->integrator == CH_PSEUDOSPECTRAL_MOBILITY)
        return fail(ctx, CH_EUNSUPPORTED, "pseudospectral_mobility supports single-species models only");
    if(d->integrator == CH_EULER_MOBILITY && d->with_noise)
        return fail(ctx, CH_EUNSUPPORTED, "noise supports single-species models only");
    if(m->model == CH_SALEH_WERTHEIM && (m->saleh_valence[0] < 0 || m->saleh_valence[1] < 0 || m->saleh_valence[2] < 0))
        return fail(ctx, CH_EINVAL, "saleh.valence must not be negative");
    if(m->model == CH_GENERIC_WERTHEIM) {
        if(m->n_patches < 1 || m->n_patches > CH_MAX_PATCHES)
            return fail(ctx, CH_EINVAL, "generic_wertheim needs 1 .. %d patch types (got %d)", CH_MAX_PATCHES, m->n_patches);
        for(int a = 0; a < m->n_species; a++)
            for(int q = 0; q < CH_MAX_PATCHES; q++)
                if(m->patch_count[a][q] < 0 || (q >= m->n_patches && m->patch_count[a][q] != 0))
                    return fail(ctx, CH_EINVAL, "bad patch count for species %d, patch %d", a, q);
        for(int a = 0; a < m->n_patches; a++)
            for(int b = 0; b < m->n_patches; b++)
                if(!(m->delta[a][b] >= 0.0) || m->delta[a][b] != m->delta[b][a])
                    return fail(ctx, CH_EINVAL, "generic_wertheim deltas must be symmetric and not negative");
    }
    return CH_OK;
}

extern "C" ch_status ch_sys_create_species(ch_ctx *ctx, const ch_desc *desc, const ch_species_model *model,
                                           ch_sys **out) {
    if(!ctx || !out || !model) return CH_EINVAL;
    *out = nullptr;
    ch_status st = validate(ctx, desc);
    if(st == CH_OK) st = validate_species(ctx, desc, model);
    if(st != CH_OK) return st;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ch_sys *s = new ch_sys();
    s->ctx = ctx;
    s->d = *desc;
    s->p = make_params(desc);
    s->bins = 1;
    for(int i = 0; i < desc->dims; i++) s->bins *= desc->N;
    s->elt = desc->precision == CH_FP64 ? 8 : 4;
    s->hat = s->bins / desc->N * (desc->N / 2 + 1);
    s->tb = 1;
    const char *g = std::getenv("CH_GRAPH_STEPS");
    s->graph_steps = std::max(0, g ? std::atoi(g) : 64);
    s->spm_host = new chd::SpModel;
    chd::sp_build_model(*model, *s->spm_host);
    s->S = model->n_species;
    const size_t SB = (size_t)s->S * s->bins;
    const int P = std::max(1, s->spm_host->P);
    auto alloc = [&](void **p, size_t bytes) -> bool { return hipMalloc(p, bytes) == hipSuccess; };
    bool ok = alloc(&s->psi[0], SB * s->elt) && alloc(&s->stage, SB * 8) &&
              alloc_partials(&s->partials) &&
              alloc((void **)&s->result, 2 * sizeof(double)) && alloc((void **)&s->spm_dev, sizeof(chd::SpModel)) &&
              alloc((void **)&s->sp_X, (size_t)P * s->bins * sizeof(double)) && alloc(&s->sp_mu, SB * s->elt) &&
              alloc(&s->sp_mob, SB * s->elt);
    ok = ok && hipMemcpy(s->spm_dev, s->spm_host, sizeof(chd::SpModel), hipMemcpyHostToDevice) == hipSuccess;
    if(ok && desc->integrator == CH_PSEUDOSPECTRAL) {
        ok = alloc(&s->work, s->hat * 2 * s->elt) && alloc(&s->sp_rho_hat, (size_t)s->S * s->hat * 2 * s->elt) &&
             alloc(&s->sp_spec[0], s->hat * 2 * s->elt) && alloc(&s->sp_spec[1], s->hat * 2 * s->elt) &&
             upload_twiddles(s);
    }
    if(!ok) {
        free_sys(s);
        return fail(ctx, CH_ENOMEM, "device allocation failed for N=%d dims=%d species=%d", desc->N, desc->dims,
                    model->n_species);
    }
    *out = s;
    return CH_OK;
}

extern "C" int32_t ch_sys_species(const ch_sys *s) { return s ? s->S : 0; }

// X = 1 (GenericWertheim.cpp:252-254), the set-up evaluation of the model
// (CahnHilliard.cpp:226-228), rho_hat of every species for the spectral path
static ch_status sp_after_upload(ch_sys *s) {
    ch_ctx *c = s->ctx;
    const int P = s->spm_host->P;
    if(P > 0) {
        std::vector<double> ones((size_t)P * s->bins, 1.0);
        HIPCHK(c, hipMemcpyAsync(s->sp_X, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    const chd::SpArgs a = sp_args(s);
    HIPCHK(c, chd::sp_warm(a, c->stream));
    if(s->d.integrator == CH_PSEUDOSPECTRAL) HIPCHK(c, chd::sp_init_spectral(a, spectral_args(s), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

static ch_status sp_evolve(ch_sys *s, int64_t nsteps) {
    ch_ctx *c = s->ctx;
    const chd::SpArgs a = sp_args(s);
    const chd::SpectralArgs sa = spectral_args(s);
    auto step = [&]() -> hipError_t {
        if(s->d.integrator == CH_PSEUDOSPECTRAL) return chd::sp_step_spectral(a, sa, c->stream);
        return chd::sp_step_stencil(a, s->d.integrator, c->stream);
    };
    const int G = s->graph_steps;
    if(G > 0 && nsteps >= G) {
        if(!s->gexec[0]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for(int i = 0; i < G && e == hipSuccess; i++) e = step();
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&s->gexec[0], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(s->gexec[0], c->stream));
        nsteps -= reps * G;
    }
    for(int64_t i = 0; i < nsteps; i++) HIPCHK(c, step());
    return CH_OK;
}

static chd::SpectralArgs spectral_args(ch_sys *s) {
    chd::SpectralArgs a;
    a.p = s->p;
    a.precision = s->d.precision;
    a.model = s->d.model;
    a.rho = s->psi[0];
    a.work = s->work;
    a.rho_hat = s->rho_hat;
    a.twiddles = s->twiddles;
    a.rh_lines = s->d.integrator == CH_PSEUDOSPECTRAL && !is_species(s) && s->d.dims == 2 &&
                 (s->d.N == 1024 || s->d.N == 2048);
    return a;
}

static ch_status after_upload(ch_sys *s) {
    ch_ctx *c = s->ctx;
    s->cur = 0;
    if(is_spectral(s->d.integrator)) {
        chd::SpectralArgs a = spectral_args(s);
        HIPCHK(c, chd::ps_run(a, c->stream, true));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_sys_upload(ch_sys *s, const double *rho_host) {
    if(!s || !rho_host) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t n = s->S * s->bins;
    if(s->d.precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(s->psi[0], rho_host, n * 8, hipMemcpyHostToDevice, c->stream));
    }
    else {
        HIPCHK(c, hipMemcpyAsync(s->stage, rho_host, n * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, chd::convert_d2f((const double *)s->stage, (float *)s->psi[0], n, c->stream));
    }
    return is_species(s) ? sp_after_upload(s) : after_upload(s);
}

extern "C" ch_status ch_sys_upload_device(ch_sys *s, const void *rho_dev) {
    if(!s || !rho_dev) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(s->psi[0], rho_dev, s->S * s->bins * s->elt, hipMemcpyDeviceToDevice, c->stream));
    return is_species(s) ? sp_after_upload(s) : after_upload(s);
}

extern "C" ch_status ch_sys_download(ch_sys *s, double *rho_host) {
    if(!s || !rho_host) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = s->S * s->bins;
    if(s->d.precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(rho_host, s->psi[s->cur], n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    else {
        HIPCHK(c, chd::convert_f2d((const float *)s->psi[s->cur], (double *)s->stage, n, c->stream));
        HIPCHK(c, hipMemcpyAsync(rho_host, s->stage, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_sys_download_device(ch_sys *s, void *rho_dev) {
    if(!s || !rho_dev) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(rho_dev, s->psi[s->cur], s->S * s->bins * s->elt, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_sys_download_mobility(ch_sys *s, double *mob_host) {
    if(!s || !mob_host) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if(is_species(s)) {
        HIPCHK(c, chd::sp_mobility_field(sp_args(s), (double *)s->stage, c->stream));
        HIPCHK(c, hipMemcpyAsync(mob_host, s->stage, s->S * s->bins * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return CH_OK;
    }
    HIPCHK(c, chd::pointwise_mobility(s->p, s->d.precision, s->d.mobility, s->psi[s->cur], s->stage2, s->bins, c->stream));
    if(s->d.precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(mob_host, s->stage2, s->bins * 8, hipMemcpyDeviceToHost, c->stream));
    }
    else {
        HIPCHK(c, chd::convert_f2d((const float *)s->stage2, (double *)s->stage, s->bins, c->stream));
        HIPCHK(c, hipMemcpyAsync(mob_host, s->stage, s->bins * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

// K time steps (one launch) queued on the stream
// (with noise: K = 1, step index *noise_step + noise_off)
static hipError_t launch_steps(ch_sys *s, int &cur, int K, int noise_off = 0) {
    hipStream_t st = s->ctx->stream;
    if(s->d.integrator == CH_PSEUDOSPECTRAL) {
        chd::SpectralArgs a = spectral_args(s);
        return chd::ps_steps(a, st, K);
    }
    const void *in = s->psi[cur];
    void *out = s->psi[cur ^ 1];
    const ch_desc &d = s->d;
    hipError_t e;
    if(has_noise(s)) {
        // K = 1; the caller advances the device step index after the launches
        e = chd::em_noise_step(s->p, d.precision, d.model, d.mobility, in, out, s->noise_step, noise_off, st);
        cur ^= 1;
        return e;
    }
    chd::Params p2 = s->p;
    if(K < s->tb) p2.wave_cfg = s->tail_cfg;
    switch(d.dims) {
    case 1: e = chd::stencil_step<1>(s->p, d.precision, d.integrator, d.model, d.mobility, K, in, out, st); break;
    case 2: e = chd::stencil_step<2>(p2, d.precision, d.integrator, d.model, d.mobility, K, in, out, st); break;
    default:
        if(s->p.tile3d == 3 && chd::split3d_ok(s->p, d.integrator, d.mobility))
            e = chd::stencil_split3d(s->p, d.precision, d.integrator, d.model, d.mobility, in, out, s->stage2,
                                     s->split_mob, st);
        else
            e = chd::stencil_step<3>(s->p, d.precision, d.integrator, d.model, d.mobility, K, in, out, st);
        break;
    }
    cur ^= 1;
    return e;
}

// pseudospectral: chains of steps joined by the fused row pass; a chain of G
// steps is captured once (one buffer, no parity) and replayed
static ch_status evolve_spectral(ch_sys *s, int64_t nsteps) {
    ch_ctx *c = s->ctx;
    chd::SpectralArgs a = spectral_args(s);
    const int G = s->graph_steps;
    // up to 4 G steps, a recurring count is one cached graph of its chain (below)
    if(G > 0 && nsteps > 4 * G) {
        if(!s->gexec[0]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = chd::ps_steps(a, c->stream, G);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&s->gexec[0], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(s->gexec[0], c->stream));
        nsteps -= reps * G;
    }
    if(nsteps <= 0) return CH_OK;
    if(G == 0) {
        HIPCHK(c, chd::ps_steps(a, c->stream, nsteps));
        return CH_OK;
    }
    // a count (or remainder) that recurs: one graph of its chain (chains split
    // anywhere give the same bits), kept until another count recurs
    if(!s->rexec[0] || s->rsteps[0] != nsteps) {
        if(s->rseen[0] != nsteps) {
            s->rseen[0] = nsteps;
            HIPCHK(c, chd::ps_steps(a, c->stream, nsteps));
            return CH_OK;
        }
        if(s->rexec[0]) {
            HIPCHK(c, hipGraphExecDestroy(s->rexec[0]));
            s->rexec[0] = nullptr;
        }
        hipGraph_t graph;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = chd::ps_steps(a, c->stream, nsteps);
        hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
        if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
        e = hipGraphInstantiate(&s->rexec[0], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        s->rsteps[0] = nsteps;
    }
    HIPCHK(c, hipGraphLaunch(s->rexec[0], c->stream));
    return CH_OK;
}

static chd::PsmArgs psm_args(ch_sys *s) {
    chd::PsmArgs B{};
    B.a = spectral_args(s);
    B.mob_rule = s->d.mobility;
    B.fh = s->psm_spec[0];
    B.muh = s->psm_spec[1];
    B.gh = s->psm_spec[2];
    B.divh = s->psm_spec[3];
    B.g = s->psm_field[0];
    B.mob = s->psm_field[1];
    B.tr = s->psm_field[2];
    B.restart = s->d.gmres_restart > 0 ? s->d.gmres_restart : 30;
    B.max_iter = s->d.gmres_max_iter > 0 ? s->d.gmres_max_iter : 200;
    B.tol = s->d.gmres_tol > 0.0 ? s->d.gmres_tol : 1e-10;
    if(s->d.precision == CH_FP32) B.tol = std::max(B.tol, 1e-6);
    const int m = std::max(2, B.restart);
    B.partials = s->psm_dev;
    B.scal = s->psm_dev + chd::kPsmParts;
    B.Hcol = B.scal + 4;
    B.ydev = B.Hcol + (m + 2);
    B.host_partials = s->psm_host;
    B.V = s->psm_vec[0];
    B.w = s->psm_vec[1];
    B.b = s->psm_vec[2];
    B.x = s->psm_vec[3];
    B.rhs = s->psm_vec[4];
    B.r = s->psm_vec[5];
    B.av = s->psm_vec[6];
    B.fuse = s->psm_fuse;
    return B;
}

// pseudospectral_mobility: the splitting scheme captures a chain of G steps
// once (no parity: one real buffer) and replays it; the GMRES scheme runs
// step by step (its restarts are decided on the host)
static ch_status evolve_psm(ch_sys *s, int64_t nsteps) {
    ch_ctx *c = s->ctx;
    const chd::PsmArgs B = psm_args(s);
    if(s->d.use_gmres) {
        for(int64_t i = 0; i < nsteps; i++) {
            int it = 0;
            HIPCHK(c, chd::psm_step_full(B, c->stream, it));
            s->gmres_iters = it;
        }
        return CH_OK;
    }
    const int G = s->graph_steps;
    if(G > 0 && nsteps >= G) {
        if(!s->gexec[0]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for(int i = 0; i < G && e == hipSuccess; i++) e = chd::psm_step_simple(B, c->stream);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&s->gexec[0], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(s->gexec[0], c->stream));
        nsteps -= reps * G;
    }
    for(int64_t i = 0; i < nsteps; i++) HIPCHK(c, chd::psm_step_simple(B, c->stream));
    return CH_OK;
}

extern "C" int32_t ch_sys_gmres_iterations(const ch_sys *s) { return s ? s->gmres_iters : 0; }

// euler_mobility with noise: one launch per step, the step index read from
// the device counter plus the launch's offset in its chain; the counter is
// advanced by one small kernel after each chain (G steps captured once)
static ch_status evolve_noise(ch_sys *s, int64_t nsteps) {
    ch_ctx *c = s->ctx;
    const int G = s->graph_steps;
    if(G > 0 && nsteps >= G) {
        const int par = s->cur;
        if(!s->gexec[par]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int cur = s->cur;
            hipError_t e = hipSuccess;
            for(int i = 0; i < G && e == hipSuccess; i++) e = launch_steps(s, cur, 1, i);
            if(e == hipSuccess) e = chd::noise_step_advance(s->noise_step, G, c->stream);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&s->gexec[par], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(s->gexec[par], c->stream));
        nsteps -= reps * G;  // G is even: parity unchanged
    }
    if(nsteps > 0) {
        for(int64_t i = 0; i < nsteps; i++) HIPCHK(c, launch_steps(s, s->cur, 1, (int)i));
        HIPCHK(c, chd::noise_step_advance(s->noise_step, (int)nsteps, c->stream));
    }
    return CH_OK;
}

extern "C" ch_status ch_sys_evolve(ch_sys *s, int64_t nsteps) {
    if(!s || nsteps < 0) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if(is_species(s)) return sp_evolve(s, nsteps);
    if(s->d.integrator == CH_PSEUDOSPECTRAL) return evolve_spectral(s, nsteps);
    if(s->d.integrator == CH_PSEUDOSPECTRAL_MOBILITY) return evolve_psm(s, nsteps);
    if(has_noise(s)) return evolve_noise(s, nsteps);
    const int G = s->graph_steps;
    // up to 4 G steps, a recurring count is one cached graph of all its launches
    // (below); longer runs replay the G-step graph, then the remainder
    if(G > 0 && nsteps > 4 * G) {
        const int par = s->cur;
        if(!s->gexec[par]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int cur = s->cur;
            hipError_t e = hipSuccess;
            for(int i = 0; i < G && e == hipSuccess; i += s->tb) e = launch_steps(s, cur, s->tb);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&s->gexec[par], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(s->gexec[par], c->stream));
        nsteps -= reps * G;  // G / tb launches, an even number: parity unchanged
    }
    // the rest in launches of tb steps, then of decreasing depth: at most log2(tb)
    auto rest = [&](int &cur, int64_t n) -> hipError_t {
        hipError_t e = hipSuccess;
        for(; n >= s->tb && e == hipSuccess; n -= s->tb) e = launch_steps(s, cur, s->tb);
        for(int K = s->tb / 2; K >= 1 && e == hipSuccess; K /= 2)
            if(n >= K) {
                e = launch_steps(s, cur, K);
                n -= K;
            }
        return e;
    };
    if(G == 0 || nsteps < s->tb) {
        HIPCHK(c, rest(s->cur, nsteps));
        return CH_OK;
    }
    // a count (or remainder) of at least one full launch that recurs from the
    // same buffer: one graph, kept until another count recurs
    const int par = s->cur;
    if(!s->rexec[par] || s->rsteps[par] != nsteps) {
        if(s->rseen[par] != nsteps) {
            s->rseen[par] = nsteps;
            HIPCHK(c, rest(s->cur, nsteps));
            return CH_OK;
        }
        if(s->rexec[par]) {
            HIPCHK(c, hipGraphExecDestroy(s->rexec[par]));
            s->rexec[par] = nullptr;
        }
        hipGraph_t graph;
        int cur = par;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = rest(cur, nsteps);
        hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "step launch during capture failed: %s", hipGetErrorString(e));
        if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
        e = hipGraphInstantiate(&s->rexec[par], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        s->rsteps[par] = nsteps;
        s->rcur[par] = cur;
    }
    HIPCHK(c, hipGraphLaunch(s->rexec[par], c->stream));
    s->cur = s->rcur[par];
    return CH_OK;
}

extern "C" int32_t ch_sys_steps_per_launch(const ch_sys *s) { return s ? s->tb : 0; }

extern "C" int32_t ch_sys_region_shape(const ch_sys *s) {
    if(!s || s->d.dims != 2 || s->d.N < 64 || is_spectral(s->d.integrator) || is_species(s) || has_noise(s)) return -1;
    return s->p.wave_cfg;
}

static hipError_t ensure_obs_host(ch_sys *s) {
    if(s->obs_host) return hipSuccess;
    return hipHostMalloc((void **)&s->obs_host, 2 * (1 + CH_OBS_QUEUE) * sizeof(double));
}

// the FP64 sums of the current field into pinned host memory, straight from
// the finalize kernel (no copy behind the reduction): the synchronous pair
// (queued = false) or the next queue slot, which the device's count of queued
// pairs selects (so the same launches can be captured once and replayed)
static hipError_t enqueue_observables(ch_sys *s, bool queued) {
    ch_ctx *c = s->ctx;
    hipError_t e = ensure_obs_host(s);
    if(e != hipSuccess) return e;
    double *ring = s->obs_host + 2;
    unsigned *count = reinterpret_cast<unsigned *>(s->partials + 2 * chd::kObsParts);
    if(is_species(s))
        return queued ? chd::sp_observables(sp_args(s), s->partials, chd::kObsParts, nullptr, c->stream, ring, count)
                      : chd::sp_observables(sp_args(s), s->partials, chd::kObsParts, s->obs_host, c->stream);
    if(queued)
        return chd::observables_ring(s->p, s->d.precision, s->d.model, s->psi[s->cur], s->partials, chd::kObsParts,
                                     ring, count, c->stream);
    return chd::observables(s->p, s->d.precision, s->d.model, s->psi[s->cur], s->partials, chd::kObsParts,
                            s->obs_host, c->stream);
}
static ch_status observables_result(ch_sys *s, const double *pair, double *mass, double *free_energy);

extern "C" ch_status ch_sys_observables(ch_sys *s, double *mass, double *free_energy) {
    if(!s) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, enqueue_observables(s, false));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return observables_result(s, s->obs_host, mass, free_energy);
}

extern "C" ch_status ch_sys_observables_queue(ch_sys *s) {
    if(!s) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(s->obs_count == CH_OBS_QUEUE)
        return fail(c, CH_EINVAL, "%d observables already queued: fetch before queueing more", CH_OBS_QUEUE);
    HIPCHK(c, hipSetDevice(c->device));
    const int slot = (s->obs_head + s->obs_count) % CH_OBS_QUEUE;
    if(!s->obs_ev[slot]) HIPCHK(c, hipEventCreateWithFlags(&s->obs_ev[slot], hipEventDisableTiming));
    HIPCHK(c, enqueue_observables(s, true));
    HIPCHK(c, hipEventRecord(s->obs_ev[slot], c->stream));
    s->obs_count++;
    return CH_OK;
}

// nsteps of ch_sys_evolve, then ch_sys_observables_queue.  For the stencil
// integrators (no noise) a count that recurs from the same buffer is one
// captured graph of its launches and of the reductions: the host issues one
// graph launch per print interval, with no launch gap before the reductions.
extern "C" ch_status ch_sys_evolve_observe(ch_sys *s, int64_t nsteps) {
    if(!s || nsteps < 0) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(s->obs_count == CH_OBS_QUEUE)
        return fail(c, CH_EINVAL, "%d observables already queued: fetch before queueing more", CH_OBS_QUEUE);
    const bool fusable = !is_species(s) && !is_spectral(s->d.integrator) && !has_noise(s) && s->graph_steps > 0 &&
                         nsteps >= s->tb && nsteps <= 4 * (int64_t)s->graph_steps;
    const int par = s->cur;
    if(!fusable || (!(s->oexec[par] && s->osteps[par] == nsteps) && s->oseen[par] != nsteps)) {
        // first sight of this count (or not capturable): issue it directly
        if(fusable) s->oseen[par] = nsteps;
        ch_status st = ch_sys_evolve(s, nsteps);
        if(st != CH_OK) return st;
        return ch_sys_observables_queue(s);
    }
    HIPCHK(c, hipSetDevice(c->device));
    if(!s->oexec[par] || s->osteps[par] != nsteps) {
        if(s->oexec[par]) {
            HIPCHK(c, hipGraphExecDestroy(s->oexec[par]));
            s->oexec[par] = nullptr;
        }
        HIPCHK(c, ensure_obs_host(s));
        hipGraph_t graph;
        int cur = par;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        int64_t n = nsteps;
        for(; n >= s->tb && e == hipSuccess; n -= s->tb) e = launch_steps(s, cur, s->tb);
        for(int K = s->tb / 2; K >= 1 && e == hipSuccess; K /= 2)
            if(n >= K) {
                e = launch_steps(s, cur, K);
                n -= K;
            }
        const int saved = s->cur;
        s->cur = cur;  // the reductions read the chain's last buffer
        if(e == hipSuccess) e = enqueue_observables(s, true);
        s->cur = saved;
        hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "launch during capture failed: %s", hipGetErrorString(e));
        if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
        e = hipGraphInstantiate(&s->oexec[par], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        s->osteps[par] = nsteps;
        s->ocur[par] = cur;
    }
    const int slot = (s->obs_head + s->obs_count) % CH_OBS_QUEUE;
    if(!s->obs_ev[slot]) HIPCHK(c, hipEventCreateWithFlags(&s->obs_ev[slot], hipEventDisableTiming));
    HIPCHK(c, hipGraphLaunch(s->oexec[par], c->stream));
    s->cur = s->ocur[par];
    HIPCHK(c, hipEventRecord(s->obs_ev[slot], c->stream));
    s->obs_count++;
    return CH_OK;
}

extern "C" ch_status ch_sys_observables_fetch(ch_sys *s, double *mass, double *free_energy) {
    if(!s) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(s->obs_count == 0) return fail(c, CH_EINVAL, "no observables queued");
    HIPCHK(c, hipSetDevice(c->device));
    const int slot = s->obs_head;
    HIPCHK(c, hipEventSynchronize(s->obs_ev[slot]));
    s->obs_head = (s->obs_head + 1) % CH_OBS_QUEUE;
    s->obs_count--;
    return observables_result(s, s->obs_host + 2 * (1 + slot), mass, free_energy);
}

extern "C" int32_t ch_sys_observables_pending(const ch_sys *s) { return s ? s->obs_count : 0; }

static ch_status observables_result(ch_sys *s, const double *pair, double *mass, double *free_energy) {
    ch_ctx *c = s->ctx;
    const double h[2] = {pair[0], pair[1]};
    const double m = h[0] * s->p.vbin_over_bins, f = h[1] * s->p.vbin_over_bins;
    if(mass) *mass = m;
    if(free_energy) *free_energy = f;
    if(std::isnan(h[0])) return fail(c, CH_ENAN, "Encountered a nan while computing the total mass");
    if(std::isnan(h[1])) return fail(c, CH_ENAN, "Encountered a nan while computing the total free energy");
    return CH_OK;
}

// The reductions of ch_sys_observables on a caller-held device field, with
// no allocation: the partial sums go to the library's fixed per-device
// scratch array (reduce.hip), the final pair is copied to the stack.
extern "C" ch_status ch_observables_device_noalloc(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev,
                                                   double *mass, double *free_energy) {
    if(!ctx) return CH_EINVAL;
    if(!desc || !rho_dev) return fail(ctx, CH_EINVAL, "null descriptor or field");
    ch_status st = validate(ctx, desc);
    if(st != CH_OK) return st;
    if(desc->model != CH_LANDAU && desc->model != CH_SIMPLE_WERTHEIM)
        return fail(ctx, CH_EUNSUPPORTED, "allocation-free observables take single-species models only");
    int64_t bins = 1;
    for(int i = 0; i < desc->dims; i++) bins *= desc->N;
    if(bins > chd::kNoAllocMaxCells)
        return fail(ctx, CH_EINVAL, "field of %lld cells exceeds the %lld the allocation-free reduction takes",
                    (long long)bins, (long long)chd::kNoAllocMaxCells);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    double *scr = chd::noalloc_scratch(ctx->device);
    if(!scr) return fail(ctx, CH_ESTATE, "no scratch address for device %d", ctx->device);
    const chd::Params p = make_params(desc);
    double *result = scr + 2 * chd::kObsParts;  // after the partials
    HIPCHK(ctx, chd::observables(p, desc->precision, desc->model, rho_dev, scr, chd::kObsParts, result, ctx->stream));
    double h[2];
    HIPCHK(ctx, hipMemcpyAsync(h, result, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if(mass) *mass = h[0] * p.vbin_over_bins;
    if(free_energy) *free_energy = h[1] * p.vbin_over_bins;
    if(std::isnan(h[0])) return fail(ctx, CH_ENAN, "Encountered a nan while computing the total mass");
    if(std::isnan(h[1])) return fail(ctx, CH_ENAN, "Encountered a nan while computing the total free energy");
    return CH_OK;
}

// ------------------------------------------------------------------ S(q)
extern "C" int32_t ch_sq_nshells(int32_t N) { return N < 2 ? 0 : (int32_t)std::lrint(std::sqrt(2.0) * (N / 2)) + 1; }

// spectral scratch + twiddles for systems created without them (stencil paths)
static ch_status ensure_spectral(ch_sys *s) {
    ch_ctx *c = s->ctx;
    const int N = s->d.N;
    if(!s->work && hipMalloc(&s->work, s->hat * 2 * s->elt) != hipSuccess) {
        s->work = nullptr;
        return fail(c, CH_ENOMEM, "no device memory for the spectrum");
    }
    if(!s->twiddles) {
        if(hipMalloc(&s->twiddles, (size_t)N * 2 * s->elt) != hipSuccess) {
            s->twiddles = nullptr;
            return fail(c, CH_ENOMEM, "no device memory for the twiddles");
        }
        std::vector<double> twd(2 * N);
        std::vector<float> twf(2 * N);
        for(int m = 0; m < N; m++) {
            long double cc, sn;
            twiddle(m, N, cc, sn);
            twd[2 * m] = (double)cc; twd[2 * m + 1] = (double)sn;
            twf[2 * m] = (float)twd[2 * m]; twf[2 * m + 1] = (float)twd[2 * m + 1];
        }
        const void *src = s->d.precision == CH_FP64 ? (const void *)twd.data() : (const void *)twf.data();
        HIPCHK(c, hipMemcpy(s->twiddles, src, (size_t)N * 2 * s->elt, hipMemcpyHostToDevice));
    }
    return CH_OK;
}

// S(q) of the current field into sq_dev[0..ns) (and shell counts into cnt_dev)
static ch_status sq_compute(ch_sys *s, int ns, double *sq_dev, double *cnt_dev) {
    ch_ctx *c = s->ctx;
    chd::SpectralArgs a = spectral_args(s);
    a.rho = s->psi[s->cur];
    HIPCHK(c, chd::fft2_forward(a, c->stream));
    HIPCHK(c, chd::sq_shells(s->d.precision, s->work, s->d.N, ns, sq_dev, cnt_dev, c->stream));
    return CH_OK;
}

extern "C" ch_status ch_sys_structure_factor(ch_sys *s, double *Sq, int32_t nshells) {
    if(!s || !Sq) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(is_species(s)) return fail(c, CH_EUNSUPPORTED, "S(q) covers single-species systems only");
    if(s->d.dims != 2) return fail(c, CH_EUNSUPPORTED, "the structure factor is defined for 2D fields");
    const int ns = ch_sq_nshells(s->d.N);
    if(nshells != ns) return fail(c, CH_EINVAL, "N = %d has %d shells (got %d)", s->d.N, ns, nshells);
    HIPCHK(c, hipSetDevice(c->device));
    ch_status st = ensure_spectral(s);
    if(st != CH_OK) return st;
    double *buf = nullptr;
    HIPCHK(c, hipMalloc((void **)&buf, 2 * ns * sizeof(double)));
    st = sq_compute(s, ns, buf, buf + ns);
    if(st == CH_OK) {
        hipError_t e = hipMemcpyAsync(Sq, buf, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if(e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if(e != hipSuccess) st = fail(c, CH_EHIP, "structure factor copy failed: %s", hipGetErrorString(e));
    }
    (void)hipFree(buf);
    return st;
}

extern "C" ch_status ch_sys_sq_loss(ch_sys *s, const double *target, int32_t nshells, double *loss, double *grad) {
    if(!s || !target) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(is_species(s)) return fail(c, CH_EUNSUPPORTED, "S(q) covers single-species systems only");
    if(s->d.dims != 2) return fail(c, CH_EUNSUPPORTED, "the structure factor is defined for 2D fields");
    const int ns = ch_sq_nshells(s->d.N);
    if(nshells != ns) return fail(c, CH_EINVAL, "N = %d has %d shells (got %d)", s->d.N, ns, nshells);
    HIPCHK(c, hipSetDevice(c->device));
    ch_status st = ensure_spectral(s);
    if(st != CH_OK) return st;
    double *buf = nullptr;
    HIPCHK(c, hipMalloc((void **)&buf, 3 * ns * sizeof(double)));
    double *sq = buf, *cnt = buf + ns, *tg = buf + 2 * ns;
    std::vector<double> h(ns);
    hipError_t e = hipMemcpyAsync(tg, target, ns * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if(e == hipSuccess) {
        st = sq_compute(s, ns, sq, cnt);
        if(st == CH_OK) e = hipMemcpyAsync(h.data(), sq, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if(st == CH_OK && e == hipSuccess && grad) {
            e = chd::sq_scale(s->d.precision, s->work, s->d.N, sq, tg, cnt, c->stream);
            chd::SpectralArgs a = spectral_args(s);
            if(e == hipSuccess) e = chd::fft2_inverse(a, s->stage2, c->stream);
            if(e == hipSuccess && s->d.precision == CH_FP32) e = chd::convert_f2d((const float *)s->stage2, (double *)s->stage, s->bins, c->stream);
            if(e == hipSuccess)
                e = hipMemcpyAsync(grad, s->d.precision == CH_FP64 ? s->stage2 : s->stage, s->bins * 8, hipMemcpyDeviceToHost, c->stream);
        }
        if(e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(buf);
    if(st != CH_OK) return st;
    if(e != hipSuccess) return fail(c, CH_EHIP, "structure-factor loss failed: %s", hipGetErrorString(e));
    if(loss) {
        double L = 0;
        for(int i = 0; i < ns; i++) L += (h[i] - target[i]) * (h[i] - target[i]);
        *loss = L;
    }
    return CH_OK;
}

extern "C" ch_status ch_sys_free_energy_grad(ch_sys *s, double *F, double *dF_dpsi, double *dF_dparams) {
    if(!s) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(is_species(s)) return fail(c, CH_EUNSUPPORTED, "the free-energy gradient covers single-species systems only");
    HIPCHK(c, hipSetDevice(c->device));
    constexpr int NS = chd::kFeSlots;
    const int max_parts = chd::kObsParts * 2 / NS;
    int nb = 0;
    HIPCHK(c, chd::free_energy_grad(s->p, s->d.precision, s->d.model, s->psi[s->cur], (double *)s->stage, s->partials,
                                    max_parts, &nb, c->stream));
    std::vector<double> h((size_t)NS * nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), s->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if(dF_dpsi) HIPCHK(c, hipMemcpyAsync(dF_dpsi, s->stage, s->bins * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double sum[NS] = {0, 0, 0, 0, 0};
    for(int b = 0; b < nb; b++)
        for(int j = 0; j < NS; j++) sum[j] += h[(size_t)NS * b + j];
    const double vb = s->p.vbin_over_bins;
    const double Fv = vb * (sum[0] + s->d.k * sum[1]);
    if(std::isnan(Fv)) return fail(c, CH_ENAN, "Encountered a nan while computing the free energy");
    if(F) *F = Fv;
    if(dF_dparams) {
        const bool landau = s->d.model == CH_LANDAU;
        dF_dparams[CH_GRAD_EPSILON] = landau ? vb * sum[2] : 0.0;
        dF_dparams[CH_GRAD_K] = vb * sum[1];
        dF_dparams[CH_GRAD_M] = 0.0;
        dF_dparams[CH_GRAD_DELTA] = landau ? 0.0 : vb * sum[3];
        dF_dparams[CH_GRAD_B2] = landau ? 0.0 : vb * sum[4];
    }
    return CH_OK;
}

// ------------------------------------------------------------------ batches
// the steps-per-launch depths an item list exists for (index = log2 K)
static constexpr int kBatchDepths = 4;

struct ch_batch {
    ch_ctx *ctx = nullptr;
    int n = 0, precision = CH_FP64, model = CH_LANDAU, graph_steps = 0;
    int tb = 4, shape = 0;  // steps per launch and region split (autotuned)
    bool tuned = false;     // false: tune on the first evolve
    size_t elt = 8;
    int64_t total = 0;                       // elements over all replicas
    std::vector<chd::BatchRep> host_reps;
    chd::BatchRep *reps = nullptr;
    // work list per region width (64, 128 or 192 columns) and K = 1, 2, 4, 8
    int2 *items[chd::kBatchFamilies][kBatchDepths] = {};
    int nitems[chd::kBatchFamilies][kBatchDepths] = {};
    void *buf[2] = {nullptr, nullptr};
    int compensation = CH_COMP_NONE;  // CH_COMP_* (FP32 only)
    // CH_COMP_STORED / CH_COMP_LOW8: the low-order parts of each replica's
    // field (FP32 or int8 per element), ping-pong beside buf
    void *comp[2] = {nullptr, nullptr};
    size_t comp_elt() const { return compensation == CH_COMP_LOW8 ? 1 : elt; }
    int cur = 0;
    void *stage = nullptr;                   // largest replica, doubles
    double *result = nullptr;                // 2 n doubles
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
};

static int log2i(int k) { return k >= 8 ? 3 : (k >= 4 ? 2 : (k >= 2 ? 1 : 0)); }

// Autotuned (shape, steps per launch) per batch composition: device,
// precision, model, compensation and the number of replicas at each N.  Kept
// for the process, and in the file named by CH_TUNE_CACHE when set (one line
// "key shape K" per entry, read on first use, appended after each tuning), so
// a batch of a composition already tuned starts at once.
static std::mutex g_tune_mu;
static std::map<std::string, std::pair<int, int>> g_tune;
static bool g_tune_loaded = false;

static std::string tune_key(const ch_batch *b) {
    std::map<int, int> count;
    for(const chd::BatchRep &r : b->host_reps) count[r.p.N]++;
    char buf[96];
    std::snprintf(buf, sizeof buf, "dev%d/p%d/m%d/c%d", b->ctx->device, b->precision, b->model, b->compensation);
    std::string k = buf;
    for(const auto &e : count) k += "/" + std::to_string(e.first) + "x" + std::to_string(e.second);
    return k;
}

static bool tune_lookup(const std::string &key, int &shape, int &K) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if(!g_tune_loaded) {
        g_tune_loaded = true;
        if(const char *path = std::getenv("CH_TUNE_CACHE")) {
            if(FILE *f = std::fopen(path, "r")) {
                char k[512];
                int sh, kk;
                while(std::fscanf(f, "%511s %d %d", k, &sh, &kk) == 3)
                    if(sh >= 0 && sh < chd::kBatchShapes && (kk == 1 || kk == 2 || kk == 4 || kk == 8)) g_tune[k] = {sh, kk};
                std::fclose(f);
            }
        }
    }
    auto it = g_tune.find(key);
    if(it == g_tune.end()) return false;
    shape = it->second.first;
    K = it->second.second;
    return true;
}

static void tune_store(const std::string &key, int shape, int K) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[key] = {shape, K};
    if(const char *path = std::getenv("CH_TUNE_CACHE"))
        if(FILE *f = std::fopen(path, "a")) {
            std::fprintf(f, "%s %d %d\n", key.c_str(), shape, K);
            std::fclose(f);
        }
}

static void free_batch(ch_batch *b) {
    for(int i = 0; i < 2; i++) {
        if(b->gexec[i]) (void)hipGraphExecDestroy(b->gexec[i]);
        if(b->buf[i]) (void)hipFree(b->buf[i]);
        if(b->comp[i]) (void)hipFree(b->comp[i]);
    }
    for(auto &fam : b->items)
        for(int2 *it : fam)
            if(it) (void)hipFree(it);
    void *bufs[] = {b->reps, b->stage, b->result};
    for(void *x : bufs)
        if(x) (void)hipFree(x);
    delete b;
}

extern "C" ch_status ch_batch_create(ch_ctx *ctx, int32_t n, const ch_desc *descs, ch_batch **out) {
    if(!ctx || !out || !descs || n < 1) return CH_EINVAL;
    return ch_batch_create_ex(ctx, n, descs, descs[0].precision == CH_FP32 ? CH_COMP_LAUNCH : CH_COMP_NONE, out);
}

extern "C" int32_t ch_batch_compensation(const ch_batch *b) { return b ? b->compensation : -1; }

extern "C" ch_status ch_batch_create_ex(ch_ctx *ctx, int32_t n, const ch_desc *descs, int32_t compensation,
                                        ch_batch **out) {
    if(!ctx || !out || !descs || n < 1) return CH_EINVAL;
    *out = nullptr;
    if(compensation < CH_COMP_NONE || compensation > CH_COMP_LOW8)
        return fail(ctx, CH_EINVAL, "unknown batch compensation %d", compensation);
    if(descs[0].precision == CH_FP64 && compensation != CH_COMP_NONE)
        return fail(ctx, CH_EINVAL, "FP64 batches take no compensation (CH_COMP_NONE)");
    for(int r = 0; r < n; r++) {
        ch_status st = validate(ctx, &descs[r]);
        if(st != CH_OK) return fail(ctx, st, "replica %d: %s", r, ctx->err.c_str());
        const ch_desc &d = descs[r];
        if(d.dims != 2 || d.integrator != CH_EULER || d.N < 64)
            return fail(ctx, CH_EUNSUPPORTED, "replica %d: batches hold 2D euler systems with N >= 64", r);
        if(d.precision != descs[0].precision || d.model != descs[0].model)
            return fail(ctx, CH_EINVAL, "replica %d: all replicas share one precision and one model", r);
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ch_batch *b = new ch_batch();
    b->ctx = ctx;
    b->n = n;
    b->precision = descs[0].precision;
    b->model = descs[0].model;
    b->elt = b->precision == CH_FP64 ? 8 : 4;
    b->compensation = compensation;
    // CH_BATCH_TB / CH_BATCH_SHAPE pin the configuration (no autotuning)
    const char *etb = std::getenv("CH_BATCH_TB"), *esh = std::getenv("CH_BATCH_SHAPE");
    if(etb || esh) {
        b->tuned = true;
        if(etb) {
            const int want = std::atoi(etb);
            b->tb = want >= 8 ? 8 : (want >= 4 ? 4 : (want >= 2 ? 2 : 1));
        }
        if(esh) b->shape = std::max(0, std::min(chd::kBatchShapes - 1, std::atoi(esh)));
    }
    std::vector<int2> items[chd::kBatchFamilies][kBatchDepths];
    int64_t maxbins = 0;
    for(int r = 0; r < n; r++) {
        chd::BatchRep rep{};
        rep.p = make_params(&descs[r]);
        rep.off = b->total;
        const int N = descs[r].N;
        rep.whole = N == chd::kBatchRegion ? 1 : 0;
        for(int fam = 0; fam < chd::kBatchFamilies; fam++)
            for(int i = 0; i < kBatchDepths; i++)
                for(int code : chd::batch_items(chd::batch_family_shape(fam), N, 1 << i))
                    items[fam][i].push_back(int2{r, code});
        b->total += (int64_t)N * N;
        maxbins = std::max<int64_t>(maxbins, (int64_t)N * N);
        b->host_reps.push_back(rep);
    }
    const char *g = std::getenv("CH_GRAPH_STEPS");
    b->graph_steps = g ? std::atoi(g) : 64;
    if(b->graph_steps < 0) b->graph_steps = 0;
    if(!b->tuned && tune_lookup(tune_key(b), b->shape, b->tb)) b->tuned = true;
    bool ok = hipMalloc(&b->buf[0], b->total * b->elt) == hipSuccess && hipMalloc(&b->buf[1], b->total * b->elt) == hipSuccess &&
              hipMalloc((void **)&b->reps, n * sizeof(chd::BatchRep)) == hipSuccess &&
              hipMalloc(&b->stage, maxbins * 8) == hipSuccess && hipMalloc((void **)&b->result, 2 * n * sizeof(double)) == hipSuccess;
    if(ok && (compensation == CH_COMP_STORED || compensation == CH_COMP_LOW8)) {
        const size_t bytes = b->total * b->comp_elt();
        ok = hipMalloc(&b->comp[0], bytes) == hipSuccess && hipMalloc(&b->comp[1], bytes) == hipSuccess &&
             hipMemset(b->comp[0], 0, bytes) == hipSuccess && hipMemset(b->comp[1], 0, bytes) == hipSuccess;
    }
    for(int fam = 0; fam < chd::kBatchFamilies; fam++)
        for(int i = 0; ok && i < kBatchDepths; i++) {
            const std::vector<int2> &v = items[fam][i];
            b->nitems[fam][i] = (int)v.size();
            ok = hipMalloc((void **)&b->items[fam][i], v.size() * sizeof(int2)) == hipSuccess &&
                 hipMemcpy(b->items[fam][i], v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
        }
    if(ok) ok = hipMemcpy(b->reps, b->host_reps.data(), n * sizeof(chd::BatchRep), hipMemcpyHostToDevice) == hipSuccess;
    if(!ok) {
        free_batch(b);
        return fail(ctx, CH_ENOMEM, "device allocation failed for a batch of %d replicas", n);
    }
    *out = b;
    return CH_OK;
}

extern "C" void ch_batch_destroy(ch_batch *b) {
    if(!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    free_batch(b);
}

extern "C" int32_t ch_batch_steps_per_launch(const ch_batch *b) { return b ? b->tb : 0; }

extern "C" ch_status ch_batch_upload(ch_batch *b, int32_t r, const double *rho_host) {
    if(!b || !rho_host || r < 0 || r >= b->n) return CH_EINVAL;
    ch_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const chd::BatchRep &rep = b->host_reps[r];
    const int64_t bins = (int64_t)rep.p.N * rep.p.N;
    char *dst = (char *)b->buf[b->cur] + rep.off * b->elt;
    // a new state: nothing of the replica's earlier updates is owed to it
    if(b->comp[0])
        HIPCHK(c, hipMemsetAsync((char *)b->comp[b->cur] + rep.off * b->comp_elt(), 0, bins * b->comp_elt(), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if(b->precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(dst, rho_host, bins * 8, hipMemcpyHostToDevice, c->stream));
    }
    else {
        HIPCHK(c, hipMemcpyAsync(b->stage, rho_host, bins * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, chd::convert_d2f((const double *)b->stage, (float *)dst, bins, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_batch_download(ch_batch *b, int32_t r, double *rho_host) {
    if(!b || !rho_host || r < 0 || r >= b->n) return CH_EINVAL;
    ch_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const chd::BatchRep &rep = b->host_reps[r];
    const int64_t bins = (int64_t)rep.p.N * rep.p.N;
    const char *src = (const char *)b->buf[b->cur] + rep.off * b->elt;
    if(b->precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(rho_host, src, bins * 8, hipMemcpyDeviceToHost, c->stream));
    }
    else {
        HIPCHK(c, chd::convert_f2d((const float *)src, (double *)b->stage, bins, c->stream));
        HIPCHK(c, hipMemcpyAsync(rho_host, b->stage, bins * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

static hipError_t batch_launch(ch_batch *b, int shape, int K, int from) {
    const int i = log2i(K), fam = chd::batch_family(shape);
    return chd::batch_launch(b->precision, b->model, b->compensation, shape, K, b->buf[from], b->buf[from ^ 1], b->reps,
                             b->items[fam][i], b->nitems[fam][i], b->comp[from], b->comp[from ^ 1], b->ctx->stream);
}

static hipError_t batch_step(ch_batch *b, int &cur, int K) {
    hipError_t e = batch_launch(b, b->shape, K, cur);
    cur ^= 1;
    return e;
}

// Times every (shape, K) candidate on the batch's own data: each trial
// launch reads the current buffer and writes the other one, so the state is
// unchanged.  A candidate's time is the best of three runs of 8 launches (the
// candidates differ by a few per cent); keeps the fastest per time step.
static ch_status batch_tune(ch_batch *b) {
    ch_ctx *c = b->ctx;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    float best = 1e30f;
    int bs = b->shape, bk = b->tb;
    hipError_t err = hipSuccess;
    // a compensated batch rounds psi once per launch, so its accuracy
    // depends on K: its candidates keep K >= 4
    const bool comp = b->compensation != CH_COMP_NONE;
    for(int shape = 0; shape < chd::kBatchShapes && err == hipSuccess; shape++) {
        for(int K : {2, 4, 8}) {
            if(comp && K < 4) continue;
            constexpr int warm = 2, reps = 8, runs = 3;
            for(int r = 0; r < warm && err == hipSuccess; r++) err = batch_launch(b, shape, K, b->cur);
            float ms = 1e30f;
            for(int run = 0; run < runs && err == hipSuccess; run++) {
                err = hipEventRecord(e0, c->stream);
                for(int r = 0; r < reps && err == hipSuccess; r++) err = batch_launch(b, shape, K, b->cur);
                if(err == hipSuccess) err = hipEventRecord(e1, c->stream);
                if(err == hipSuccess) err = hipEventSynchronize(e1);
                float t = 0;
                if(err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
                ms = std::min(ms, t);
            }
            if(err != hipSuccess) break;
            const float per_step = ms / (reps * K);
            if(per_step < best) {
                best = per_step;
                bs = shape;
                bk = K;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if(err != hipSuccess) return fail(c, CH_EHIP, "batch autotuning failed: %s", hipGetErrorString(err));
    if(bs != b->shape || bk != b->tb) {
        for(int i = 0; i < 2; i++)
            if(b->gexec[i]) {
                (void)hipGraphExecDestroy(b->gexec[i]);
                b->gexec[i] = nullptr;
            }
    }
    b->shape = bs;
    b->tb = bk;
    b->tuned = true;
    tune_store(tune_key(b), bs, bk);
    return CH_OK;
}

extern "C" ch_status ch_batch_autotune(ch_batch *b, int32_t *shape, int32_t *steps_per_launch) {
    if(!b) return CH_EINVAL;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    ch_status st = batch_tune(b);
    if(st != CH_OK) return st;
    if(shape) *shape = b->shape;
    if(steps_per_launch) *steps_per_launch = b->tb;
    return CH_OK;
}

extern "C" ch_status ch_batch_evolve(ch_batch *b, int64_t nsteps) {
    if(!b || nsteps < 0) return CH_EINVAL;
    ch_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if(!b->tuned && nsteps > 0) {
        ch_status st = batch_tune(b);
        if(st != CH_OK) return st;
    }
    // graph chunk: a multiple of 2 tb launches' worth of steps (parity restored)
    const int G = b->graph_steps - b->graph_steps % (2 * b->tb);
    if(G > 0 && nsteps >= G) {
        const int par = b->cur;
        if(!b->gexec[par]) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int cur = b->cur;
            hipError_t e = hipSuccess;
            for(int i = 0; i < G && e == hipSuccess; i += b->tb) e = batch_step(b, cur, b->tb);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "batch launch during capture failed: %s", hipGetErrorString(e));
            if(e2 != hipSuccess) return fail(c, CH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&b->gexec[par], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if(e != hipSuccess) return fail(c, CH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        const int64_t reps = nsteps / G;
        for(int64_t r = 0; r < reps; r++) HIPCHK(c, hipGraphLaunch(b->gexec[par], c->stream));
        nsteps -= reps * G;
    }
    for(; nsteps >= b->tb; nsteps -= b->tb) HIPCHK(c, batch_step(b, b->cur, b->tb));
    for(int K = b->tb / 2; K >= 1; K /= 2)
        if(nsteps >= K) {
            HIPCHK(c, batch_step(b, b->cur, K));
            nsteps -= K;
        }
    return CH_OK;
}

extern "C" ch_status ch_batch_observables(ch_batch *b, double *mass, double *free_energy) {
    if(!b) return CH_EINVAL;
    ch_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, chd::batch_observables(b->precision, b->model, b->buf[b->cur], b->reps, b->n, b->result, c->stream));
    std::vector<double> h(2 * (size_t)b->n);
    HIPCHK(c, hipMemcpyAsync(h.data(), b->result, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    bool nan = false;
    for(int r = 0; r < b->n; r++) {
        const double s = b->host_reps[r].p.vbin_over_bins;
        if(mass) mass[r] = h[2 * r] * s;
        if(free_energy) free_energy[r] = h[2 * r + 1] * s;
        nan = nan || std::isnan(h[2 * r]) || std::isnan(h[2 * r + 1]);
    }
    if(nan) return fail(c, CH_ENAN, "Encountered a nan while computing the observables of a replica");
    return CH_OK;
}

// ------------------------------------------------------------------ reverse mode
static bool grow(void **buf, int64_t &cap, int64_t fields, size_t field_bytes) {
    if(fields <= cap && *buf) return true;
    if(*buf) (void)hipFree(*buf);
    *buf = nullptr;
    cap = 0;
    if(hipMalloc(buf, (size_t)fields * field_bytes) != hipSuccess) {
        *buf = nullptr;
        return false;
    }
    cap = fields;
    return true;
}

static char *field_at(void *base, int64_t i, size_t field_bytes) { return (char *)base + (size_t)i * field_bytes; }

// the 3D euler_mobility reverse sweep in three launches per step
// (chd::adjoint_split3d) and its window recompute in two
static bool adj_split(const ch_sys *s) {
    return s->d.dims == 3 && s->d.integrator == CH_EULER_MOBILITY && chd::adjoint_split3d_ok(s->p, s->d.mobility);
}

// Tape slots: psi_t, and with the three-launch reverse step also mu_t and the
// mobilities of psi_t behind it (the recompute's two-launch step computes
// them on its way to psi_{t+1}, so the reverse step reads them back).
static size_t tape_slot(const ch_sys *s) { return s->bins * s->elt * (adj_split(s) ? 3 : 1); }

// one recorded state: psi_{t+1} = step(psi_t) from the slot at `in` into `out`
static hipError_t tape_step1(ch_sys *s, const void *in, void *out, hipStream_t st) {
    const ch_desc &d = s->d;
    if(adj_split(s)) {
        const size_t fb = s->bins * s->elt;
        return chd::stencil_split3d(s->p, d.precision, d.integrator, d.model, d.mobility, in, out, (char *)in + fb,
                                    (char *)in + 2 * fb, st);
    }
    return chd::tape_launch(s->p, d.precision, d.model, d.mobility, d.integrator == CH_EULER_MOBILITY, 1, in, out, st);
}

// `count` states after tape slot `from` of `tape` into slots from+1 .. from+count
static hipError_t tape_steps(ch_sys *s, void *tape, int64_t from, int64_t count, hipStream_t st) {
    const size_t sb = tape_slot(s);
    const bool em = s->d.integrator == CH_EULER_MOBILITY;
    while(count > 0) {
        const int K = chd::tape_depth(s->d.dims, s->d.N, em, count);
        hipError_t e = K == 1 ? tape_step1(s, field_at(tape, from, sb), field_at(tape, from + 1, sb), st)
                              : chd::tape_launch(s->p, s->d.precision, s->d.model, s->d.mobility, em, K,
                                                 field_at(tape, from, sb), field_at(tape, from + 1, sb), st);
        if(e != hipSuccess) return e;
        from += K;
        count -= K;
    }
    return hipSuccess;
}

extern "C" ch_status ch_sys_record(ch_sys *s, int64_t nsteps, int32_t ckpt_every) {
    if(!s || nsteps < 1 || ckpt_every < 0) return CH_EINVAL;
    ch_ctx *c = s->ctx;
    if(has_noise(s)) return fail(c, CH_EUNSUPPORTED, "reverse mode does not cover euler_mobility with noise");
    if(s->d.integrator == CH_PSEUDOSPECTRAL_MOBILITY)
        return fail(c, CH_EUNSUPPORTED, "reverse mode does not cover pseudospectral_mobility");
    if(is_species(s)) return fail(c, CH_EUNSUPPORTED, "reverse mode covers single-species models only");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t fb = s->bins * s->elt;
    const bool ps = s->d.integrator == CH_PSEUDOSPECTRAL;
    const size_t hb = (size_t)s->hat * 2 * s->elt;
    s->rec_valid = false;
    s->rec_full_tape = ckpt_every == 0;
    s->rec_window = ckpt_every == 0 ? nsteps : std::min<int64_t>(ckpt_every, nsteps);
    const int64_t nck = (nsteps + s->rec_window - 1) / s->rec_window;
    // checkpointed: two window tapes, one recomputed while the other is swept;
    // the pseudospectral full tape also keeps the final state
    const int64_t tape_fields = s->rec_full_tape ? nsteps + (ps ? 1 : 0) : 2 * s->rec_window;
    if(adj_split(s))
        for(void *&b : s->adj_split)
            if(!b && hipMalloc(&b, fb) != hipSuccess) {
                b = nullptr;
                return fail(c, CH_ENOMEM, "no device memory for the reverse-sweep work fields");
            }
    if(!grow(&s->tape, s->tape_cap, tape_fields, tape_slot(s)) ||
       (!s->rec_full_tape && !grow(&s->ckpt, s->ckpt_cap, nck, fb)) ||
       (ps && !s->rec_full_tape && !grow(&s->ckpt_hat, s->ckpt_hat_cap, nck, hb)))
        return fail(c, CH_ENOMEM, "no device memory for %lld checkpoints / a %lld-state tape", (long long)nck,
                    (long long)tape_fields);
    if(ps) {
        chd::SpectralArgs a = spectral_args(s);
        if(s->rec_full_tape) {
            // psi_0 .. psi_T on the tape, the live spectrum evolved to step T
            HIPCHK(c, hipMemcpyAsync(s->tape, s->psi[0], fb, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, chd::ps_steps_tape(a, c->stream, nsteps, s->tape));
            HIPCHK(c, hipMemcpyAsync(s->psi[0], field_at(s->tape, nsteps, fb), fb, hipMemcpyDeviceToDevice, c->stream));
        }
        else {
            for(int64_t k = 0; k < nck; k++) {
                HIPCHK(c, hipMemcpyAsync(field_at(s->ckpt, k, fb), s->psi[0], fb, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(field_at(s->ckpt_hat, k, hb), s->rho_hat, hb, hipMemcpyDeviceToDevice, c->stream));
                const int64_t steps = std::min<int64_t>(s->rec_window, nsteps - k * s->rec_window);
                ch_status st = ch_sys_evolve(s, steps);
                if(st != CH_OK) return st;
            }
        }
        s->rec_steps = nsteps;
        s->rec_valid = true;
        return CH_OK;
    }
    if(s->rec_full_tape) {
        // every state: psi_0 .. psi_{T-1} on the tape, psi_T in the field buffer
        HIPCHK(c, hipMemcpyAsync(s->tape, s->psi[s->cur], fb, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, tape_steps(s, s->tape, 0, nsteps - 1, c->stream));
        HIPCHK(c, tape_step1(s, field_at(s->tape, nsteps - 1, tape_slot(s)), s->psi[s->cur], c->stream));
    }
    else {
        for(int64_t k = 0; k < nck; k++) {
            HIPCHK(c, hipMemcpyAsync(field_at(s->ckpt, k, fb), s->psi[s->cur], fb, hipMemcpyDeviceToDevice, c->stream));
            const int64_t steps = std::min<int64_t>(s->rec_window, nsteps - k * s->rec_window);
            ch_status st = ch_sys_evolve(s, steps);
            if(st != CH_OK) return st;
        }
    }
    s->rec_steps = nsteps;
    s->rec_valid = true;
    return CH_OK;
}

// Reverse sweep, windows last to first.  Checkpointed records recompute window
// k - 1 on the auxiliary stream into one tape buffer while the adjoint of
// window k runs on the context stream from the other; events order the reuse
// of each buffer.
static ch_status adjoint_core(ch_sys *s, double *grads) {
    ch_ctx *c = s->ctx;
    const size_t fb = s->bins * s->elt;
    const bool em = s->d.integrator == CH_EULER_MOBILITY;
    const bool ps = s->d.integrator == CH_PSEUDOSPECTRAL;
    const int ntiles = ps ? chd::kPsGradBlocks : chd::adjoint_tiles(s->p, em);
    const size_t hb = (size_t)s->hat * 2 * s->elt;
    const int64_t W = s->rec_window, T = s->rec_steps;
    const int64_t nck = (T + W - 1) / W;
    constexpr int NS = chd::kGradSlots;
    const bool recompute = !s->rec_full_tape;
    hipStream_t ms = c->stream;
    int li = 0;  // lambda lives in lam[li]
    HIPCHK(c, chd::fill_zero(s->gacc, NS * ntiles, ms));
    if(recompute) {
        HIPCHK(c, hipEventRecord(s->ev_start, ms));
        HIPCHK(c, hipStreamWaitEvent(s->aux, s->ev_start, 0));
    }
    bool used[2] = {false, false};
    for(int64_t k = nck - 1; k >= 0; k--) {
        const int64_t start = k * W, len = std::min<int64_t>(W, T - start);
        const int b = (int)(k & 1);
        void *tape = recompute ? field_at(s->tape, b * W, tape_slot(s)) : s->tape;
        if(recompute) {
            if(used[b]) HIPCHK(c, hipStreamWaitEvent(s->aux, s->ev_adj[b], 0));
            HIPCHK(c, hipMemcpyAsync(tape, field_at(s->ckpt, k, fb), fb, hipMemcpyDeviceToDevice, s->aux));
            if(ps) {
                // the chain from (psi, rho_hat) of checkpoint k, in scratch spectra
                HIPCHK(c, hipMemcpyAsync(s->ps_rh, field_at(s->ckpt_hat, k, hb), hb, hipMemcpyDeviceToDevice, s->aux));
                chd::SpectralArgs a = spectral_args(s);
                a.rho = tape;
                a.rho_hat = s->ps_rh;
                a.work = s->ps_w;
                HIPCHK(c, chd::ps_steps_tape(a, s->aux, len - 1, tape));
            }
            else {
                HIPCHK(c, tape_steps(s, tape, 0, len - 1, s->aux));
                if(adj_split(s)) {
                    // mu and the mobilities of the window's last state
                    char *last = field_at(tape, len - 1, tape_slot(s));
                    HIPCHK(c, chd::split3d_mu(s->p, s->d.precision, s->d.integrator, s->d.model, s->d.mobility, last,
                                              last + fb, last + 2 * fb, s->aux));
                }
            }
            HIPCHK(c, hipEventRecord(s->ev_rec[b], s->aux));
            HIPCHK(c, hipStreamWaitEvent(ms, s->ev_rec[b], 0));
        }
        if(ps) {
            const chd::SpectralArgs a = spectral_args(s);
            for(int64_t t = len - 1; t >= 0; t--) {
                HIPCHK(c, chd::ps_adjoint_step(a, s->lam[li], s->lam[li ^ 1], field_at(tape, t, fb), s->ps_spec,
                                               s->ps_real, s->gacc, ms));
                li ^= 1;
            }
        }
        if(adj_split(s)) {
            for(int64_t t = len - 1; t >= 0; t--) {
                char *slot = field_at(tape, t, tape_slot(s));
                void *const w[4] = {slot + fb, slot + 2 * fb, s->adj_split[0], nullptr};
                HIPCHK(c, chd::adjoint_split3d(s->p, s->d.precision, s->d.model, s->d.mobility, s->lam[li],
                                               s->lam[li ^ 1], slot, w, true, s->gacc, ms));
                li ^= 1;
            }
        }
        for(int64_t t = (ps || adj_split(s)) ? -1 : len - 1; t >= 0;) {
            const int ka = chd::adjoint_depth(s->d.dims, s->d.N, em, t + 1);
            HIPCHK(c, chd::adjoint_launch(s->p, s->d.precision, s->d.model, s->d.mobility, em, ka, s->lam[li],
                                          s->lam[li ^ 1], tape, (int)t, s->gacc, ms));
            li ^= 1;
            t -= ka;
        }
        if(recompute) {
            HIPCHK(c, hipEventRecord(s->ev_adj[b], ms));
            used[b] = true;
        }
    }
    std::vector<double> g(NS * (size_t)ntiles);
    HIPCHK(c, hipMemcpyAsync(g.data(), s->gacc, g.size() * sizeof(double), hipMemcpyDeviceToHost, ms));
    HIPCHK(c, hipStreamSynchronize(ms));
    double sum[NS] = {0, 0, 0, 0};
    for(int i = 0; i < ntiles; i++)
        for(int j = 0; j < NS; j++) sum[j] += g[NS * (size_t)i + j];
    if(grads && ps) {
        // slots: <b, rho>, k term, M term, delta term (dt is inside the spectral multipliers)
        const bool landau = s->d.model == CH_LANDAU;
        grads[CH_GRAD_EPSILON] = landau ? -sum[0] : 0.0;
        grads[CH_GRAD_K] = sum[1];
        grads[CH_GRAD_M] = sum[2];
        grads[CH_GRAD_DELTA] = landau ? 0.0 : sum[3];
        grads[CH_GRAD_B2] = landau ? 0.0 : 2.0 * sum[0];
    }
    else if(grads) {
        // slots: <nu, psi>, <nu, L psi>, M term, delta term; euler keeps w = nu / M
        const double dt = s->d.dt, f = em ? dt : dt * s->d.M;
        const bool landau = s->d.model == CH_LANDAU;
        grads[CH_GRAD_EPSILON] = landau ? -f * sum[0] : 0.0;
        grads[CH_GRAD_K] = -2.0 * f * sum[1];
        grads[CH_GRAD_M] = dt * sum[2];
        grads[CH_GRAD_DELTA] = landau ? 0.0 : f * sum[3];
        grads[CH_GRAD_B2] = landau ? 0.0 : 2.0 * f * sum[0];
    }
    if(li == 1) HIPCHK(c, hipMemcpyAsync(s->lam[0], s->lam[1], fb, hipMemcpyDeviceToDevice, ms));
    return CH_OK;
}

static ch_status adjoint_prepare(ch_sys *s) {
    ch_ctx *c = s->ctx;
    if(!s->rec_valid) return fail(c, CH_ESTATE, "ch_sys_adjoint needs a ch_sys_record first");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t fb = s->bins * s->elt;
    if(!s->aux) {
        // default priority, like the context stream
        HIPCHK(c, hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&s->ev_start, hipEventDisableTiming));
        for(int i = 0; i < 2; i++) {
            HIPCHK(c, hipEventCreateWithFlags(&s->ev_rec[i], hipEventDisableTiming));
            HIPCHK(c, hipEventCreateWithFlags(&s->ev_adj[i], hipEventDisableTiming));
        }
    }
    const bool ps = s->d.integrator == CH_PSEUDOSPECTRAL;
    const int ntiles = ps ? chd::kPsGradBlocks : chd::adjoint_tiles(s->p, s->d.integrator == CH_EULER_MOBILITY);
    int64_t cap0 = s->lam[0] ? 1 : 0, cap1 = s->lam[1] ? 1 : 0, capg = s->gacc ? 1 : 0;
    if(!grow(&s->lam[0], cap0, 1, fb) || !grow(&s->lam[1], cap1, 1, fb) ||
       !grow((void **)&s->gacc, capg, 1, chd::kGradSlots * sizeof(double) * ntiles))
        return fail(c, CH_ENOMEM, "no device memory for the adjoint");
    if(ps && !s->ps_rh) {
        const size_t hb = (size_t)s->hat * 2 * s->elt;
        bool ok = hipMalloc(&s->ps_rh, hb) == hipSuccess && hipMalloc(&s->ps_w, hb) == hipSuccess;
        for(int i = 0; ok && i < 3; i++) ok = hipMalloc(&s->ps_spec[i], hb) == hipSuccess;
        for(int i = 0; ok && i < 2; i++) ok = hipMalloc(&s->ps_real[i], fb) == hipSuccess;
        if(!ok) return fail(c, CH_ENOMEM, "no device memory for the pseudospectral adjoint");
    }
    return CH_OK;
}

extern "C" ch_status ch_sys_adjoint(ch_sys *s, const double *dL_dpsiT, double *dL_dpsi0, double *grads) {
    if(!s || !dL_dpsiT) return CH_EINVAL;
    ch_status st = adjoint_prepare(s);
    if(st != CH_OK) return st;
    ch_ctx *c = s->ctx;
    if(s->d.precision == CH_FP64) {
        HIPCHK(c, hipMemcpyAsync(s->lam[0], dL_dpsiT, s->bins * 8, hipMemcpyHostToDevice, c->stream));
    }
    else {
        HIPCHK(c, hipMemcpyAsync(s->stage, dL_dpsiT, s->bins * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, chd::convert_d2f((const double *)s->stage, (float *)s->lam[0], s->bins, c->stream));
    }
    st = adjoint_core(s, grads);
    if(st != CH_OK) return st;
    if(dL_dpsi0) {
        if(s->d.precision == CH_FP64) {
            HIPCHK(c, hipMemcpyAsync(dL_dpsi0, s->lam[0], s->bins * 8, hipMemcpyDeviceToHost, c->stream));
        }
        else {
            HIPCHK(c, chd::convert_f2d((const float *)s->lam[0], (double *)s->stage, s->bins, c->stream));
            HIPCHK(c, hipMemcpyAsync(dL_dpsi0, s->stage, s->bins * 8, hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_sys_adjoint_device(ch_sys *s, const void *dL_dpsiT_dev, void *dL_dpsi0_dev, double *grads) {
    if(!s || !dL_dpsiT_dev) return CH_EINVAL;
    ch_status st = adjoint_prepare(s);
    if(st != CH_OK) return st;
    ch_ctx *c = s->ctx;
    const size_t fb = s->bins * s->elt;
    HIPCHK(c, hipMemcpyAsync(s->lam[0], dL_dpsiT_dev, fb, hipMemcpyDeviceToDevice, c->stream));
    st = adjoint_core(s, grads);
    if(st != CH_OK) return st;
    if(dL_dpsi0_dev) HIPCHK(c, hipMemcpyAsync(dL_dpsi0_dev, s->lam[0], fb, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CH_OK;
}

extern "C" ch_status ch_der_bulk_free_energy(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev, void *out_dev,
                                             int64_t n) {
    if(!ctx || !desc || !rho_dev || !out_dev || n < 0) return CH_EINVAL;
    if(desc->model != CH_LANDAU && desc->model != CH_SIMPLE_WERTHEIM) return fail(ctx, CH_EINVAL, "bad model");
    if(desc->precision != CH_FP64 && desc->precision != CH_FP32) return fail(ctx, CH_EINVAL, "bad precision");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    chd::Params p = make_params(desc);
    HIPCHK(ctx, chd::pointwise_fprime(p, desc->precision, desc->model, rho_dev, out_dev, n, ctx->stream));
    return CH_OK;
}

extern "C" ch_status ch_set_mobility(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev, void *mob_dev, int64_t n) {
    if(!ctx || !desc || !rho_dev || !mob_dev || n < 0) return CH_EINVAL;
    if(desc->mobility < CH_MOB_CONSTANT || desc->mobility > CH_MOB_LINEAR) return fail(ctx, CH_EINVAL, "bad mobility");
    if(desc->precision != CH_FP64 && desc->precision != CH_FP32) return fail(ctx, CH_EINVAL, "bad precision");
    if(desc->mobility == CH_MOB_FREE_ENERGY && desc->model != CH_SIMPLE_WERTHEIM)
        return fail(ctx, CH_EUNSUPPORTED, "this model does not support mobility coupling");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    chd::Params p = make_params(desc);
    HIPCHK(ctx, chd::pointwise_mobility(p, desc->precision, desc->mobility, rho_dev, mob_dev, n, ctx->stream));
    return CH_OK;
}
