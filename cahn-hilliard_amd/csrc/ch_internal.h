// ch_internal.h -- host-side declarations shared by the library's .hip files.
#pragma once

#include <vector>

#include "ch_device.h"

namespace chd {

// K fused Euler / EulerMobility steps in one launch, in -> out (distinct
// buffers).  K = 1 always exists; K in {2, 4, 8} for 2D grids with N >= 64.
template <int DIMS>
hipError_t stencil_step(const Params &p, int precision, int integrator, int model, int mob, int K, const void *in,
                        void *out, hipStream_t s);
// the 3D euler / euler_mobility step as two launches (mu of every cell into
// `work`, N^3 elements, and a variable mobility into `mob_work`, N^3 elements,
// when split3d_needs_mobility; then the update): Params::tile3d == 3; not for
// free-energy mobility (split3d_ok)
bool split3d_ok(const Params &p, int integrator, int mob);
bool split3d_needs_mobility(int integrator, int mob);
hipError_t stencil_split3d(const Params &p, int precision, int integrator, int model, int mob, const void *in,
                           void *out, void *work, void *mob_work, hipStream_t s);
// the first launch alone: mu (and the mobilities when split3d_needs_mobility)
// of `in`
hipError_t split3d_mu(const Params &p, int precision, int integrator, int model, int mob, const void *in, void *mu,
                      void *mob_out, hipStream_t s);
// the 3D euler_mobility reverse step as three launches (Params::tile3d == 3,
// mobility not free_energy): lambda_in -> lambda_out with psi = psi_n; work =
// three N^3 fields (mu, mobility, nu; work[3] unused), the first two already
// holding mu_n and the mobilities of psi_n when have_mu (then two launches); gradient
// partials into the first adjoint_split3d_blocks(p) slot groups of gacc
bool adjoint_split3d_ok(const Params &p, int mob);
int adjoint_split3d_blocks(const Params &p);
hipError_t adjoint_split3d(const Params &p, int precision, int model, int mob, const void *lam_in, void *lam_out,
                           const void *psi, void *const work[4], bool have_mu, double *gacc, hipStream_t s);
inline int stencil_max_k(int dims, int N) { return (dims == 2 && N >= 64) ? 8 : 1; }
// one euler_mobility step with conserved noise (noise.hip), in -> out; the
// step index is *step_base + step_off.  noise_step_advance adds n to *step_base.
hipError_t em_noise_step(const Params &p, int precision, int model, int mob, const void *in, void *out,
                         const int64_t *step_base, int step_off, hipStream_t s);
hipError_t noise_step_advance(int64_t *step_base, int n, hipStream_t s);

// pseudospectral pipeline
struct SpectralArgs {
    Params p;
    int precision;
    int model;
    void *rho;            // N^d real field (T)
    void *work;           // half spectrum scratch (T2)
    void *rho_hat;        // half spectrum state (T2)
    const void *twiddles; // W_N^m, m = 0..N-1 (T2)
    // rho_hat stored line-major (entry j of column line l at l N + j) where the
    // 2D column pass runs the two-wave kernel (N = 1024, 2048): set for the
    // pseudospectral integrator, whose chain alone reads rho_hat
    bool rh_lines = false;
};
// init = true: rho_hat = r2c(rho); otherwise one time step
hipError_t ps_run(const SpectralArgs &a, hipStream_t s, bool init);
// nsteps time steps, consecutive steps joined by the fused c2r + f' + r2c row
// pass (same results as nsteps calls of ps_run)
hipError_t ps_steps(const SpectralArgs &a, hipStream_t s, int64_t nsteps);
// the same chain from the real state a.rho and spectral state a.rho_hat, with
// state t (1 <= t <= nsteps) written to field t of `tape` (a.rho untouched)
hipError_t ps_steps_tape(const SpectralArgs &a, hipStream_t s, int64_t nsteps, void *tape);
// whole-field transforms (any dims, unnormalised): r2c of in (of f'(in) when
// fprime) into spec; c2r of spec (overwritten) into out
hipError_t spec_r2c(const SpectralArgs &a, const void *real_in, void *spec_out, bool fprime, hipStream_t s);
hipError_t spec_c2r(const SpectralArgs &a, void *spec, void *real_out, hipStream_t s);
// pseudospectral_mobility: divh = sum_d (i k_d) deal r2c(w c2r((i k_d) deal src) / N^d),
// w = M - M0 (sub, M0 = *M0p) or M, with the element-wise work inside the
// transform passes; hipErrorNotSupported outside 2D N = 1024, 2048
hipError_t psm_div_fused(const SpectralArgs &a, const void *src, void *gh, void *divh, const void *mob,
                         const double *M0p, int sub, hipStream_t s);
// pseudospectral reverse step (ps_adjoint.hip): lam_in = dL/drho_{n+1} ->
// lam_out = dL/drho_n with rho_n = `rho`; spec[3] half spectra and real[2]
// fields of scratch; kPsGradBlocks x kGradSlots partial sums added to gacc
constexpr int kPsGradBlocks = 256;
hipError_t ps_adjoint_step(const SpectralArgs &a, const void *lam_in, void *lam_out, const void *rho, void *const spec[3],
                           void *const real[2], double *gacc, hipStream_t s);
// 2D transforms on a.work: forward r2c of a.rho, inverse (c2r) into out
hipError_t fft2_forward(const SpectralArgs &a, hipStream_t s);
hipError_t fft2_inverse(const SpectralArgs &a, void *out, hipStream_t s);
// structure factor shells of a.work and the loss-gradient scaling
hipError_t sq_shells(int precision, const void *W, int N, int nshells, double *sq, double *cnt, hipStream_t s);
hipError_t sq_scale(int precision, void *W, int N, const double *sq, const double *target, const double *cnt,
                    hipStream_t s);

// pseudospectral_mobility (psm.hip): buffers of one system
constexpr int kPsmParts = 1024;
struct PsmArgs {
    SpectralArgs a;               // rho = the real field, rho_hat = the spectral state, twiddles
    int mob_rule;                 // CH_MOB_*
    void *fh, *muh, *gh, *divh;   // half spectra (T2)
    void *g, *mob, *tr;           // fields (T)
    double *partials;             // kPsmParts reduction partials
    double *scal;                 // scal[0] = M0 = max(0, max M)
    // full scheme (GMRES)
    int restart, max_iter;
    double tol;
    void *V;                      // restart + 1 Krylov vectors (T)
    void *w, *b, *x, *rhs, *r, *av;
    double *Hcol, *ydev;          // restart + 2 doubles each (device)
    double *host_partials;        // pinned host memory, kPsmParts doubles
    bool fuse;                    // divergence term through psm_div_fused where it applies
};
// one splitting step, asynchronous (capturable)
hipError_t psm_step_simple(const PsmArgs &B, hipStream_t s);
// one GMRES step, synchronous; iters = GMRES iterations
hipError_t psm_step_full(const PsmArgs &B, hipStream_t s, int &iters);

// multi-species models (species.hip)
constexpr int kSpMaxSpecies = CH_MAX_SPECIES, kSpMaxPatches = CH_MAX_PATCHES;
constexpr int kSpMaxTerms = kSpMaxPatches * kSpMaxSpecies * kSpMaxPatches;
struct SpModel {
    int model, S, P;               // CH_*_WERTHEIM, species, patch types (Generic)
    double k[kSpMaxSpecies];
    int val[3], lh;                // Saleh valences, linker half valence
    double sB2, sB3, dAA, dBB;     // Saleh
    double rB2, d00, d12;          // Ricci
    // Generic: unique patch types ascending, their mass-action terms
    int n_unique, unique[kSpMaxPatches], term_start[kSpMaxPatches + 1];
    int t_species[kSpMaxTerms], t_other[kSpMaxTerms];
    double t_coeff[kSpMaxTerms];
    int pc[kSpMaxSpecies][kSpMaxPatches];  // patch multiplicities per species
    double B2[kSpMaxSpecies][kSpMaxSpecies];
};
void sp_build_model(const ch_species_model &in, SpModel &m);
struct SpArgs {
    Params p;
    int precision, mob_rule;
    const SpModel *m;       // device copy
    const SpModel *m_host;  // host copy
    void *rho;              // S * bins (T), species-major
    double *X;              // P * bins, patch-major (Generic)
    void *mu, *mob;         // S * bins (T) each
    void *rho_hat;          // S * hat (T2), pseudospectral state
    void *spec[2];          // hat (T2) scratch
    int64_t bins, hat;
};
hipError_t sp_warm(const SpArgs &a, hipStream_t s);
hipError_t sp_step_stencil(const SpArgs &a, int integrator, hipStream_t s);
hipError_t sp_step_spectral(const SpArgs &a, const SpectralArgs &sa, hipStream_t s);
hipError_t sp_init_spectral(const SpArgs &a, const SpectralArgs &sa, hipStream_t s);
hipError_t sp_mobility_field(const SpArgs &a, double *out, hipStream_t s);
// ring != nullptr: the pair goes to the ring slot of *count (observables_ring)
hipError_t sp_observables(const SpArgs &a, double *partials, int nparts, double *result, hipStream_t s,
                          double *ring = nullptr, unsigned *count = nullptr);
hipError_t finalize_pairs(double *partials, int nparts, double *result, hipStream_t s);

// FP64 observables: partials (2 nparts doubles) -> result[0] = sum rho,
// result[1] = sum f-density.  The _ring forms write the pair to slot
// (*count % CH_OBS_QUEUE) of `ring` and increment *count on the device.
hipError_t observables(const Params &p, int precision, int model, const void *rho, double *partials, int nparts,
                       double *result, hipStream_t s);
hipError_t observables_ring(const Params &p, int precision, int model, const void *rho, double *partials, int nparts,
                            double *ring, unsigned *count, hipStream_t s);
hipError_t finalize_pairs_ring(double *partials, int nparts, double *ring, unsigned *count, hipStream_t s);
constexpr int kObsParts = 1024;
// allocation-free observables: the device address of the fixed scratch array
// (2 kObsParts + 2 doubles: partials, result), looked up per device by
// noalloc_scratch_init
constexpr int kMaxDevices = 64;
constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 31;
hipError_t noalloc_scratch_init(int device);
double *noalloc_scratch(int device);

// batched replicas (batch.hip)
struct BatchRep {
    Params p;
    int64_t off;   // element offset of the replica inside each ping-pong buffer
    int whole;     // 1: N == 64, one periodic item
};
// batch regions (shape): 64 rows high and 64, 128 or 192 columns wide
// (shapes 0-6), or 80 rows and 192 columns (shape 7); a whole N = 64 replica
// is one 64 x 64 item.  The items of a K-step launch cover each replica with
// the interiors of one family (region size) of halo regions.
constexpr int kBatchRegion = 64;
constexpr int kBatchShapes = 9;
constexpr int kBatchFamilies = 5;
int batch_family(int shape);         // 0: 64 x 64, 1: 128 x 64, 2: 192 x 64, 3: 192 x 80, 4: fitted per N (shape 8)
int batch_family_shape(int family);  // a shape of the family (its region size)
int batch_tile_x(int shape, int K);  // interior of a K-step halo item
int batch_tile_y(int shape, int K);
std::vector<int> batch_items(int shape, int N, int K);  // tx | ty << 16 per item of one replica
hipError_t batch_launch(int precision, int model, int upd, int shape, int K, const void *a, void *b,
                        const BatchRep *reps, const int2 *items, int nitems, const void *ca, void *cb, hipStream_t s);
// per-replica FP64 observables: result[2 r] = sum rho, result[2 r + 1] = sum f-density
hipError_t batch_observables(int precision, int model, const void *buf, const BatchRep *reps, int n, double *result,
                             hipStream_t s);

// adjoint (adjoint.hip).  `em`: the euler_mobility integrator (any mobility
// rule); otherwise euler with constant mobility.  Every adjoint tile owns
// kGradSlots FP64 accumulators.
constexpr int kGradSlots = 4;
// depth (steps) of the next adjoint / recompute launch with `remaining` steps left
int adjoint_depth(int dims, int N, bool em, int64_t remaining);
int tape_depth(int dims, int N, bool em, int64_t remaining);
int adjoint_tiles(const Params &p, bool em);
hipError_t adjoint_launch(const Params &p, int precision, int model, int mob, bool em, int KA, const void *lam_in,
                          void *lam_out, const void *tape, int s_top, double *gacc, hipStream_t s);
// recompute: K states from `in` into K consecutive tape fields (K from tape_depth)
hipError_t tape_launch(const Params &p, int precision, int model, int mob, bool em, int K, const void *in,
                       void *tape_out, hipStream_t s);
hipError_t fill_zero(double *a, int n, hipStream_t s);
// free-energy gradient: grad[i] = V_bin / bins * mu_i (double), and per-block
// partial sums partials[kFeSlots * b + j] of {f, sum_d (grad_d+ rho)^2,
// d f / d eps, d f / d delta, rho^2}; returns the number of blocks in *nb
constexpr int kFeSlots = 5;
hipError_t free_energy_grad(const Params &p, int precision, int model, const void *rho, double *grad,
                            double *partials, int max_parts, int *nb, hipStream_t s);

hipError_t pointwise_fprime(const Params &p, int precision, int model, const void *in, void *out, int64_t n,
                            hipStream_t s);
hipError_t pointwise_mobility(const Params &p, int precision, int mob, const void *in, void *out, int64_t n,
                              hipStream_t s);
// double <-> float conversions on device
hipError_t convert_d2f(const double *in, float *out, int64_t n, hipStream_t s);
hipError_t convert_f2d(const float *in, double *out, int64_t n, hipStream_t s);

}  // namespace chd
