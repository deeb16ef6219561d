/*
 * ch_hip.h -- C ABI of the MI355X (gfx950) Cahn-Hilliard time-stepping library.
 *
 * The drop-in boundary for the reference's in-process plugin API:
 *   Integrator<dims>          src/integrators/Integrator.h:10-41
 *   CUDAIntegrator<dims>      src/CUDA/integrators/CUDAIntegrator.{h,cu}
 *   FreeEnergyModel (device)  src/models/FreeEnergyModel.h:29,48
 *   IMobility<dims>           src/mobility/IMobility.h:17-30
 *   CahnHilliard observables  src/CahnHilliard.cpp:285-319
 * Plain C types only: pointers, sizes, status codes.  Nothing here calls
 * exit(); every failure is a ch_status plus a message from ch_ctx_last_error.
 *
 * Layout of every field: one species, idx = x + N*y + N*N*z (x fastest), the
 * reference's MultiField layout (src/utils/MultiField.h:62).  Host arrays are
 * always double; device buffers are double (CH_FP64) or float (CH_FP32).
 *
 * Threading: calls on one ch_ctx are serialised by the caller; work is queued
 * on the context's stream (default priority).  Distinct contexts are
 * independent.
 */
#ifndef CH_HIP_H
#define CH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CH_ABI_VERSION 9

typedef enum ch_status {
    CH_OK = 0,
    CH_EINVAL = 1,        /* bad argument or descriptor                         */
    CH_EHIP = 2,          /* HIP runtime error (message in ch_ctx_last_error)   */
    CH_ENAN = 3,          /* NaN met in a reduction (CahnHilliard.cpp:290-293)  */
    CH_EUNSUPPORTED = 4,  /* valid request the library does not implement       */
    CH_ENOMEM = 5,        /* device allocation failed                           */
    CH_ESTATE = 6         /* call out of order (e.g. adjoint before record)     */
} ch_status;

enum { CH_FP64 = 0, CH_FP32 = 1 };
/* `integrator` key: src/CahnHilliard.cpp:169-220 */
enum { CH_EULER = 0, CH_EULER_MOBILITY = 1, CH_PSEUDOSPECTRAL = 2, CH_PSEUDOSPECTRAL_MOBILITY = 3 };
/* `free_energy` key: src/main.cpp:44-62 */
enum { CH_LANDAU = 0, CH_SIMPLE_WERTHEIM = 1, CH_SALEH_WERTHEIM = 2, CH_RICCI_WERTHEIM = 3, CH_GENERIC_WERTHEIM = 4 };
/* `mobility.type` key: src/mobility/IMobility.cpp:18-57; LINEAR (M rho / rho_ref) is new */
enum { CH_MOB_CONSTANT = 0, CH_MOB_FREE_ENERGY = 1, CH_MOB_REGULARISED = 2, CH_MOB_LINEAR = 3 };

/* One system.  All values in internal units (the caller applies
 * distance_scaling_factor as src/integrators/Integrator.cpp:11-28 does). */
typedef struct ch_desc {
    int32_t dims;          /* 1, 2 or 3                                           */
    int32_t N;             /* cells per side, a power of two >= 4                 */
    int32_t precision;     /* CH_FP64 or CH_FP32 (field arithmetic)              */
    int32_t integrator;    /* CH_EULER, CH_EULER_MOBILITY, CH_PSEUDOSPECTRAL,    
                              CH_PSEUDOSPECTRAL_MOBILITY                         */
    int32_t model;         /* CH_LANDAU, CH_SIMPLE_WERTHEIM                      */
    int32_t mobility;      /* CH_MOB_*                                            */
    int32_t valence;       /* wertheim.valence                                    */
    int32_t use_dealias;   /* pseudospectral.use_dealias                          */
    double epsilon;        /* landau.epsilon (or 1 - landau.T)                    */
    double B2;             /* wertheim.B2                                         */
    double delta;          /* wertheim.delta (already resolved by Delta)          */
    double regularisation_delta; /* wertheim.regularisation_delta                 */
    double M;              /* mobility.M                                          */
    double rho_min;        /* mobility.rho_min (regularised)                      */
    double rho_ref;        /* mobility.rho_ref (linear)                           */
    double dt, dx, k;      /* dt, dx, k (the step uses mu = f' - 2 k lap)         */
    double S;              /* pseudospectral.S                                    */
    /* conserved noise of euler_mobility (EulerMobilityCPU.cpp:10-18, 65-68):
     * each face flux gains sqrt(max(M_face, 0)) sqrt(2 / (dx^dims dt))
     * noise_rescale_factor eta with eta a standard normal.  eta comes from a
     * counter-based generator, Philox4x32-10 keyed by `seed` with the counter
     * (cell index, step index), so a run is reproducible for any launch shape;
     * the reference's mt19937 / curand streams are not reproduced.  Ignored
     * by the other integrators, as in the reference. */
    int32_t with_noise;            /* mobility.with_noise (0 / 1)                 */
    int32_t reserved0;             /* 0                                           */
    double noise_rescale_factor;   /* mobility.noise_rescale_factor (default 1)   */
    uint64_t seed;                 /* seed                                        */
    /* pseudospectral_mobility (PseudospectralMobilityCPU.cpp:203-208): 0 =
     * mobility splitting, 1 = restarted GMRES; 0 in the other three fields
     * selects the reference's defaults (30, 200, 1e-10).  FP32 systems solve
     * to max(gmres_tol, 1e-6). */
    int32_t use_gmres;             /* pseudospectral.use_gmres                    */
    int32_t gmres_restart;         /* pseudospectral.gmres_restart                */
    int32_t gmres_max_iter;        /* pseudospectral.gmres_max_iter               */
    int32_t reserved1;             /* 0                                           */
    double gmres_tol;              /* pseudospectral.gmres_tol                    */
    /* distance_scaling_factor^-3: the Wertheim models' ideal term is
     * rho (ln(rho * density_conversion) - 1) (FreeEnergyModel.h:20-21); 0 means 1 */
    double density_conversion;
} ch_desc;

/* ---- multi-species models (src/models/SalehWertheim.cpp, RicciWertheim.cpp,
 * GenericWertheim.cpp).  Fields of such a system are species-major: species s
 * occupies elements [s N^dims, (s + 1) N^dims) of every host buffer. */
#define CH_MAX_SPECIES 8
#define CH_MAX_PATCHES 16
typedef struct ch_species_model {
    int32_t model;                 /* CH_SALEH_WERTHEIM (3 species), CH_RICCI_WERTHEIM (2), CH_GENERIC_WERTHEIM */
    int32_t n_species;             /* 3, 2, or the number of generic_wertheim.species (<= CH_MAX_SPECIES) */
    double k[CH_MAX_SPECIES];      /* the `k` array of the integrator, per species (after unit scaling)   */
    /* saleh.*: valence[3], B2, B3, delta_AA, delta_BB (internal units) */
    int32_t saleh_valence[3];
    int32_t reserved0;
    double saleh_B2, saleh_B3, saleh_delta_AA, saleh_delta_BB;
    /* ricci.*: B2, delta_00, delta_12 */
    double ricci_B2, ricci_delta_00, ricci_delta_12;
    /* generic_wertheim.*: patch_count[s][p] = how many patches of type p species s carries
     * (the `patches` lists), n_patches = largest patch type + 1, delta[p][q] and B2[s][t]
     * symmetric (the `deltas` and `B2s` tables) */
    int32_t n_patches;
    int32_t reserved1;
    int32_t patch_count[CH_MAX_SPECIES][CH_MAX_PATCHES];
    double delta[CH_MAX_PATCHES][CH_MAX_PATCHES];
    double B2[CH_MAX_SPECIES][CH_MAX_SPECIES];
} ch_species_model;

typedef struct ch_ctx ch_ctx;
typedef struct ch_sys ch_sys;
typedef struct ch_batch ch_batch;

int ch_abi_version(void);
const char *ch_status_string(ch_status s);

/* device + stream; replaces the implicit CUDA context and legacy default
 * stream of src/CUDA/integrators/CUDAIntegrator.cu:13-33 */
ch_status ch_ctx_create(int device, ch_ctx **out);
/* every system and batch created on ctx must be destroyed first */
void ch_ctx_destroy(ch_ctx *ctx);
const char *ch_ctx_last_error(const ch_ctx *ctx);
ch_status ch_ctx_synchronize(ch_ctx *ctx);
/* the hipStream_t work is queued on (for interop, e.g. events) */
void *ch_ctx_stream(ch_ctx *ctx);

/* == CUDAIntegrator ctor + EulerCUDA / EulerMobilityCUDA / PseudospectralCUDA
 * ctor (device buffers, constants, FFT set-up): CUDAIntegrator.cu:13-33,
 * EulerCUDA.cu:37-55, PseudospectralCUDA.cu:190-230 */
ch_status ch_sys_create(ch_ctx *ctx, const ch_desc *desc, ch_sys **out);
void ch_sys_destroy(ch_sys *sys);
/* == the same for a multi-species model (CahnHilliard.cpp:169-228 with
 * N_species() > 1): desc->model equals model->model; the euler (constant
 * mobility), euler_mobility (constant, regularised, linear; free_energy with
 * CH_GENERIC_WERTHEIM, GenericWertheim.cpp:276-291) and pseudospectral
 * integrators.  Upload / download / observables use n_species N^dims values;
 * reverse mode, noise, S(q) and batches are single-species only. */
ch_status ch_sys_create_species(ch_ctx *ctx, const ch_desc *desc, const ch_species_model *model, ch_sys **out);
/* number of species of a system (1 for the single-species models) */
int32_t ch_sys_species(const ch_sys *sys);

/* host -> device copy of rho (N^dims doubles); == CUDAIntegrator::_CPU_GPU
 * (CUDAIntegrator.cu:43-53).  For the pseudospectral integrator it also sets
 * rho_hat = r2c(rho) (PseudospectralCPU.cpp:98-101). */
ch_status ch_sys_upload(ch_sys *sys, const double *rho_host);
/* device -> host copy of rho; == Integrator::sync() (CUDAIntegrator.cu:55-70) */
ch_status ch_sys_download(ch_sys *sys, double *rho_host);
/* device-to-device variants; `rho_dev` holds N^dims elements of the system's
 * precision (double for CH_FP64, float for CH_FP32) */
ch_status ch_sys_upload_device(ch_sys *sys, const void *rho_dev);
ch_status ch_sys_download_device(ch_sys *sys, void *rho_dev);
/* the mobility field used by the last step (N^dims doubles); the host copy
 * CUDAIntegrator::sync() refreshes (CUDAIntegrator.cu:61-66) */
ch_status ch_sys_download_mobility(ch_sys *sys, double *mob_host);

/* nsteps x { IMobility::update_mobility(); Integrator::evolve(); }
 * == CahnHilliard::evolve (src/CahnHilliard.cpp:279-282) called nsteps times.
 * Asynchronous on the context stream. */
ch_status ch_sys_evolve(ch_sys *sys, int64_t nsteps);

/* Time steps one stencil launch advances (temporal blocking depth; 1 for
 * the pseudospectral integrator).  No reference counterpart: introspection for
 * measurement (the reference always does one step per launch,
 * EulerCUDA.cu:60-69).  Override with the CH_TB_STEPS environment variable. */
int32_t ch_sys_steps_per_launch(const ch_sys *sys);
/* Region shape index (0-8) of the 2D register/DPP stencil kernel, -1 for the
 * systems that do not use it (1D, 3D, N < 64, spectral, species, noise).  No
 * reference counterpart: introspection for measurement.  Override with the
 * CH_WAVE_CFG environment variable. */
int32_t ch_sys_region_shape(const ch_sys *sys);
/* GMRES iterations of the last pseudospectral_mobility step with use_gmres
 * (0 otherwise); the reference computes the count and discards it
 * (PseudospectralMobilityCPU.cpp:375-376). */
int32_t ch_sys_gmres_iterations(const ch_sys *sys);

/* FP64 reductions of the current field: average_mass and average_free_energy
 * (src/CahnHilliard.cpp:285-319, V_bin = dx^3 as :163).  Synchronous.
 * Returns CH_ENAN when either sum is NaN. */
ch_status ch_sys_observables(ch_sys *sys, double *mass, double *free_energy);

/* Queued observables (no reference counterpart).  ch_sys_observables_queue
 * enqueues the reductions of ch_sys_observables on the current field behind
 * the work already issued and returns at once; ch_sys_observables_fetch waits
 * for the oldest queued pair and returns it (first in, first out), with
 * CH_ENAN as ch_sys_observables.  A driver that prints every few steps
 * (src/main.cpp:120-128) queues at each print point and fetches one interval
 * later, so the device does not wait for the host between intervals; the
 * values and their order are those of the synchronous call.  At most
 * CH_OBS_QUEUE pairs are pending: queueing beyond that, or fetching with none
 * pending, returns CH_EINVAL. */
#define CH_OBS_QUEUE 8
ch_status ch_sys_observables_queue(ch_sys *sys);
ch_status ch_sys_observables_fetch(ch_sys *sys, double *mass, double *free_energy);
int32_t ch_sys_observables_pending(const ch_sys *sys);
/* ch_sys_evolve(sys, nsteps) followed by ch_sys_observables_queue(sys): one
 * print interval of the reference's run loop (src/main.cpp:120-128, the
 * reductions of CahnHilliard.cpp:285-319 after CahnHilliard::evolve).  For the
 * stencil integrators a count that recurs from the same buffer becomes one
 * captured graph of the steps and the reductions (no launch gap between
 * them); same values as the two calls.  No reference counterpart. */
ch_status ch_sys_evolve_observe(ch_sys *sys, int64_t nsteps);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The reductions of ch_sys_observables (same kernels, same
 * bits) on a field the caller holds on the device: `rho_dev` is N^dims
 * elements of desc->precision, desc a single-species model (CH_EUNSUPPORTED
 * otherwise).  No allocation: the partial sums live in one fixed array of
 * 16,400 bytes per device, set up by ch_ctx_create; fields of more than 2^31
 * cells are refused (CH_EINVAL).  Calls on one device share that array, so
 * calls that overlap on one device race.  Synchronous; CH_ENAN as
 * ch_sys_observables. */
ch_status ch_observables_device_noalloc(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev, double *mass,
                                        double *free_energy);

/* ---- structure factor (2D): the circular average of utils/sq_from_traj.py:72-103
 * (mean removed, |FFT|^2 / N^2, shell = rint(|k| / dq)), computed with the
 * library's own FFT.  nshells must equal ch_sq_nshells(N) = rint(N sqrt(2) / 2) + 1. */
int32_t ch_sq_nshells(int32_t N);
ch_status ch_sys_structure_factor(ch_sys *sys, double *Sq, int32_t nshells);
/* loss = sum_s (S_s - target_s)^2 of the current field and, when grad is not
 * NULL, dloss/dpsi (N^2 doubles): the seed of ch_sys_adjoint for the
 * structure-factor loss.  Synchronous. */
ch_status ch_sys_sq_loss(ch_sys *sys, const double *target, int32_t nshells, double *loss, double *grad);

/* ---- batches of independent replicas (no reference counterpart: the
 * reference runs one system per process, src/main.cpp:182-200).  Every
 * replica is a 2D euler system with N >= 64 and its own N, epsilon (or
 * Wertheim parameters), k, dt, dx and M; all share one model and precision.
 * One launch advances every replica by ch_batch_steps_per_launch steps. */
ch_status ch_batch_create(ch_ctx *ctx, int32_t n, const ch_desc *descs, ch_batch **out);
/* How an FP32 batch rounds its updates psi + dt M lap(mu).
 *   CH_COMP_NONE    every step (the single-system FP32 arithmetic: each
 *                   replica gets the bits of the same system run alone)
 *   CH_COMP_LAUNCH  once per launch: a launch sums each cell's increments in
 *                   registers and rounds psi once at its store (no extra
 *                   memory traffic); the FP32 default of ch_batch_create.
 *                   Results depend on the launch depth, not on the region split.
 *   CH_COMP_STORED  every update is a compensated sum whose lost low-order
 *                   parts are kept in a second FP32 field (twice the traffic);
 *                   ch_batch_upload clears a replica's parts and
 *                   ch_batch_download returns psi without them, so a download
 *                   followed by an upload is not exact in this mode.
 *   CH_COMP_LOW8    CH_COMP_LAUNCH over a field that keeps an 8-bit low-order
 *                   part beside each FP32 value (ulp / 256 resolution: about
 *                   32 significant bits, 2 more bytes per cell and launch): a
 *                   launch starts from psi + its low-order part and stores the
 *                   remainder of its final rounding there, so no rounding
 *                   accumulates over launches.  Upload clears the parts.
 * FP64 batches take CH_COMP_NONE only (CH_EINVAL otherwise). */
enum { CH_COMP_NONE = 0, CH_COMP_LAUNCH = 1, CH_COMP_STORED = 2, CH_COMP_LOW8 = 3 };
ch_status ch_batch_create_ex(ch_ctx *ctx, int32_t n, const ch_desc *descs, int32_t compensation, ch_batch **out);
int32_t ch_batch_compensation(const ch_batch *batch);
void ch_batch_destroy(ch_batch *batch);
ch_status ch_batch_upload(ch_batch *batch, int32_t replica, const double *rho_host);
ch_status ch_batch_download(ch_batch *batch, int32_t replica, double *rho_host);
/* nsteps x CahnHilliard::evolve for every replica (asynchronous) */
ch_status ch_batch_evolve(ch_batch *batch, int64_t nsteps);
/* FP64 average mass and free energy of every replica (arrays of n); synchronous */
ch_status ch_batch_observables(ch_batch *batch, double *mass, double *free_energy);
int32_t ch_batch_steps_per_launch(const ch_batch *batch);
/* Autotuner (no reference counterpart): times every region split of the
 * batch kernel (rows per wave x waves over 64-row regions: shape 0: 16 x 4,
 * 1: 8 x 8, 2: 4 x 16 with 64 columns; 3: 4 x 16, 4: 8 x 8 with 128 columns,
 * two per lane; 5: 4 x 16, 6: 8 x 8 with 192 columns, three per lane; 7: 5 x 16
 * over 192 columns and 80 rows; 8: regions fitted per replica size -- whole
 * N = 64 and 128 replicas, N = 256 in regions of all 256 columns, others as 7)
 * at 2, 4
 * and 8 steps per launch (4 and 8 for a compensated batch, whose accuracy
 * depends on the depth) on the batch's current fields and keeps the fastest
 * per time step; the fields are not changed.  Runs by itself on the first
 * ch_batch_evolve unless the environment pins the choice (CH_BATCH_TB,
 * CH_BATCH_SHAPE) or a batch of the same composition (device, precision,
 * model, compensation, replicas per N) was tuned before in this process or
 * is listed in the file named by CH_TUNE_CACHE, to which each tuning appends
 * its result.  Synchronous; either output may be NULL. */
ch_status ch_batch_autotune(ch_batch *batch, int32_t *shape, int32_t *steps_per_launch);

/* ---- reverse mode (no reference counterpart: the reference has no
 * gradients).  The euler integrator (constant mobility) and the
 * euler_mobility integrator (every mobility rule), both models.
 *
 * Parameter gradients come in arrays of CH_NGRADS doubles, indexed by
 * CH_GRAD_*: epsilon (Landau), k, M (mobility.M), delta and B2
 * (SimpleWertheim; wertheim.delta as resolved into ch_desc.delta).  Entries a
 * model does not have are 0. */
enum { CH_GRAD_EPSILON = 0, CH_GRAD_K = 1, CH_GRAD_M = 2, CH_GRAD_DELTA = 3, CH_GRAD_B2 = 4, CH_NGRADS = 5 };
/* ch_sys_record: advance the current state nsteps times (as ch_sys_evolve)
 * keeping a checkpoint every `ckpt_every` steps; ckpt_every = 0 keeps every
 * state (no recomputation).  The system ends at psi_T. */
ch_status ch_sys_record(ch_sys *sys, int64_t nsteps, int32_t ckpt_every);
/* Reverse sweep over the last record: given dL/dpsi_T (N^dims doubles on the
 * host), writes dL/dpsi_0 (may be NULL) and grads[CH_NGRADS] (may be NULL).
 * Windows between checkpoints are recomputed on the device.  Synchronous. */
ch_status ch_sys_adjoint(ch_sys *sys, const double *dL_dpsiT, double *dL_dpsi0, double *grads);
/* Same with device buffers of the system's precision. */
ch_status ch_sys_adjoint_device(ch_sys *sys, const void *dL_dpsiT_dev, void *dL_dpsi0_dev, double *grads);
/* The free-energy loss: F = average_free_energy of the current field (as
 * ch_sys_observables, FP64), dF/dpsi (N^dims doubles; may be NULL) -- the
 * seed of ch_sys_adjoint for a loss on F -- and the explicit partial
 * derivatives dF/dparams[CH_NGRADS] (may be NULL).  Synchronous. */
ch_status ch_sys_free_energy_grad(ch_sys *sys, double *F, double *dF_dpsi, double *dF_dparams);

/* == FreeEnergyModel::der_bulk_free_energy(field_type*, float*, int) and
 * FreeEnergyModel::set_mobility(field_type*, double, field_type*, int)
 * (src/models/FreeEnergyModel.h:29,48; Landau.cu:5-19, SimpleWertheim.cu:10-42):
 * pointwise kernels on caller-owned device buffers of n elements of
 * desc->precision.  set_mobility evaluates desc->mobility (not only
 * free_energy), so every IMobility type has a device form. */
ch_status ch_der_bulk_free_energy(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev, void *out_dev, int64_t n);
ch_status ch_set_mobility(ch_ctx *ctx, const ch_desc *desc, const void *rho_dev, void *mob_dev, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* CH_HIP_H */
