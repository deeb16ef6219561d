/*
 * ch_oracle.h -- FP64 CPU restatement of the cahn-hilliard time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Layout everywhere: one species (Landau and SimpleWertheim are single-species
 * models, src/models/Landau.cpp:40-42, src/models/SimpleWertheim.h:21-23),
 * idx = x + N*y + N*N*z (src/CahnHilliard.cpp:245-253, MultiField.h:62).
 */
#ifndef CH_ORACLE_H
#define CH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OCH_LANDAU = 0, OCH_SIMPLE_WERTHEIM = 1 };
enum { OCH_MOB_CONSTANT = 0, OCH_MOB_FREE_ENERGY = 1, OCH_MOB_REGULARISED = 2, OCH_MOB_LINEAR = 3 };

typedef struct och_model {
    int model;                 /* OCH_LANDAU / OCH_SIMPLE_WERTHEIM */
    double epsilon;            /* landau.epsilon                        */
    double B2;                 /* wertheim.B2                           */
    double delta;              /* wertheim.delta                        */
    double regularisation_delta;
    int valence;               /* wertheim.valence                      */
} och_model;

typedef struct och_params {
    int dims, N;
    double dt, dx, k;          /* k: the TOML "k" (the step uses 2k)    */
    int mobility;              /* OCH_MOB_*                             */
    double M;                  /* mobility.M                            */
    double rho_min;            /* mobility.rho_min (regularised)        */
    double rho_ref;            /* mobility.rho_ref (linear)             */
    double S;                  /* pseudospectral.S                      */
    int use_dealias;           /* pseudospectral.use_dealias            */
} och_params;

/* ---- POSIX drand48 (the reference seeds it at src/main.cpp:36) ---- */
typedef struct och_rng { uint64_t x; } och_rng;
void och_srand48(och_rng *r, int64_t seed);
double och_drand48(och_rng *r);

/* initial field, src/CahnHilliard.cpp:133-151 */
void och_init_density(int dims, int N, double density, double A, int N_peaks, int64_t seed, double *rho);

/* ---- models ---- */
double och_bulk_free_energy(const och_model *m, double rho);
double och_der_bulk_free_energy(const och_model *m, double rho);
double och_second_der_bulk_free_energy(const och_model *m, double rho);
double och_wertheim_X(const och_model *m, double rho);
double och_pressure(const och_model *m, double rho);
void och_der_bulk_free_energy_field(const och_model *m, const double *rho, double *out, int64_t n);

/* mobility field from rho, src/mobility/*.h */
void och_update_mobility(const och_params *p, const och_model *m, const double *rho, double *mob, int64_t n);

/* ---- integrators: one step, in place ---- */
void och_euler_step(const och_params *p, const och_model *m, double *rho);
void och_euler_mobility_step(const och_params *p, const och_model *m, double *rho);

/* pseudospectral: opaque state carrying rho_hat (src/integrators/PseudospectralCPU.h:45-53) */
typedef struct och_ps och_ps;
och_ps *och_ps_create(const och_params *p, const och_model *m, const double *rho);
void och_ps_step(och_ps *s, double *rho);
void och_ps_destroy(och_ps *s);
int64_t och_ps_hat_size(const och_ps *s);
void och_ps_get_hat(const och_ps *s, double *re_im);

/* n steps of the selected integrator (0 euler, 1 euler_mobility, 2 pseudospectral) */
int och_run(int integrator, const och_params *p, const och_model *m, double *rho, int64_t steps);

/* ---- observables, src/CahnHilliard.cpp:285-319 ---- */
double och_average_mass(const och_params *p, const double *rho);
double och_average_free_energy(const och_params *p, const och_model *m, const double *rho);

/* ---- FFT (FP64, radix-2, exact-index twiddles) ---- */
void och_fft_c2c(int n, int inverse, double *re_im, int stride);
void och_dft_brute(int n, int inverse, const double *in_re_im, double *out_re_im);

/* ---- S(q), utils/sq_from_traj.py:72-103, 2D L x L ---- */
int och_sq_nshells(int L);
void och_radial_sq(int L, double dx, const double *field, double *Sq);
double och_sq_loss(int L, double dx, const double *field, const double *Sq_target, double *grad_field);

/* ---- FP64 adjoint of the Landau Euler step (constant mobility) ----
 * forward: psi' = psi + dt*M*Lap(mu), mu = f'(psi) - 2k Lap(psi)
 * given lambda' = dLoss/dpsi', returns lambda = dLoss/dpsi in place and
 * accumulates dLoss/d(eps, k, M) into grad[0..2]. */
void och_euler_adjoint_step(const och_params *p, const och_model *m, const double *psi, double *lambda, double *grad);
/* full forward + reverse sweep with every state stored: loss = <w, psi_T> */
void och_euler_adjoint(const och_params *p, const och_model *m, const double *psi0, int64_t steps,
                       const double *dL_dpsiT, double *dL_dpsi0, double *grad);

/* ---- parameter derivatives and the adjoint of both stencil integrators ----
 * grad[] order: epsilon (Landau), k, M (the mobility.M scale), delta and B2
 * (SimpleWertheim); entries a model does not have stay 0. */
enum { OCH_G_EPSILON = 0, OCH_G_K = 1, OCH_G_M = 2, OCH_G_DELTA = 3, OCH_G_B2 = 4, OCH_NGRADS = 5 };
/* d f'/d delta and d f/d delta (SimpleWertheim; 0 for Landau) */
double och_der_bulk_free_energy_ddelta(const och_model *m, double rho);
double och_bulk_free_energy_ddelta(const och_model *m, double rho);
/* d M/d rho and d M/d delta of the mobility rule (och_update_mobility) */
double och_mobility_drho(const och_params *p, const och_model *m, double rho);
double och_mobility_ddelta(const och_params *p, const och_model *m, double rho);
/* integrator 0 (euler, constant mobility) or 1 (euler_mobility, any
 * mobility rule): lambda' -> lambda in place, grad[OCH_NGRADS] accumulated */
void och_adjoint_step(int integrator, const och_params *p, const och_model *m, const double *psi, double *lambda,
                      double *grad);
void och_adjoint(int integrator, const och_params *p, const och_model *m, const double *psi0, int64_t steps,
                 const double *dL_dpsiT, double *dL_dpsi0, double *grad);
/* adjoint of the pseudospectral step on the real field (k and M through the
 * spectral multipliers; S is held fixed), and the full forward + reverse sweep */
void och_ps_adjoint_step(const och_params *p, const och_model *m, const double *rho, double *lambda, double *grad);
void och_ps_adjoint(const och_params *p, const och_model *m, const double *rho0, int64_t steps,
                    const double *dL_drhoT, double *dL_drho0, double *grad);
/* ---- conserved noise of euler_mobility (counter-based Philox4x32-10) ---- */
void och_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void och_cell_normals(uint64_t seed, int64_t cell, int64_t step, double nrm[4]);
/* one euler_mobility step with noise; `step` is the step's index since the
 * system was created */
void och_euler_mobility_noise_step(const och_params *p, const och_model *m, double noise_rescale, uint64_t seed,
                                   int64_t step, double *rho);
/* ---- pseudospectral_mobility (src/integrators/PseudospectralMobilityCPU.cpp) ----
 * use_gmres = 0: _evolve_simple (mobility splitting, M0 = max M); 1:
 * _evolve_full (restarted GMRES with the FFT-diagonal preconditioner).  The
 * state carries rho_hat; och_psm_step returns the GMRES iterations. */
typedef struct och_psm och_psm;
och_psm *och_psm_create(const och_params *p, const och_model *m, const double *rho, int use_gmres, int restart,
                        int max_iter, double tol);
int och_psm_step(och_psm *s, double *rho);
double och_psm_last_M0(const och_psm *s);
void och_psm_destroy(och_psm *s);
/* the spectral div(mob grad g) of the integrator (c2r semantics included) */
void och_psm_div_mob_grad(const och_params *p, const double *mob, const double *g, double *out);
/* ---- multi-species Wertheim models (src/models/SalehWertheim.cpp,
 * RicciWertheim.cpp, GenericWertheim.cpp) and the stencil / pseudospectral
 * integrators over species-major fields rho[s * bins + i].  The struct has
 * the layout of ch_species_model. */
enum { OCH_SALEH = 2, OCH_RICCI = 3, OCH_GENERIC = 4 };
#define OCH_MAX_SPECIES 8
#define OCH_MAX_PATCHES 16
typedef struct och_species_model {
    int32_t model, n_species;
    double k[OCH_MAX_SPECIES];
    int32_t saleh_valence[3], reserved0;
    double saleh_B2, saleh_B3, saleh_delta_AA, saleh_delta_BB;
    double ricci_B2, ricci_delta_00, ricci_delta_12;
    int32_t n_patches, reserved1;
    int32_t patch_count[OCH_MAX_SPECIES][OCH_MAX_PATCHES];
    double delta[OCH_MAX_PATCHES][OCH_MAX_PATCHES];
    double B2[OCH_MAX_SPECIES][OCH_MAX_SPECIES];
} och_species_model;
/* Generic: the mass-action fixed point, warm-started from X (n_patches values,
 * in/out), GenericWertheim.cpp:184-208 */
void och_ms_update_X(const och_species_model *m, const double *rhos, double *X);
/* f'_s for every species at one cell (X in/out for Generic); a species whose
 * density is exactly 0 gets 0 */
void och_ms_der(const och_species_model *m, const double *rhos, double *X, double *out);
/* the bulk free-energy density at one cell (Generic: X from 0, as bonding_free_energy) */
double och_ms_bulk(const och_species_model *m, double density_conversion, const double *rhos);
typedef struct och_ms och_ms;
/* integrator 0 euler, 1 euler_mobility, 2 pseudospectral; p->k is unused (m->k per species) */
och_ms *och_ms_create(int integrator, const och_params *p, const och_species_model *m, const double *rho);
void och_ms_step(och_ms *s, double *rho);
void och_ms_observables(const och_ms *s, double density_conversion, const double *rho, double *mass, double *fe);
const double *och_ms_X(const och_ms *s);
void och_ms_destroy(och_ms *s);
/* F = average_free_energy(rho) with dF/drho (may be NULL) and the explicit
 * dF/dparams[OCH_NGRADS] (may be NULL) */
double och_free_energy_grad(const och_params *p, const och_model *m, const double *rho, double *dF_drho,
                            double *dF_dparams);

#ifdef __cplusplus
}
#endif
#endif
