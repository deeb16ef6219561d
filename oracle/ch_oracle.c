/*
 * ch_oracle.c -- FP64 CPU restatement of the cahn-hilliard time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.  Loops follow the reference's loop order so that the
 * single-core timing is representative of the reference CPU integrators.
 *
 * Pins (what checks each function; the tests live in tests/test_oracle_*.py):
 *   drand48 ............ glibc drand48 (python ctypes) + known first values
 *   init ............... reference fixture header/first values (6 digits)
 *   euler, euler_mob ... oracle/_ref golden vectors (reference sources, 1e-12),
 *                        reference .dat fixtures (5e-6), Fourier eigenmode,
 *                        brute-force dense operator on tiny grids
 *   pseudospectral ..... brute DFT, per-mode closed form, reference .dat
 *                        fixtures (5e-6); FFTW itself is unavailable here
 *   observables ........ closed forms (uniform field, single cosine)
 *   S(q) ............... numpy.fft on random fields
 *   adjoint ............ central finite differences of och_run + dot-product test
 */
#include "ch_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQR(x) ((x) * (x))

static const double OCH_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* drand48: X_{n+1} = (0x5DEECE66D X_n + 0xB) mod 2^48, srand48 sets   */
/* X = (seed << 16) | 0x330E (POSIX).  Seeded at src/main.cpp:36.       */
/* ------------------------------------------------------------------ */
void och_srand48(och_rng *r, int64_t seed) {
    uint64_t s = (uint64_t)(uint32_t)seed;
    r->x = ((s << 16) | 0x330Eull) & 0xFFFFFFFFFFFFull;
}

double och_drand48(och_rng *r) {
    r->x = (0x5DEECE66Dull * r->x + 0xBull) & 0xFFFFFFFFFFFFull;
    return (double)r->x / 281474976710656.0; /* 2^48 */
}

static int64_t ipow(int N, int dims) {
    int64_t b = 1;
    for(int i = 0; i < dims; i++) b *= N;
    return b;
}

/* src/CahnHilliard.cpp:133-151 (one species, unit scaling factor 1) */
void och_init_density(int dims, int N, double density, double A, int N_peaks, int64_t seed, double *rho) {
    och_rng r;
    och_srand48(&r, seed);
    int64_t bins = ipow(N, dims);
    double initial_k = 2 * OCH_PI * N_peaks / (double)N;
    for(int64_t bin = 0; bin < bins; bin++) {
        double modulation = A * cos(initial_k * bin);
        double random_factor = (N_peaks == 0) ? (och_drand48(&r) - 0.5) : 1.0 + 0.02 * (och_drand48(&r) - 0.5);
        if(density != 0.0) {
            rho[bin] = density * (1.0 + 2.0 * modulation * random_factor);
        }
        else {
            rho[bin] = 2.0 * modulation * random_factor;
        }
    }
}

/* ------------------------------------------------------------------ */
/* models                                                              */
/* ------------------------------------------------------------------ */
static double two_valence_delta(const och_model *m) { return 2.0 * m->valence * m->delta; }

/* src/models/SimpleWertheim.cpp:88-90 */
double och_wertheim_X(const och_model *m, double rho) {
    double tvd = two_valence_delta(m);
    return (-1.0 + sqrt(1.0 + 2.0 * tvd * rho)) / (tvd * rho);
}

/* src/models/Landau.cpp:54-57, src/models/SimpleWertheim.cpp:71-86 */
double och_bulk_free_energy(const och_model *m, double rho) {
    if(m->model == OCH_LANDAU) {
        return -0.5 * m->epsilon * SQR(rho) + 0.25 * SQR(SQR(rho));
    }
    double reg = m->regularisation_delta;
    double log_delta = log(reg);
    double f_ref = m->B2 * SQR(rho);
    if(rho < reg) {
        double a = 1 / (2.0 * reg);
        double b = log_delta - 1.0;
        double c = reg * (log_delta - 1.0) - a * SQR(reg) - b * reg;
        f_ref += a * SQR(rho) + b * rho + c;
    }
    else {
        f_ref += rho * (log(rho) - 1.0);
    }
    double X = och_wertheim_X(m, rho);
    double f_bond = (rho > 0.) ? m->valence * rho * (log(X) + 0.5 * (1. - X)) : 0.0;
    return f_ref + f_bond;
}

/* src/models/Landau.cpp:46-52, src/models/SimpleWertheim.cpp:40-53 */
double och_der_bulk_free_energy(const och_model *m, double rho) {
    if(m->model == OCH_LANDAU) {
        return -m->epsilon * rho + rho * rho * rho;
    }
    double reg = m->regularisation_delta;
    double der_f_ref = 2 * m->B2 * rho;
    if(rho < reg) {
        der_f_ref += log(reg) + (rho - reg) / reg;
    }
    else {
        der_f_ref += log(rho);
    }
    double der_f_bond = (rho > 0.) ? m->valence * log(och_wertheim_X(m, rho)) : 0.0;
    return der_f_ref + der_f_bond;
}

/* d^2 f / d rho^2.  Landau: -eps + 3 psi^2.  SimpleWertheim (rho > reg):
 * 2 B2 + 1/rho + valence X'/X with X' from the mass-action law
 * c rho X^2 + X - 1 = 0 (c = tvd/2): X' = -c X^2 / (2 c rho X + 1). */
double och_second_der_bulk_free_energy(const och_model *m, double rho) {
    if(m->model == OCH_LANDAU) {
        return -m->epsilon + 3.0 * rho * rho;
    }
    double reg = m->regularisation_delta;
    double d = 2 * m->B2;
    d += (rho < reg) ? 1.0 / reg : 1.0 / rho;
    if(rho > 0.) {
        double c = 0.5 * two_valence_delta(m);
        double X = och_wertheim_X(m, rho);
        double Xp = -c * X * X / (2.0 * c * rho * X + 1.0);
        d += m->valence * Xp / X;
    }
    return d;
}

/* src/models/Landau.cpp:65-70 */
double och_pressure(const och_model *m, double rho) {
    return rho * och_der_bulk_free_energy(m, rho) - och_bulk_free_energy(m, rho);
}

void och_der_bulk_free_energy_field(const och_model *m, const double *rho, double *out, int64_t n) {
    for(int64_t i = 0; i < n; i++) out[i] = och_der_bulk_free_energy(m, rho[i]);
}

/* src/mobility/ConstantMobility.h:19-23, FreeEnergyMobility.h:25-34 with
 * SimpleWertheim.cpp:104-111, RegularisedMobility.h:26-33; "linear" is new:
 * M rho / rho_ref. */
void och_update_mobility(const och_params *p, const och_model *m, const double *rho, double *mob, int64_t n) {
    for(int64_t i = 0; i < n; i++) {
        switch(p->mobility) {
        case OCH_MOB_CONSTANT: mob[i] = p->M; break;
        case OCH_MOB_FREE_ENERGY: {
            double X = och_wertheim_X(m, rho[i]);
            double D = p->M * pow(X, (double)m->valence);
            mob[i] = rho[i] * D;
            break;
        }
        case OCH_MOB_REGULARISED: mob[i] = p->M * rho[i] / (rho[i] + p->rho_min); break;
        case OCH_MOB_LINEAR: mob[i] = p->M * rho[i] / p->rho_ref; break;
        default: mob[i] = NAN;
        }
    }
}

/* ------------------------------------------------------------------ */
/* grid helpers, src/integrators/EulerCPU.cpp:39-88                    */
/* ------------------------------------------------------------------ */
static int ilog2(int N) { int b = 0; while((1 << b) < N) b++; return b; }

static void fill_coords(int dims, int N, int64_t idx, int *c) {
    int bits = ilog2(N);
    for(int d = 0; d < dims; d++) { c[d] = (int)(idx & (N - 1)); idx >>= bits; }
}

static int64_t cell_idx(int dims, int N, const int *c) {
    int bits = ilog2(N);
    int64_t idx = 0, mul = 1;
    for(int d = 0; d < dims; d++) { idx += c[d] * mul; mul <<= bits; }
    return idx;
}

/* periodic (2d+1)-point Laplacian, neighbours summed as -x,+x,-y,+y,-z,+z */
static double cell_laplacian(int dims, int N, double dx, const double *f, int64_t idx) {
    if(dims == 1) {
        int64_t m = (idx - 1 + N) & (N - 1), pp = (idx + 1) & (N - 1);
        return (f[m] + f[pp] - 2.0 * f[idx]) / SQR(dx);
    }
    int c[3], cn[3];
    fill_coords(dims, N, idx, c);
    memcpy(cn, c, sizeof(c));
    double sum = 0.0;
    for(int d = 0; d < dims; d++) {
        cn[d] = (c[d] - 1 + N) & (N - 1);
        sum += f[cell_idx(dims, N, cn)];
        cn[d] = (c[d] + 1) & (N - 1);
        sum += f[cell_idx(dims, N, cn)];
        cn[d] = c[d];
    }
    return (sum - 2.0 * dims * f[idx]) / SQR(dx);
}

/* src/integrators/EulerCPU.cpp:18-36 */
void och_euler_step(const och_params *p, const och_model *m, double *rho) {
    int64_t bins = ipow(p->N, p->dims);
    double *rho_der = (double *)malloc(sizeof(double) * bins);
    for(int64_t i = 0; i < bins; i++) rho_der[i] = och_der_bulk_free_energy(m, rho[i]);
    for(int64_t i = 0; i < bins; i++) rho_der[i] -= 2 * p->k * cell_laplacian(p->dims, p->N, p->dx, rho, i);
    double M = p->M; /* constant mobility: mobility(0, 0) */
    for(int64_t i = 0; i < bins; i++) {
        double total_derivative = M * cell_laplacian(p->dims, p->N, p->dx, rho_der, i);
        rho[i] += total_derivative * p->dt;
    }
    free(rho_der);
}

/* src/integrators/EulerMobilityCPU.cpp:31-83 (no noise) with the mobility
 * refreshed from rho first, as CahnHilliard::evolve does (src/CahnHilliard.cpp:279-282) */
void och_euler_mobility_step(const och_params *p, const och_model *m, double *rho) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double *mob = (double *)malloc(sizeof(double) * bins);
    double *mu = (double *)malloc(sizeof(double) * bins);
    double *flux = (double *)malloc(sizeof(double) * bins * dims);
    och_update_mobility(p, m, rho, mob, bins);
    for(int64_t i = 0; i < bins; i++) mu[i] = och_der_bulk_free_energy(m, rho[i]);
    for(int64_t i = 0; i < bins; i++) mu[i] -= 2.0 * p->k * cell_laplacian(dims, N, p->dx, rho, i);
    for(int64_t i = 0; i < bins; i++) {
        int c[3], cp[3];
        fill_coords(dims, N, i, c);
        for(int d = 0; d < dims; d++) {
            memcpy(cp, c, sizeof(c));
            cp[d] = (c[d] + 1) & (N - 1);
            int64_t ip = cell_idx(dims, N, cp);
            double M_face = 0.5 * (mob[i] + mob[ip]);
            flux[i * dims + d] = M_face * (mu[ip] - mu[i]) / p->dx;
        }
    }
    for(int64_t i = 0; i < bins; i++) {
        /* src/integrators/EulerCPU.cpp:114-128 */
        int c[3], cm[3];
        fill_coords(dims, N, i, c);
        memcpy(cm, c, sizeof(c));
        double res = 0;
        for(int d = 0; d < dims; d++) {
            cm[d] = (c[d] - 1 + N) & (N - 1);
            int64_t im = cell_idx(dims, N, cm);
            res += (flux[i * dims + d] - flux[im * dims + d]) / p->dx;
            cm[d] = c[d];
        }
        rho[i] += res * p->dt;
    }
    free(mob); free(mu); free(flux);
}

/* ------------------------------------------------------------------ */
/* FFT: iterative radix-2, twiddles from exact integer indices         */
/* ------------------------------------------------------------------ */
static void twiddle(int n, int j, int inverse, double *c, double *s) {
    double ang = 2.0 * OCH_PI * (double)j / (double)n;
    *c = cos(ang);
    *s = inverse ? sin(ang) : -sin(ang);
}

/* unnormalised DFT, forward sign e^{-i}, inverse e^{+i} (FFTW convention) */
void och_fft_c2c(int n, int inverse, double *a, int stride) {
    /* gather into a contiguous buffer */
    double *b = (double *)malloc(sizeof(double) * 2 * n);
    for(int j = 0; j < n; j++) { b[2 * j] = a[2 * (int64_t)j * stride]; b[2 * j + 1] = a[2 * (int64_t)j * stride + 1]; }
    int bits = ilog2(n);
    for(int j = 0; j < n; j++) {
        int r = 0;
        for(int q = 0; q < bits; q++) if(j & (1 << q)) r |= 1 << (bits - 1 - q);
        if(r > j) {
            double t0 = b[2 * j], t1 = b[2 * j + 1];
            b[2 * j] = b[2 * r]; b[2 * j + 1] = b[2 * r + 1];
            b[2 * r] = t0; b[2 * r + 1] = t1;
        }
    }
    for(int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for(int i = 0; i < n; i += len) {
            for(int j = 0; j < half; j++) {
                double wr, wi;
                twiddle(len, j, inverse, &wr, &wi);
                double *u = b + 2 * (i + j), *v = b + 2 * (i + j + half);
                double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr; v[1] = u[1] - vi;
                u[0] += vr; u[1] += vi;
            }
        }
    }
    for(int j = 0; j < n; j++) { a[2 * (int64_t)j * stride] = b[2 * j]; a[2 * (int64_t)j * stride + 1] = b[2 * j + 1]; }
    free(b);
}

void och_dft_brute(int n, int inverse, const double *in, double *out) {
    for(int k = 0; k < n; k++) {
        double sr = 0, si = 0;
        for(int j = 0; j < n; j++) {
            double wr, wi;
            twiddle(n, (int)(((int64_t)j * k) % n), inverse, &wr, &wi);
            sr += in[2 * j] * wr - in[2 * j + 1] * wi;
            si += in[2 * j] * wi + in[2 * j + 1] * wr;
        }
        out[2 * k] = sr; out[2 * k + 1] = si;
    }
}

/* r2c over a dims-dimensional N^dims real array into N^(dims-1) x (N/2+1) complex */
static void r2c(int dims, int N, const double *in, double *out) {
    int Nh = N / 2 + 1;
    int64_t lines = ipow(N, dims - 1);
    double *line = (double *)malloc(sizeof(double) * 2 * N);
    for(int64_t l = 0; l < lines; l++) {
        for(int x = 0; x < N; x++) { line[2 * x] = in[l * N + x]; line[2 * x + 1] = 0.0; }
        och_fft_c2c(N, 0, line, 1);
        memcpy(out + 2 * l * Nh, line, sizeof(double) * 2 * Nh);
    }
    free(line);
    if(dims >= 2) { /* y axis: stride Nh */
        int64_t planes = ipow(N, dims - 2);
        for(int64_t pz = 0; pz < planes; pz++)
            for(int kx = 0; kx < Nh; kx++) och_fft_c2c(N, 0, out + 2 * (pz * N * Nh + kx), Nh);
    }
    if(dims == 3) { /* z axis: stride N*Nh */
        for(int y = 0; y < N; y++)
            for(int kx = 0; kx < Nh; kx++) och_fft_c2c(N, 0, out + 2 * ((int64_t)y * Nh + kx), N * Nh);
    }
}

/* c2r, unnormalised; imaginary parts of the kx = 0 and kx = N/2 entries of
 * each x-line are dropped (the Hermitian-completion semantics of a c2r). */
static void c2r(int dims, int N, double *in, double *out) {
    int Nh = N / 2 + 1;
    if(dims == 3) {
        for(int y = 0; y < N; y++)
            for(int kx = 0; kx < Nh; kx++) och_fft_c2c(N, 1, in + 2 * ((int64_t)y * Nh + kx), N * Nh);
    }
    if(dims >= 2) {
        int64_t planes = ipow(N, dims - 2);
        for(int64_t pz = 0; pz < planes; pz++)
            for(int kx = 0; kx < Nh; kx++) och_fft_c2c(N, 1, in + 2 * (pz * N * Nh + kx), Nh);
    }
    int64_t lines = ipow(N, dims - 1);
    double *line = (double *)malloc(sizeof(double) * 2 * N);
    for(int64_t l = 0; l < lines; l++) {
        const double *h = in + 2 * l * Nh;
        for(int k = 0; k < Nh; k++) { line[2 * k] = h[2 * k]; line[2 * k + 1] = h[2 * k + 1]; }
        line[1] = 0.0;
        line[2 * (N / 2) + 1] = 0.0;
        for(int k = 1; k < N / 2; k++) { line[2 * (N - k)] = h[2 * k]; line[2 * (N - k) + 1] = -h[2 * k + 1]; }
        och_fft_c2c(N, 1, line, 1);
        for(int x = 0; x < N; x++) out[l * N + x] = line[2 * x];
    }
    free(line);
}

/* ------------------------------------------------------------------ */
/* pseudospectral, src/integrators/PseudospectralCPU.cpp:13-136        */
/* ------------------------------------------------------------------ */
struct och_ps {
    och_params p;
    och_model m;
    int64_t bins, hat;
    double *rho_hat, *rho_hat_copy, *f_der, *f_der_hat, *k2, *dealiaser;
};

och_ps *och_ps_create(const och_params *p, const och_model *m, const double *rho) {
    och_ps *s = (och_ps *)calloc(1, sizeof(och_ps));
    s->p = *p; s->m = *m;
    int N = p->N, dims = p->dims, Nh = N / 2 + 1;
    s->bins = ipow(N, dims);
    s->hat = ipow(N, dims - 1) * Nh;
    s->rho_hat = (double *)malloc(sizeof(double) * 2 * s->hat);
    s->rho_hat_copy = (double *)malloc(sizeof(double) * 2 * s->hat);
    s->f_der_hat = (double *)malloc(sizeof(double) * 2 * s->hat);
    s->f_der = (double *)malloc(sizeof(double) * s->bins);
    s->k2 = (double *)malloc(sizeof(double) * s->hat);
    s->dealiaser = (double *)malloc(sizeof(double) * s->hat);
    double k_cut = N * OCH_PI / (N * p->dx) * 2.0 / 3.0;
    int64_t k_idx = 0;
    if(dims == 1) {
        for(int i = 0; i < Nh; i++) {
            double k = 2.0 * OCH_PI * i / (N * p->dx);
            s->dealiaser[k_idx] = (k <= k_cut) ? 1.0 : 0.0;
            s->k2[k_idx++] = SQR(k);
        }
    }
    else if(dims == 2) {
        for(int kx_idx = 0; kx_idx < N; kx_idx++) {
            int kx = (kx_idx < N / 2) ? kx_idx : N - kx_idx;
            for(int ky = 0; ky < Nh; ky++) {
                double k = 2.0 * OCH_PI * sqrt((double)(SQR(kx) + SQR(ky))) / (N * p->dx);
                s->dealiaser[k_idx] = (k < k_cut) ? 1.0 : 0.0;
                s->k2[k_idx++] = SQR(k);
            }
        }
    }
    else {
        for(int kx_idx = 0; kx_idx < N; kx_idx++) {
            int kx_i = (kx_idx < N / 2) ? kx_idx : (kx_idx - N);
            double kx = 2.0 * OCH_PI * kx_i / (N * p->dx);
            for(int ky_idx = 0; ky_idx < N; ky_idx++) {
                int ky_i = (ky_idx < N / 2) ? ky_idx : (ky_idx - N);
                double ky = 2.0 * OCH_PI * ky_i / (N * p->dx);
                for(int kz_idx = 0; kz_idx < Nh; kz_idx++) {
                    double kz = 2.0 * OCH_PI * kz_idx / (N * p->dx);
                    double k2 = SQR(kx) + SQR(ky) + SQR(kz);
                    double k = sqrt(k2);
                    s->dealiaser[k_idx] = (k <= k_cut) ? 1.0 : 0.0;
                    s->k2[k_idx++] = k2;
                }
            }
        }
    }
    r2c(dims, N, rho, s->rho_hat);
    memcpy(s->rho_hat_copy, s->rho_hat, sizeof(double) * 2 * s->hat);
    return s;
}

void och_ps_step(och_ps *s, double *rho) {
    const och_params *p = &s->p;
    double M = p->M;
    for(int64_t i = 0; i < s->bins; i++) s->f_der[i] = och_der_bulk_free_energy(&s->m, rho[i]);
    r2c(p->dims, p->N, s->f_der, s->f_der_hat);
    for(int64_t k = 0; k < s->hat; k++) {
        double *fh = s->f_der_hat + 2 * k, *rh = s->rho_hat + 2 * k, *rc = s->rho_hat_copy + 2 * k;
        if(p->use_dealias) { fh[0] *= s->dealiaser[k]; fh[1] *= s->dealiaser[k]; }
        double k2 = s->k2[k];
        double k4 = SQR(k2);
        double denom = 1.0 + p->dt * M * (p->S * k2 + 2.0 * p->k * k4);
        double re = (rh[0] - p->dt * M * k2 * (fh[0] - p->S * rh[0])) / denom;
        double im = (rh[1] - p->dt * M * k2 * (fh[1] - p->S * rh[1])) / denom;
        rh[0] = rc[0] = re;
        rh[1] = rc[1] = im;
        rc[0] /= s->bins;
        rc[1] /= s->bins;
    }
    c2r(p->dims, p->N, s->rho_hat_copy, rho);
}

int64_t och_ps_hat_size(const och_ps *s) { return s->hat; }
void och_ps_get_hat(const och_ps *s, double *re_im) { memcpy(re_im, s->rho_hat, sizeof(double) * 2 * s->hat); }

void och_ps_destroy(och_ps *s) {
    if(!s) return;
    free(s->rho_hat); free(s->rho_hat_copy); free(s->f_der_hat); free(s->f_der); free(s->k2); free(s->dealiaser);
    free(s);
}

int och_run(int integrator, const och_params *p, const och_model *m, double *rho, int64_t steps) {
    if(integrator == 0) {
        for(int64_t t = 0; t < steps; t++) och_euler_step(p, m, rho);
    }
    else if(integrator == 1) {
        for(int64_t t = 0; t < steps; t++) och_euler_mobility_step(p, m, rho);
    }
    else if(integrator == 2) {
        och_ps *s = och_ps_create(p, m, rho);
        for(int64_t t = 0; t < steps; t++) och_ps_step(s, rho);
        och_ps_destroy(s);
    }
    else return -1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* observables, src/CahnHilliard.cpp:256-319 (V_bin = dx^3, :163)      */
/* ------------------------------------------------------------------ */
double och_average_mass(const och_params *p, const double *rho) {
    int64_t bins = ipow(p->N, p->dims);
    double mass = 0.;
    for(int64_t i = 0; i < bins; i++) mass += rho[i];
    double V_bin = p->dx * p->dx * p->dx;
    return mass * V_bin / bins;
}

double och_average_free_energy(const och_params *p, const och_model *m, const double *rho) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double fe = 0.;
    for(int64_t i = 0; i < bins; i++) {
        int c[3], cp[3];
        fill_coords(dims, N, i, c);
        double interfacial = 0.;
        for(int d = 0; d < dims; d++) {
            memcpy(cp, c, sizeof(c));
            cp[d] = (c[d] + 1) & (N - 1);
            double g = (rho[cell_idx(dims, N, cp)] - rho[i]) / p->dx;
            interfacial += p->k * g * g;
        }
        fe += och_bulk_free_energy(m, rho[i]) + interfacial;
    }
    double V_bin = p->dx * p->dx * p->dx;
    return fe * V_bin / bins;
}

/* ------------------------------------------------------------------ */
/* S(q), utils/sq_from_traj.py:72-103                                  */
/* ------------------------------------------------------------------ */
static int fftfreq_index(int i, int L) { return (i < (L + 1) / 2) ? i : i - L; }

static int shell_of(int L, double dx, int a, int b) {
    double k = 2.0 * OCH_PI / (L * dx);
    double kx = k * fftfreq_index(a, L), ky = k * fftfreq_index(b, L);
    double K = sqrt(kx * kx + ky * ky);
    double dq = 2.0 * OCH_PI / (L * dx);
    return (int)rint(K / dq);
}

int och_sq_nshells(int L) {
    int mx = 0;
    for(int a = 0; a < L; a++)
        for(int b = 0; b < L; b++) { int s = shell_of(L, 1.0, a, b); if(s > mx) mx = s; }
    return mx + 1;
}

static void sq_spectrum(int L, const double *field, double *F) {
    int64_t n = (int64_t)L * L;
    double mean = 0;
    for(int64_t i = 0; i < n; i++) mean += field[i];
    mean /= n;
    for(int64_t i = 0; i < n; i++) { F[2 * i] = field[i] - mean; F[2 * i + 1] = 0; }
    for(int r = 0; r < L; r++) och_fft_c2c(L, 0, F + 2 * (int64_t)r * L, 1);
    for(int c = 0; c < L; c++) och_fft_c2c(L, 0, F + 2 * c, L);
}

void och_radial_sq(int L, double dx, const double *field, double *Sq) {
    int ns = och_sq_nshells(L);
    int64_t n = (int64_t)L * L;
    double *F = (double *)malloc(sizeof(double) * 2 * n);
    double *cnt = (double *)calloc(ns, sizeof(double));
    sq_spectrum(L, field, F);
    for(int s = 0; s < ns; s++) Sq[s] = 0;
    for(int a = 0; a < L; a++)
        for(int b = 0; b < L; b++) {
            int64_t i = (int64_t)a * L + b;
            int s = shell_of(L, dx, a, b);
            Sq[s] += (F[2 * i] * F[2 * i] + F[2 * i + 1] * F[2 * i + 1]) / n;
            cnt[s] += 1;
        }
    for(int s = 0; s < ns; s++) Sq[s] = cnt[s] > 0 ? Sq[s] / cnt[s] : 0.0;
    free(F); free(cnt);
}

/* loss = sum_s (S_s - T_s)^2 and its gradient with respect to the field */
double och_sq_loss(int L, double dx, const double *field, const double *T, double *grad) {
    int ns = och_sq_nshells(L);
    int64_t n = (int64_t)L * L;
    double *Sq = (double *)malloc(sizeof(double) * ns);
    double *cnt = (double *)calloc(ns, sizeof(double));
    double *F = (double *)malloc(sizeof(double) * 2 * n);
    och_radial_sq(L, dx, field, Sq);
    double loss = 0;
    for(int s = 0; s < ns; s++) loss += SQR(Sq[s] - T[s]);
    if(grad) {
        sq_spectrum(L, field, F);
        for(int a = 0; a < L; a++)
            for(int b = 0; b < L; b++) cnt[shell_of(L, dx, a, b)] += 1;
        /* G_k = 2 (S_s - T_s) / (c_s n); dL/dphi = 2 Re(sum_k G_k F_k e^{+ikx}) */
        for(int a = 0; a < L; a++)
            for(int b = 0; b < L; b++) {
                int64_t i = (int64_t)a * L + b;
                int s = shell_of(L, dx, a, b);
                double G = 2.0 * (Sq[s] - T[s]) / (cnt[s] * n);
                F[2 * i] *= G; F[2 * i + 1] *= G;
            }
        for(int r = 0; r < L; r++) och_fft_c2c(L, 1, F + 2 * (int64_t)r * L, 1);
        for(int c = 0; c < L; c++) och_fft_c2c(L, 1, F + 2 * c, L);
        double mean = 0;
        for(int64_t i = 0; i < n; i++) { grad[i] = 2.0 * F[2 * i]; mean += grad[i]; }
        mean /= n;
        for(int64_t i = 0; i < n; i++) grad[i] -= mean;
    }
    free(Sq); free(cnt); free(F);
    return loss;
}

/* ------------------------------------------------------------------ */
/* adjoint of the Landau Euler step                                    */
/* J^T l = l + dt M (f''(psi) (L l) - 2k L (L l))   (L symmetric)       */
/* ------------------------------------------------------------------ */
void och_euler_adjoint_step(const och_params *p, const och_model *m, const double *psi, double *lambda, double *grad) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double *w = (double *)malloc(sizeof(double) * bins);
    double *lp = (double *)malloc(sizeof(double) * bins);
    double *mu = (double *)malloc(sizeof(double) * bins);
    for(int64_t i = 0; i < bins; i++) w[i] = cell_laplacian(dims, N, p->dx, lambda, i);
    for(int64_t i = 0; i < bins; i++) lp[i] = cell_laplacian(dims, N, p->dx, psi, i);
    for(int64_t i = 0; i < bins; i++) mu[i] = och_der_bulk_free_energy(m, psi[i]) - 2.0 * p->k * lp[i];
    double ge = 0, gk = 0, gM = 0;
    for(int64_t i = 0; i < bins; i++) {
        ge += w[i] * psi[i];
        gk += w[i] * lp[i];
        gM += w[i] * mu[i];
    }
    grad[0] += -p->dt * p->M * ge;
    grad[1] += -2.0 * p->dt * p->M * gk;
    grad[2] += p->dt * gM;
    for(int64_t i = 0; i < bins; i++) {
        double Lw = cell_laplacian(dims, N, p->dx, w, i);
        lambda[i] += p->dt * p->M * (och_second_der_bulk_free_energy(m, psi[i]) * w[i] - 2.0 * p->k * Lw);
    }
    free(w); free(lp); free(mu);
}

void och_euler_adjoint(const och_params *p, const och_model *m, const double *psi0, int64_t steps,
                       const double *dL_dpsiT, double *dL_dpsi0, double *grad) {
    int64_t bins = ipow(p->N, p->dims);
    double *tape = (double *)malloc(sizeof(double) * bins * (steps + 1));
    memcpy(tape, psi0, sizeof(double) * bins);
    for(int64_t t = 0; t < steps; t++) {
        memcpy(tape + (t + 1) * bins, tape + t * bins, sizeof(double) * bins);
        och_euler_step(p, m, tape + (t + 1) * bins);
    }
    memcpy(dL_dpsi0, dL_dpsiT, sizeof(double) * bins);
    grad[0] = grad[1] = grad[2] = 0;
    for(int64_t t = steps - 1; t >= 0; t--) och_euler_adjoint_step(p, m, tape + t * bins, dL_dpsi0, grad);
    free(tape);
}

/* ------------------------------------------------------------------ */
/* parameter derivatives (SimpleWertheim.cpp:40-111, mobility/*.h)      */
/* With c = valence delta the bond fraction solves c rho X^2 + X - 1 = 0, */
/* so dX/ddelta = -valence rho X^2 / (2 c rho X + 1) and                */
/* dX/drho = -c X^2 / (2 c rho X + 1).                                  */
/* ------------------------------------------------------------------ */
static double wertheim_dX_ddelta(const och_model *m, double rho, double X) {
    double c = 0.5 * two_valence_delta(m);
    return -m->valence * rho * X * X / (2.0 * c * rho * X + 1.0);
}

static double wertheim_dX_drho(const och_model *m, double rho, double X) {
    double c = 0.5 * two_valence_delta(m);
    return -c * X * X / (2.0 * c * rho * X + 1.0);
}

double och_der_bulk_free_energy_ddelta(const och_model *m, double rho) {
    if(m->model == OCH_LANDAU || !(rho > 0.)) return 0.0;
    double X = och_wertheim_X(m, rho);
    return m->valence * wertheim_dX_ddelta(m, rho, X) / X;
}

double och_bulk_free_energy_ddelta(const och_model *m, double rho) {
    if(m->model == OCH_LANDAU || !(rho > 0.)) return 0.0;
    double X = och_wertheim_X(m, rho);
    return m->valence * rho * wertheim_dX_ddelta(m, rho, X) * (1.0 / X - 0.5);
}

double och_mobility_drho(const och_params *p, const och_model *m, double rho) {
    switch(p->mobility) {
    case OCH_MOB_FREE_ENERGY: {
        double X = och_wertheim_X(m, rho), v = m->valence;
        return p->M * (pow(X, v) + rho * v * pow(X, v - 1.0) * wertheim_dX_drho(m, rho, X));
    }
    case OCH_MOB_REGULARISED: return p->M * p->rho_min / SQR(rho + p->rho_min);
    case OCH_MOB_LINEAR: return p->M / p->rho_ref;
    default: return 0.0;
    }
}

double och_mobility_ddelta(const och_params *p, const och_model *m, double rho) {
    if(p->mobility != OCH_MOB_FREE_ENERGY) return 0.0;
    double X = och_wertheim_X(m, rho), v = m->valence;
    return p->M * rho * v * pow(X, v - 1.0) * wertheim_dX_ddelta(m, rho, X);
}

static int64_t neighbour(int dims, int N, int64_t i, int d, int s) {
    int c[3];
    fill_coords(dims, N, i, c);
    c[d] = (c[d] + s + N) & (N - 1);
    return cell_idx(dims, N, c);
}

/* Transpose of one step.  euler: psi' = psi + dt M L mu.  euler_mobility:
 * psi' = psi + dt D(Mob, mu) with D(a, u)_i = sum_d [A(i,i+e)(u_{i+e} - u_i)
 * - A(i-e,i)(u_i - u_{i-e})] / dx^2, A = (a_i + a_j) / 2 (the flux form of
 * EulerMobilityCPU.cpp:44-82), Mob = mobility(psi), mu = f'(psi) - 2k L psi.
 * <lambda', D(Mob, mu)> is bilinear in (Mob, mu), so with
 *   nu  = D(Mob, lambda')                      (d/dmu, D is symmetric in u)
 *   eta_j = -sum_d [(mu_{j+e}-mu_j)(l_{j+e}-l_j) + (mu_j-mu_{j-e})(l_j-l_{j-e})] / (2 dx^2)
 * (d/dMob) the transpose is lambda = lambda' + dt (f'' nu - 2k L nu + Mob' eta). */
void och_adjoint_step(int integrator, const och_params *p, const och_model *m, const double *psi, double *lambda,
                      double *grad) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double inv_dx2 = 1.0 / SQR(p->dx);
    double *nu = (double *)malloc(sizeof(double) * bins);
    double *eta = (double *)malloc(sizeof(double) * bins);
    double *mu = (double *)malloc(sizeof(double) * bins);
    double *lp = (double *)malloc(sizeof(double) * bins);
    double *mob = (double *)malloc(sizeof(double) * bins);
    for(int64_t i = 0; i < bins; i++) lp[i] = cell_laplacian(dims, N, p->dx, psi, i);
    for(int64_t i = 0; i < bins; i++) mu[i] = och_der_bulk_free_energy(m, psi[i]) - 2.0 * p->k * lp[i];
    if(integrator == 0) {
        /* nu = M L lambda'; d/dM of <lambda', dt M L mu> = dt <L lambda', mu> */
        for(int64_t i = 0; i < bins; i++) {
            double w = cell_laplacian(dims, N, p->dx, lambda, i);
            mob[i] = p->M;
            nu[i] = p->M * w;
            eta[i] = w * mu[i];
        }
    }
    else {
        och_update_mobility(p, m, psi, mob, bins);
        for(int64_t i = 0; i < bins; i++) {
            double sn = 0.0, se = 0.0;
            for(int d = 0; d < dims; d++) {
                int64_t ip = neighbour(dims, N, i, d, 1), im = neighbour(dims, N, i, d, -1);
                sn += 0.5 * (mob[i] + mob[ip]) * (lambda[ip] - lambda[i]) - 0.5 * (mob[im] + mob[i]) * (lambda[i] - lambda[im]);
                se += (mu[ip] - mu[i]) * (lambda[ip] - lambda[i]) + (mu[i] - mu[im]) * (lambda[i] - lambda[im]);
            }
            nu[i] = sn * inv_dx2;
            eta[i] = -0.5 * se * inv_dx2;
        }
    }
    double ge = 0, gk = 0, gm = 0, gd = 0, gb = 0;
    for(int64_t i = 0; i < bins; i++) {
        ge += nu[i] * psi[i];
        gk += nu[i] * lp[i];
        gb += nu[i] * 2.0 * psi[i];
        gd += nu[i] * och_der_bulk_free_energy_ddelta(m, psi[i]);
        if(integrator == 0) gm += eta[i];
        else {
            /* the mobility is linear in mobility.M: Mob = M * (Mob / M) */
            och_params p1 = *p;
            p1.M = 1.0;
            double mhat;
            och_update_mobility(&p1, m, psi + i, &mhat, 1);
            gm += eta[i] * mhat;
            gd += eta[i] * och_mobility_ddelta(p, m, psi[i]);
        }
    }
    if(m->model == OCH_LANDAU) grad[OCH_G_EPSILON] += -p->dt * ge;
    else {
        grad[OCH_G_DELTA] += p->dt * gd;
        grad[OCH_G_B2] += p->dt * gb;
    }
    grad[OCH_G_K] += -2.0 * p->dt * gk;
    grad[OCH_G_M] += p->dt * gm;
    for(int64_t i = 0; i < bins; i++) {
        double Lnu = cell_laplacian(dims, N, p->dx, nu, i);
        double v = och_second_der_bulk_free_energy(m, psi[i]) * nu[i] - 2.0 * p->k * Lnu;
        if(integrator == 1) v += och_mobility_drho(p, m, psi[i]) * eta[i];
        lambda[i] += p->dt * v;
    }
    free(nu); free(eta); free(mu); free(lp); free(mob);
}

void och_adjoint(int integrator, const och_params *p, const och_model *m, const double *psi0, int64_t steps,
                 const double *dL_dpsiT, double *dL_dpsi0, double *grad) {
    int64_t bins = ipow(p->N, p->dims);
    double *tape = (double *)malloc(sizeof(double) * bins * (steps + 1));
    memcpy(tape, psi0, sizeof(double) * bins);
    for(int64_t t = 0; t < steps; t++) {
        memcpy(tape + (t + 1) * bins, tape + t * bins, sizeof(double) * bins);
        if(integrator == 0) och_euler_step(p, m, tape + (t + 1) * bins);
        else och_euler_mobility_step(p, m, tape + (t + 1) * bins);
    }
    memcpy(dL_dpsi0, dL_dpsiT, sizeof(double) * bins);
    for(int g = 0; g < OCH_NGRADS; g++) grad[g] = 0.0;
    for(int64_t t = steps - 1; t >= 0; t--) och_adjoint_step(integrator, p, m, tape + t * bins, dL_dpsi0, grad);
    free(tape);
}

/* F = V_bin / bins sum_i [f(rho_i) + k sum_d ((rho_{i+e} - rho_i) / dx)^2]
 * (src/CahnHilliard.cpp:300-319).  dF/drho_i = V_bin / bins (f'(rho_i) - 2k L rho_i):
 * each forward difference appears in two cells' sums. */
double och_free_energy_grad(const och_params *p, const och_model *m, const double *rho, double *dF_drho,
                            double *dF_dparams) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double scale = p->dx * p->dx * p->dx / bins;
    double sg = 0, se = 0, sd = 0, sb = 0;
    for(int64_t i = 0; i < bins; i++) {
        for(int d = 0; d < dims; d++) {
            double g = (rho[neighbour(dims, N, i, d, 1)] - rho[i]) / p->dx;
            sg += g * g;
        }
        se += -0.5 * SQR(rho[i]);
        sd += och_bulk_free_energy_ddelta(m, rho[i]);
        sb += SQR(rho[i]);
        if(dF_drho)
            dF_drho[i] = scale * (och_der_bulk_free_energy(m, rho[i]) - 2.0 * p->k * cell_laplacian(dims, N, p->dx, rho, i));
    }
    if(dF_dparams) {
        for(int g = 0; g < OCH_NGRADS; g++) dF_dparams[g] = 0.0;
        dF_dparams[OCH_G_K] = scale * sg;
        if(m->model == OCH_LANDAU) dF_dparams[OCH_G_EPSILON] = scale * se;
        else {
            dF_dparams[OCH_G_DELTA] = scale * sd;
            dF_dparams[OCH_G_B2] = scale * sb;
        }
    }
    return och_average_free_energy(p, m, rho);
}

/* ------------------------------------------------------------------ */
/* adjoint of the pseudospectral step (PseudospectralCPU.cpp:112-136)  */
/* On the real field the step is rho' = Ainv(A F rho + B F f'(rho))    */
/* with per-mode A = (1 + dt M S q^2) / D, B = -dt M q^2 deal / D,      */
/* D = 1 + dt M (S q^2 + 2 k q^4): real multipliers even in q, so the   */
/* operators are symmetric and                                          */
/*   lambda = F^-1(A F lambda') + f''(rho) F^-1(B F lambda').           */
/* k and M enter through A and B: with rho_hat' = A rho_hat + B f_hat,  */
/*   d rho_hat'/dk = -2 dt M q^4 rho_hat' / D,                           */
/*   d rho_hat'/dM = (-dt q^2 (deal f_hat - S rho_hat)                  */
/*                    - dt (S q^2 + 2 k q^4) rho_hat') / D,              */
/* and <lambda', F^-1 X> = (1/bins) sum over the full spectrum of       */
/* conj(lambda_hat') X (half spectrum: weight 2 off the kx = 0, N/2     */
/* planes).                                                             */
/* ------------------------------------------------------------------ */
void och_ps_adjoint_step(const och_params *p, const och_model *m, const double *rho, double *lambda, double *grad) {
    int dims = p->dims, N = p->N, Nh = N / 2 + 1;
    int64_t bins = ipow(N, dims), hat = ipow(N, dims - 1) * Nh;
    och_ps *tab = och_ps_create(p, m, rho);  /* q^2 and dealias tables, rho_hat = r2c(rho) */
    double *lh = (double *)malloc(sizeof(double) * 2 * hat);
    double *fh = (double *)malloc(sizeof(double) * 2 * hat);
    double *ah = (double *)malloc(sizeof(double) * 2 * hat);
    double *bh = (double *)malloc(sizeof(double) * 2 * hat);
    double *f = (double *)malloc(sizeof(double) * bins);
    double *a = (double *)malloc(sizeof(double) * bins);
    double *b = (double *)malloc(sizeof(double) * bins);
    r2c(dims, N, lambda, lh);
    for(int64_t i = 0; i < bins; i++) f[i] = och_der_bulk_free_energy(m, rho[i]);
    r2c(dims, N, f, fh);
    const double dt = p->dt, M = p->M, S = p->S;
    double gk = 0, gM = 0;
    for(int64_t q = 0; q < hat; q++) {
        const double k2 = tab->k2[q], k4 = k2 * k2;
        const double deal = p->use_dealias ? tab->dealiaser[q] : 1.0;
        const double D = 1.0 + dt * M * (S * k2 + 2.0 * p->k * k4);
        const double A = (1.0 + dt * M * S * k2) / D, B = -dt * M * k2 * deal / D;
        const double *r = tab->rho_hat + 2 * q, *fq = fh + 2 * q, *l = lh + 2 * q;
        const double nr = A * r[0] + B * fq[0], ni = A * r[1] + B * fq[1];  /* rho_hat' */
        const int kx = (int)(q % Nh);
        const double w = ((kx == 0 || kx == N / 2) ? 1.0 : 2.0) / bins;
        const double dk = -2.0 * dt * M * k4 / D;
        gk += w * dk * (l[0] * nr + l[1] * ni);
        const double xr = (-dt * k2 * (deal * fq[0] - S * r[0]) - dt * (S * k2 + 2.0 * p->k * k4) * nr) / D;
        const double xi = (-dt * k2 * (deal * fq[1] - S * r[1]) - dt * (S * k2 + 2.0 * p->k * k4) * ni) / D;
        gM += w * (l[0] * xr + l[1] * xi);
        ah[2 * q] = A * l[0] / bins; ah[2 * q + 1] = A * l[1] / bins;
        bh[2 * q] = B * l[0] / bins; bh[2 * q + 1] = B * l[1] / bins;
    }
    c2r(dims, N, ah, a);
    c2r(dims, N, bh, b);
    double ge = 0, gd = 0;
    for(int64_t i = 0; i < bins; i++) {
        ge += b[i] * rho[i];
        gd += b[i] * och_der_bulk_free_energy_ddelta(m, rho[i]);
        lambda[i] = a[i] + och_second_der_bulk_free_energy(m, rho[i]) * b[i];
    }
    if(m->model == OCH_LANDAU) grad[OCH_G_EPSILON] += -ge;
    else {
        grad[OCH_G_DELTA] += gd;
        grad[OCH_G_B2] += 2.0 * ge;
    }
    grad[OCH_G_K] += gk;
    grad[OCH_G_M] += gM;
    och_ps_destroy(tab);
    free(lh); free(fh); free(ah); free(bh); free(f); free(a); free(b);
}

void och_ps_adjoint(const och_params *p, const och_model *m, const double *rho0, int64_t steps,
                    const double *dL_drhoT, double *dL_drho0, double *grad) {
    int64_t bins = ipow(p->N, p->dims);
    double *tape = (double *)malloc(sizeof(double) * bins * (steps + 1));
    memcpy(tape, rho0, sizeof(double) * bins);
    och_ps *s = och_ps_create(p, m, rho0);
    for(int64_t t = 0; t < steps; t++) {
        memcpy(tape + (t + 1) * bins, tape + t * bins, sizeof(double) * bins);
        och_ps_step(s, tape + (t + 1) * bins);
    }
    och_ps_destroy(s);
    memcpy(dL_drho0, dL_drhoT, sizeof(double) * bins);
    for(int g = 0; g < OCH_NGRADS; g++) grad[g] = 0.0;
    for(int64_t t = steps - 1; t >= 0; t--) och_ps_adjoint_step(p, m, tape + t * bins, dL_drho0, grad);
    free(tape);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:  */
/* as easy as 1, 2, 3", SC 2011): 10 rounds of                          */
/*   (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,                         */
/*   c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),                       */
/* with the key bumped by the Weyl constants (W0, W1) between rounds.   */
/* ------------------------------------------------------------------ */
void och_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for(int r = 0; r < 10; r++) {
        if(r > 0) { k0 += W0; k1 += W1; }
        const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The four standard normals of cell `cell` at step `step`: Philox of the
 * counter (cell lo, cell hi, step lo, step hi) under the key (seed lo, seed
 * hi), uniforms u = (x + 1/2) 2^-32 in (0, 1), Box-Muller on (u0, u1) and
 * (u2, u3).  Normal d drives the stochastic flux of the cell's +e_d face. */
void och_cell_normals(uint64_t seed, int64_t cell, int64_t step, double nrm[4]) {
    const uint32_t ctr[4] = {(uint32_t)cell, (uint32_t)((uint64_t)cell >> 32), (uint32_t)step,
                             (uint32_t)((uint64_t)step >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    och_philox4x32_10(ctr, key, x);
    for(int h = 0; h < 2; h++) {
        const double u1 = ((double)x[2 * h] + 0.5) * (1.0 / 4294967296.0);
        const double u2 = ((double)x[2 * h + 1] + 0.5) * (1.0 / 4294967296.0);
        const double r = sqrt(-2.0 * log(u1)), th = 2.0 * OCH_PI * u2;
        nrm[2 * h] = r * cos(th);
        nrm[2 * h + 1] = r * sin(th);
    }
}

/* euler_mobility with conserved noise (src/integrators/EulerMobilityCPU.cpp:
 * 10-18, 55-83): the +e_d face flux of cell i gains
 * sqrt(max(M_face, 0)) * sqrt(2 / (dx^d dt)) * rescale * eta_d(i, step),
 * and rho += dt (div J_det + div J_sto). */
void och_euler_mobility_noise_step(const och_params *p, const och_model *m, double noise_rescale, uint64_t seed,
                                   int64_t step, double *rho) {
    int dims = p->dims, N = p->N;
    int64_t bins = ipow(N, dims);
    double *mob = (double *)malloc(sizeof(double) * bins);
    double *mu = (double *)malloc(sizeof(double) * bins);
    double *flux = (double *)malloc(sizeof(double) * bins * dims);
    double *sto = (double *)malloc(sizeof(double) * bins * dims);
    const double nf = sqrt(2.0 / (pow(p->dx, dims) * p->dt)) * noise_rescale;
    och_update_mobility(p, m, rho, mob, bins);
    for(int64_t i = 0; i < bins; i++) mu[i] = och_der_bulk_free_energy(m, rho[i]) - 2.0 * p->k * cell_laplacian(dims, N, p->dx, rho, i);
    for(int64_t i = 0; i < bins; i++) {
        double nrm[4];
        och_cell_normals(seed, i, step, nrm);
        for(int d = 0; d < dims; d++) {
            int64_t ip = neighbour(dims, N, i, d, 1);
            double M_face = 0.5 * (mob[i] + mob[ip]);
            flux[i * dims + d] = M_face * (mu[ip] - mu[i]) / p->dx;
            sto[i * dims + d] = sqrt(M_face > 0.0 ? M_face : 0.0) * nf * nrm[d];
        }
    }
    for(int64_t i = 0; i < bins; i++) {
        double det = 0, st = 0;
        for(int d = 0; d < dims; d++) {
            int64_t im = neighbour(dims, N, i, d, -1);
            det += (flux[i * dims + d] - flux[im * dims + d]) / p->dx;
            st += (sto[i * dims + d] - sto[im * dims + d]) / p->dx;
        }
        rho[i] += (det + st) * p->dt;
    }
    free(mob); free(mu); free(flux); free(sto);
}

/* ------------------------------------------------------------------ */
/* pseudospectral_mobility, src/integrators/PseudospectralMobilityCPU.cpp */
/* ------------------------------------------------------------------ */
struct och_psm {
    och_params p;
    och_model m;
    int use_gmres, restart, max_iter, last_iters;
    double tol, M0;
    int64_t bins, hat;
    double *rho_hat;                   /* state, 2 * hat                          */
    double *k2, *deal, *kc[3];         /* per half-spectrum entry; kc[d] in the   */
                                       /* reference's axis order (slowest first)  */
    double *fh, *muh, *gh, *divh, *th; /* spectra, 2 * hat each                   */
    double *mob, *g, *tr;              /* real fields                             */
};

/* wave-vector tables, PseudospectralMobilityCPU.cpp:56-125: full axes use the
 * signed index (i < N/2 ? i : i - N), the x (half) axis 0 .. N/2; the
 * dealiaser keeps |k| <= k_cut in every dimension */
och_psm *och_psm_create(const och_params *p, const och_model *m, const double *rho, int use_gmres, int restart,
                        int max_iter, double tol) {
    och_psm *s = (och_psm *)calloc(1, sizeof(och_psm));
    s->p = *p; s->m = *m;
    s->use_gmres = use_gmres;
    s->restart = restart; s->max_iter = max_iter; s->tol = tol;
    int N = p->N, dims = p->dims, Nh = N / 2 + 1;
    s->bins = ipow(N, dims);
    s->hat = ipow(N, dims - 1) * Nh;
    size_t hb = sizeof(double) * 2 * s->hat, fb = sizeof(double) * s->bins;
    s->rho_hat = (double *)malloc(hb);
    s->fh = (double *)malloc(hb); s->muh = (double *)malloc(hb); s->gh = (double *)malloc(hb);
    s->divh = (double *)malloc(hb); s->th = (double *)malloc(hb);
    s->mob = (double *)malloc(fb); s->g = (double *)malloc(fb); s->tr = (double *)malloc(fb);
    s->k2 = (double *)malloc(sizeof(double) * s->hat);
    s->deal = (double *)malloc(sizeof(double) * s->hat);
    for(int d = 0; d < dims; d++) s->kc[d] = (double *)malloc(sizeof(double) * s->hat);
    const double k_cut = N * OCH_PI / (N * p->dx) * 2.0 / 3.0;
    for(int64_t q = 0; q < s->hat; q++) {
        int ix = (int)(q % Nh);
        int64_t rest = q / Nh;
        int iy = dims == 3 ? (int)(rest % N) : (int)rest, iz = dims == 3 ? (int)(rest / N) : 0;
        int idx[3], nd = 0;  /* slowest axis first, the half axis last */
        if(dims == 3) idx[nd++] = iz;
        if(dims >= 2) idx[nd++] = iy;
        double k2 = 0;
        for(int d = 0; d < dims; d++) {
            double kd;
            if(d == dims - 1) kd = 2.0 * OCH_PI * ix / (N * p->dx);
            else {
                int ki = (idx[d] < N / 2) ? idx[d] : (idx[d] - N);
                kd = 2.0 * OCH_PI * ki / (N * p->dx);
            }
            s->kc[d][q] = kd;
            k2 += SQR(kd);
        }
        s->k2[q] = k2;
        s->deal[q] = (sqrt(k2) <= k_cut) ? 1.0 : 0.0;
    }
    r2c(dims, N, rho, s->rho_hat);
    return s;
}

void och_psm_destroy(och_psm *s) {
    if(!s) return;
    free(s->rho_hat); free(s->fh); free(s->muh); free(s->gh); free(s->divh); free(s->th);
    free(s->mob); free(s->g); free(s->tr); free(s->k2); free(s->deal);
    for(int d = 0; d < s->p.dims; d++) free(s->kc[d]);
    free(s);
}

/* (i k) z */
static inline void ik_mul(double k, const double *z, double *out) {
    double re = -k * z[1], im = k * z[0];
    out[0] = re; out[1] = im;
}

/* M0 = max(0, max_i M_i), PseudospectralMobilityCPU.cpp:406-414 */
static double psm_M0(const och_psm *s) {
    double M0 = 0.0;
    for(int64_t i = 0; i < s->bins; i++) M0 = s->mob[i] > M0 ? s->mob[i] : M0;
    return M0;
}

/* divh = sum_d (i k_d) deal r2c(w(x) c2r(deal (i k_d) src)/bins); `dM` is
 * subtracted from w (the simple splitting uses M - M0, the full scheme M).
 * :257-291 (simple) and :329-357, :495-526 (full) */
static void psm_div_w_grad(och_psm *s, const double *src, const double *w, double dM) {
    const och_params *p = &s->p;
    int dims = p->dims, N = p->N;
    for(int64_t q = 0; q < s->hat; q++) { s->divh[2 * q] = 0.0; s->divh[2 * q + 1] = 0.0; }
    for(int d = 0; d < dims; d++) {
        for(int64_t q = 0; q < s->hat; q++) {
            ik_mul(s->kc[d][q], src + 2 * q, s->gh + 2 * q);
            if(p->use_dealias) { s->gh[2 * q] *= s->deal[q]; s->gh[2 * q + 1] *= s->deal[q]; }
        }
        c2r(dims, N, s->gh, s->g);
        for(int64_t i = 0; i < s->bins; i++) s->g[i] /= s->bins;
        for(int64_t i = 0; i < s->bins; i++) s->g[i] = (w[i] - dM) * s->g[i];
        r2c(dims, N, s->g, s->gh);
        for(int64_t q = 0; q < s->hat; q++) {
            double *f = s->gh + 2 * q, t[2];
            if(p->use_dealias) { f[0] *= s->deal[q]; f[1] *= s->deal[q]; }
            ik_mul(s->kc[d][q], f, t);
            s->divh[2 * q] += t[0];
            s->divh[2 * q + 1] += t[1];
        }
    }
}

void och_psm_div_mob_grad(const och_params *p, const double *mob, const double *g, double *out) {
    och_model m = {OCH_LANDAU, 0.9, 0, 0, 0, 0};
    och_psm *s = och_psm_create(p, &m, g, 0, 30, 200, 1e-10);
    psm_div_w_grad(s, s->rho_hat, mob, 0.0);  /* rho_hat = r2c(g) */
    c2r(p->dims, p->N, s->divh, out);
    for(int64_t i = 0; i < s->bins; i++) out[i] /= s->bins;
    och_psm_destroy(s);
}

/* _evolve_simple, :234-301 */
static void psm_step_simple(och_psm *s, double *rho) {
    const och_params *p = &s->p;
    int dims = p->dims, N = p->N;
    s->M0 = psm_M0(s);
    for(int64_t i = 0; i < s->bins; i++) s->g[i] = och_der_bulk_free_energy(&s->m, rho[i]);
    r2c(dims, N, s->g, s->fh);
    for(int64_t q = 0; q < s->hat; q++) {
        for(int c = 0; c < 2; c++) {
            double v = s->fh[2 * q + c] + 2.0 * p->k * s->k2[q] * s->rho_hat[2 * q + c];
            if(p->use_dealias) v *= s->deal[q];
            s->muh[2 * q + c] = v;
        }
    }
    psm_div_w_grad(s, s->muh, s->mob, s->M0);
    for(int64_t q = 0; q < s->hat; q++) {
        double k2 = s->k2[q], k4 = SQR(k2);
        double denom = 1.0 + p->dt * s->M0 * (p->S * k2 + 2.0 * p->k * k4);
        for(int c = 0; c < 2; c++) {
            double *rh = s->rho_hat + 2 * q + c;
            double v = (*rh + p->dt * s->divh[2 * q + c] -
                        p->dt * s->M0 * k2 * (s->fh[2 * q + c] - p->S * *rh)) / denom;
            *rh = v;
            s->th[2 * q + c] = v / s->bins;
        }
    }
    c2r(dims, N, s->th, rho);
}

/* A x = x - dt div(M grad((S - 2k lap) x)), :466-527 */
static void psm_apply_A(och_psm *s, const double *x, double *Ax) {
    const och_params *p = &s->p;
    r2c(p->dims, p->N, x, s->th);
    for(int64_t q = 0; q < s->hat; q++) {
        for(int c = 0; c < 2; c++) {
            double v = (p->S + 2.0 * p->k * s->k2[q]) * s->th[2 * q + c];
            if(p->use_dealias) v *= s->deal[q];
            s->muh[2 * q + c] = v;
        }
    }
    psm_div_w_grad(s, s->muh, s->mob, 0.0);
    c2r(p->dims, p->N, s->divh, s->tr);
    for(int64_t i = 0; i < s->bins; i++) Ax[i] = x[i] - p->dt * (s->tr[i] / s->bins);
}

/* z = P^-1 r, P = 1 + dt M0 (S k^2 + 2k k^4) in Fourier space, :431-464 */
static void psm_apply_Pinv(och_psm *s, const double *r, double *z) {
    const och_params *p = &s->p;
    r2c(p->dims, p->N, r, s->th);
    for(int64_t q = 0; q < s->hat; q++) {
        double k2 = s->k2[q], k4 = k2 * k2;
        double denom = 1.0 + p->dt * s->M0 * (p->S * k2 + 2.0 * p->k * k4);
        s->th[2 * q] /= denom;
        s->th[2 * q + 1] /= denom;
    }
    c2r(p->dims, p->N, s->th, s->tr);
    for(int64_t i = 0; i < s->bins; i++) z[i] = s->tr[i] / s->bins;
}

static double psm_dot(const double *a, const double *b, int64_t n) {
    double sum = 0.0;
    for(int64_t i = 0; i < n; i++) sum += a[i] * b[i];
    return sum;
}

/* left-preconditioned restarted GMRES (modified Gram-Schmidt, Givens), :529-663 */
static int psm_gmres(och_psm *s, const double *b, double *x) {
    const int64_t n = s->bins;
    const int m = s->restart > 2 ? s->restart : 2;
    double *rhs = (double *)malloc(sizeof(double) * n), *Mx = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
    double *Av = (double *)malloc(sizeof(double) * n);
    double *V = (double *)malloc(sizeof(double) * n * (m + 1));
    double *H = (double *)malloc(sizeof(double) * (m + 1) * m);
    double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
    double *g = (double *)malloc(sizeof(double) * (m + 1)), *y = (double *)malloc(sizeof(double) * m);
#define HH(i, j) H[(i) * m + (j)]
    psm_apply_Pinv(s, b, rhs);
    psm_apply_A(s, x, Av);
    psm_apply_Pinv(s, Av, Mx);
    for(int64_t i = 0; i < n; i++) r[i] = rhs[i] - Mx[i];
    double rn = sqrt(psm_dot(rhs, rhs, n));
    const double rhs_norm = rn > 1e-30 ? rn : 1e-30;
    double beta = sqrt(psm_dot(r, r, n));
    double rel = beta / rhs_norm;
    int iter_total = 0;
    if(rel >= s->tol) {
        while(iter_total < s->max_iter) {
            for(int64_t i = 0; i < n * (m + 1); i++) V[i] = 0.0;
            for(int i = 0; i < (m + 1) * m; i++) H[i] = 0.0;
            for(int i = 0; i < m; i++) { cs[i] = 0.0; sn[i] = 0.0; }
            for(int i = 0; i <= m; i++) g[i] = 0.0;
            for(int64_t i = 0; i < n; i++) V[i] = r[i] / beta;
            g[0] = beta;
            int j_end = 0;
            for(int j = 0; j < m && iter_total < s->max_iter; j++, iter_total++) {
                psm_apply_A(s, V + (int64_t)j * n, Av);
                psm_apply_Pinv(s, Av, w);
                for(int i = 0; i <= j; i++) {
                    HH(i, j) = psm_dot(w, V + (int64_t)i * n, n);
                    for(int64_t k = 0; k < n; k++) w[k] -= HH(i, j) * V[(int64_t)i * n + k];
                }
                HH(j + 1, j) = sqrt(psm_dot(w, w, n));
                if(HH(j + 1, j) > 1e-30) {
                    for(int64_t k = 0; k < n; k++) V[(int64_t)(j + 1) * n + k] = w[k] / HH(j + 1, j);
                }
                else {
                    j_end = j + 1;
                    break;
                }
                for(int i = 0; i < j; i++) {
                    double t1 = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                    double t2 = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                    HH(i, j) = t1;
                    HH(i + 1, j) = t2;
                }
                double a = HH(j, j), b2 = HH(j + 1, j);
                double den = sqrt(a * a + b2 * b2);
                cs[j] = (den > 0) ? (a / den) : 1.0;
                sn[j] = (den > 0) ? (b2 / den) : 0.0;
                HH(j, j) = cs[j] * a + sn[j] * b2;
                HH(j + 1, j) = 0.0;
                double gj = g[j];
                g[j] = cs[j] * gj;
                g[j + 1] = -sn[j] * gj;
                rel = fabs(g[j + 1]) / rhs_norm;
                if(rel < s->tol) {
                    j_end = j + 1;
                    break;
                }
                j_end = j + 1;
            }
            for(int i = j_end - 1; i >= 0; i--) {
                double sum = g[i];
                for(int k = i + 1; k < j_end; k++) sum -= HH(i, k) * y[k];
                y[i] = sum / HH(i, i);
            }
            for(int i = 0; i < j_end; i++)
                for(int64_t k = 0; k < n; k++) x[k] += V[(int64_t)i * n + k] * y[i];
            psm_apply_A(s, x, Av);
            psm_apply_Pinv(s, Av, Mx);
            for(int64_t i = 0; i < n; i++) r[i] = rhs[i] - Mx[i];
            beta = sqrt(psm_dot(r, r, n));
            rel = beta / rhs_norm;
            if(rel < s->tol) break;
        }
    }
#undef HH
    free(rhs); free(Mx); free(r); free(w); free(Av); free(V); free(H); free(cs); free(sn); free(g); free(y);
    return iter_total;
}

/* _evolve_full, :303-404 */
static void psm_step_full(och_psm *s, double *rho) {
    const och_params *p = &s->p;
    int dims = p->dims, N = p->N;
    const int64_t n = s->bins;
    s->M0 = psm_M0(s);
    double *b = (double *)malloc(sizeof(double) * n), *x = (double *)malloc(sizeof(double) * n);
    for(int64_t i = 0; i < n; i++) s->tr[i] = och_der_bulk_free_energy(&s->m, rho[i]) - p->S * rho[i];
    r2c(dims, N, s->tr, s->fh);
    psm_div_w_grad(s, s->fh, s->mob, 0.0);
    c2r(dims, N, s->divh, s->tr);
    for(int64_t i = 0; i < n; i++) {
        b[i] = rho[i] + p->dt * (s->tr[i] / n);
        x[i] = rho[i];
    }
    s->last_iters = psm_gmres(s, b, x);
    memcpy(rho, x, sizeof(double) * n);
    r2c(dims, N, rho, s->rho_hat);
    free(b); free(x);
}

/* one step: IMobility::update_mobility() on rho (CahnHilliard.cpp:279-282), then evolve() */
int och_psm_step(och_psm *s, double *rho) {
    och_update_mobility(&s->p, &s->m, rho, s->mob, s->bins);
    if(s->use_gmres) psm_step_full(s, rho);
    else {
        s->last_iters = 0;
        psm_step_simple(s, rho);
    }
    return s->last_iters;
}

double och_psm_last_M0(const och_psm *s) { return s->M0; }

/* ------------------------------------------------------------------ */
/* multi-species Wertheim models                                       */
/* ------------------------------------------------------------------ */
/* Generic tables, GenericWertheim.cpp:22-150: the unique patch types in
 * ascending order; per unique patch the terms (species, other patch,
 * multiplicity * delta) over the species in order and, within a species, its
 * patch types in ascending order; per species its patch types with
 * multiplicities */
typedef struct {
    int n_unique, unique[OCH_MAX_PATCHES];
    int n_terms[OCH_MAX_PATCHES];
    int t_species[OCH_MAX_PATCHES][OCH_MAX_SPECIES * OCH_MAX_PATCHES];
    int t_other[OCH_MAX_PATCHES][OCH_MAX_SPECIES * OCH_MAX_PATCHES];
    double t_coeff[OCH_MAX_PATCHES][OCH_MAX_SPECIES * OCH_MAX_PATCHES];
} generic_tables;

static void generic_build(const och_species_model *m, generic_tables *g) {
    memset(g, 0, sizeof(*g));
    int used[OCH_MAX_PATCHES] = {0};
    for(int s = 0; s < m->n_species; s++)
        for(int q = 0; q < m->n_patches; q++)
            if(m->patch_count[s][q] > 0) used[q] = 1;
    for(int q = 0; q < m->n_patches; q++)
        if(used[q]) g->unique[g->n_unique++] = q;
    for(int u = 0; u < g->n_unique; u++) {
        int patch = g->unique[u], n = 0;
        for(int s = 0; s < m->n_species; s++)
            for(int q = 0; q < m->n_patches; q++)
                if(m->patch_count[s][q] > 0 && m->delta[patch][q] > 0.0) {
                    g->t_species[u][n] = s;
                    g->t_other[u][n] = q;
                    g->t_coeff[u][n] = (double)m->patch_count[s][q] * m->delta[patch][q];
                    n++;
                }
        g->n_terms[u] = n;
    }
}

void och_ms_update_X(const och_species_model *m, const double *rhos, double *X) {
    generic_tables g;
    generic_build(m, &g);
    const double tolerance = 1e-8;
    for(int iter = 0; iter < 10000; iter++) {
        double max_delta = 0.0;
        for(int u = 0; u < g.n_unique; u++) {
            int patch = g.unique[u];
            double sum = 0.0;
            for(int t = 0; t < g.n_terms[u]; t++) {
                double rho = rhos[g.t_species[u][t]];
                if(rho != 0.0) sum += rho * X[g.t_other[u][t]] * g.t_coeff[u][t];
            }
            double new_X = 1.0 / (1.0 + sum);
            double d = fabs(new_X - X[patch]);
            max_delta = (max_delta < d) ? d : max_delta;
            X[patch] = new_X;
        }
        if(max_delta < tolerance) return;
    }
}

/* SalehWertheim.cpp:80-85 */
static double saleh_contribution(const och_species_model *m, const double *rhos, int species) {
    double delta = (species == 0) ? m->saleh_delta_AA : m->saleh_delta_BB;
    int lh = m->saleh_valence[2] / 2;
    double rho_factor = delta * (m->saleh_valence[species] * rhos[species] + lh * rhos[2]);
    double X = (-1.0 + sqrt(1.0 + 4.0 * rho_factor)) / (2.0 * rho_factor);
    return (rho_factor >= 0) ? log(X) : 0.0;
}

/* SalehWertheim.cpp:87-110, RicciWertheim.cpp:72-98, GenericWertheim.cpp:250-274 */
void och_ms_der(const och_species_model *m, const double *rhos, double *X, double *out) {
    const int S = m->n_species;
    double logX[OCH_MAX_PATCHES];
    if(m->model == OCH_GENERIC) {
        och_ms_update_X(m, rhos, X);
        for(int i = 0; i < m->n_patches; i++) logX[i] = log(X[i]);
    }
    for(int s = 0; s < S; s++) {
        out[s] = 0.0;
        if(rhos[s] == 0.) continue;
        double rho_tot = 0.;
        for(int t = 0; t < S; t++) rho_tot += rhos[t];
        double ref, bond = 0.0;
        if(m->model == OCH_SALEH) {
            ref = log(rhos[s]) + 2.0 * m->saleh_B2 * rho_tot + 3.0 * m->saleh_B3 * SQR(rho_tot);
            int lh = m->saleh_valence[2] / 2;
            if(s == 0) bond = m->saleh_valence[0] * saleh_contribution(m, rhos, 0);
            else if(s == 1) bond = m->saleh_valence[1] * saleh_contribution(m, rhos, 1);
            else bond = lh * (saleh_contribution(m, rhos, 0) + saleh_contribution(m, rhos, 1));
        }
        else if(m->model == OCH_RICCI) {
            ref = log(rhos[s]) + 2.0 * m->ricci_B2 * rho_tot;
            double A = rhos[1] * m->ricci_delta_12, B = rhos[0] * m->ricci_delta_12;
            if(s == 0) {
                double d00 = m->ricci_delta_00;
                double X0 = (-1 + sqrt((1 + 12 * rhos[0] * d00))) / (6 * rhos[0] * d00);
                double X1 = (B - 1 - 2 * A + sqrt(SQR(1 + 2 * A - B) + 4 * B)) / (2 * B);
                bond = 3 * log(X0) + log(X1);
            }
            else {
                double X2 = (-B + 2 * A - 1 + sqrt(SQR(B - 2 * A + 1) + 8 * A)) / (4 * A);
                bond = 2 * log(X2);
            }
        }
        else {
            double B2c = 0.;
            for(int o = 0; o < S; o++) B2c += 2.0 * rhos[o] * m->B2[s][o];
            ref = log(rhos[s]) + B2c;
            for(int q = 0; q < m->n_patches; q++)
                if(m->patch_count[s][q] > 0) bond += m->patch_count[s][q] * logX[q];
        }
        out[s] = ref + bond;
    }
}

static double wertheim_term(double X) { return log(X) - X / 2.0 + 0.5; }

/* SalehWertheim.cpp:47-78, RicciWertheim.cpp:43-70, GenericWertheim.cpp:210-248 */
double och_ms_bulk(const och_species_model *m, double dcf, const double *rhos) {
    const int S = m->n_species;
    double rho = 0.;
    for(int i = 0; i < S; i++) rho += rhos[i];
    double mixing_S = 0., B2c = 0.;
    for(int i = 0; i < S; i++) {
        double x_i = rhos[i] / rho;
        mixing_S += (x_i > 0) ? x_i * log(x_i) : 0;
        if(m->model == OCH_GENERIC)
            for(int j = 0; j < S; j++) B2c += rhos[i] * rhos[j] * m->B2[i][j];
    }
    if(m->model == OCH_SALEH) {
        int lh = m->saleh_valence[2] / 2;
        double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S + m->saleh_B2 * rho + 0.5 * m->saleh_B3 * SQR(rho));
        double rf = m->saleh_delta_AA * (m->saleh_valence[0] * rhos[0] + lh * rhos[2]);
        double X1A = (-1.0 + sqrt(1.0 + 4.0 * rf)) / (2.0 * rf);
        double p1 = log(X1A) - X1A / 2.0 + 0.5;
        rf = m->saleh_delta_BB * (m->saleh_valence[1] * rhos[1] + lh * rhos[2]);
        double X2B = (-1.0 + sqrt(1.0 + 4.0 * rf)) / (2.0 * rf);
        double p2 = log(X2B) - X2B / 2.0 + 0.5;
        double bond = rhos[0] * m->saleh_valence[0] * p1 + rhos[1] * m->saleh_valence[1] * p2 + rhos[2] * lh * (p1 + p2);
        return f_ref + bond;
    }
    if(m->model == OCH_RICCI) {
        double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S + m->ricci_B2 * rho);
        double d00 = m->ricci_delta_00, d12 = m->ricci_delta_12;
        double X0 = (-1 + sqrt((1 + 12 * rhos[0] * d00))) / (6 * rhos[0] * d00);
        double A = rhos[1] * d12, B = rhos[0] * d12;
        double X1 = (B - 1 - 2 * A + sqrt(SQR(1 + 2 * A - B) + 4 * B)) / (2 * B);
        double X2 = (-B + 2 * A - 1 + sqrt(SQR(B - 2 * A + 1) + 8 * A)) / (4 * A);
        double bond = rhos[0] * (3 * (log(X0) - X0 / 2) + log(X1) - X1 / 2 + 2);
        bond += rhos[1] * (2 * log(X2) - X2 + 1);
        return f_ref + bond;
    }
    double f_ref = rho * (log(rho * dcf) - 1.0 + mixing_S) + B2c;
    double X[OCH_MAX_PATCHES];
    for(int q = 0; q < OCH_MAX_PATCHES; q++) X[q] = 0.0;
    och_ms_update_X(m, rhos, X);
    double bond = 0;
    for(int s = 0; s < S; s++) {
        double sf = 0;
        for(int q = 0; q < m->n_patches; q++)
            if(m->patch_count[s][q] > 0) sf += m->patch_count[s][q] * wertheim_term(X[q]);
        bond += rhos[s] * sf;
    }
    return f_ref + bond;
}

struct och_ms {
    int integrator;
    och_params p;
    och_species_model m;
    int S;
    int64_t bins, hat;
    double *X;                 /* n_patches per cell (Generic), patch-major X[q * bins + i] */
    double *fder, *mob, *mu;   /* S * bins */
    double *rho_hat, *fh, *th; /* S * 2 * hat (pseudospectral) */
    och_ps *tab;               /* q^2 and dealias tables */
};

/* f' of every cell (X updated), the model's der_bulk_free_energy over the field */
static void ms_der_field(och_ms *s, const double *rho) {
    const int S = s->S, P = s->m.n_patches;
    for(int64_t i = 0; i < s->bins; i++) {
        double rhos[OCH_MAX_SPECIES], X[OCH_MAX_PATCHES], out[OCH_MAX_SPECIES];
        for(int t = 0; t < S; t++) rhos[t] = rho[t * s->bins + i];
        for(int q = 0; q < P; q++) X[q] = s->X[q * s->bins + i];
        och_ms_der(&s->m, rhos, X, out);
        for(int q = 0; q < P; q++) s->X[q * s->bins + i] = X[q];
        for(int t = 0; t < S; t++) s->fder[t * s->bins + i] = out[t];
    }
}

/* IMobility::update_mobility for every species (RegularisedMobility.h:26-33,
 * GenericWertheim.cpp:276-291 with X of the last f' evaluation) */
static void ms_mobility(och_ms *s, const double *rho) {
    const och_params *p = &s->p;
    for(int t = 0; t < s->S; t++)
        for(int64_t i = 0; i < s->bins; i++) {
            double r = rho[t * s->bins + i], v;
            switch(p->mobility) {
            case OCH_MOB_CONSTANT: v = p->M; break;
            case OCH_MOB_REGULARISED: v = p->M * r / (r + p->rho_min); break;
            case OCH_MOB_LINEAR: v = p->M * r / p->rho_ref; break;
            default: {
                double D = p->M;
                for(int q = 0; q < s->m.n_patches; q++)
                    if(s->m.patch_count[t][q] > 0) D *= pow(s->X[q * s->bins + i], (double)s->m.patch_count[t][q]);
                v = r * D;
            }
            }
            s->mob[t * s->bins + i] = v;
        }
}

och_ms *och_ms_create(int integrator, const och_params *p, const och_species_model *m, const double *rho) {
    och_ms *s = (och_ms *)calloc(1, sizeof(och_ms));
    s->integrator = integrator;
    s->p = *p;
    s->m = *m;
    s->S = m->n_species;
    s->bins = ipow(p->N, p->dims);
    s->hat = ipow(p->N, p->dims - 1) * (p->N / 2 + 1);
    int64_t SB = s->S * s->bins;
    s->X = (double *)malloc(sizeof(double) * (m->n_patches > 0 ? m->n_patches : 1) * s->bins);
    for(int64_t i = 0; i < (int64_t)(m->n_patches > 0 ? m->n_patches : 1) * s->bins; i++) s->X[i] = 1.0;
    s->fder = (double *)calloc(SB, sizeof(double));
    s->mob = (double *)calloc(SB, sizeof(double));
    s->mu = (double *)calloc(SB, sizeof(double));
    /* CahnHilliard.cpp:226-228: one f' evaluation at set-up initialises X */
    ms_der_field(s, rho);
    if(integrator == 2) {
        s->tab = och_ps_create(p, &(och_model){OCH_LANDAU, 0.9, 0, 0, 0, 0}, rho);
        s->rho_hat = (double *)malloc(sizeof(double) * 2 * s->hat * s->S);
        s->fh = (double *)malloc(sizeof(double) * 2 * s->hat * s->S);
        s->th = (double *)malloc(sizeof(double) * 2 * s->hat);
        for(int t = 0; t < s->S; t++) r2c(p->dims, p->N, rho + t * s->bins, s->rho_hat + 2 * t * s->hat);
    }
    return s;
}

void och_ms_destroy(och_ms *s) {
    if(!s) return;
    free(s->X); free(s->fder); free(s->mob); free(s->mu);
    free(s->rho_hat); free(s->fh); free(s->th);
    if(s->tab) och_ps_destroy(s->tab);
    free(s);
}

const double *och_ms_X(const och_ms *s) { return s->X; }

/* one CahnHilliard::evolve (:279-282): update_mobility, then EulerCPU::evolve
 * (:18-36), EulerMobilityCPU::evolve (:31-83) or PseudospectralCPU::evolve
 * (:112-136), every species with its own k */
void och_ms_step(och_ms *s, double *rho) {
    const och_params *p = &s->p;
    const int S = s->S, dims = p->dims, N = p->N;
    const int64_t bins = s->bins;
    ms_mobility(s, rho);
    ms_der_field(s, rho);
    if(s->integrator == 2) {
        double M = s->mob[0];
        for(int t = 0; t < S; t++) {
            double *fh = s->fh + 2 * t * s->hat, *rh = s->rho_hat + 2 * t * s->hat;
            r2c(dims, N, s->fder + t * bins, fh);
            for(int64_t k = 0; k < s->hat; k++) {
                if(p->use_dealias) { fh[2 * k] *= s->tab->dealiaser[k]; fh[2 * k + 1] *= s->tab->dealiaser[k]; }
                double k2 = s->tab->k2[k], k4 = SQR(k2);
                double denom = 1.0 + p->dt * M * (p->S * k2 + 2.0 * s->m.k[t] * k4);
                for(int c = 0; c < 2; c++) {
                    double v = (rh[2 * k + c] - p->dt * M * k2 * (fh[2 * k + c] - p->S * rh[2 * k + c])) / denom;
                    rh[2 * k + c] = v;
                    s->th[2 * k + c] = v / bins;
                }
            }
            c2r(dims, N, s->th, rho + t * bins);
        }
        return;
    }
    for(int t = 0; t < S; t++)
        for(int64_t i = 0; i < bins; i++)
            s->mu[t * bins + i] = s->fder[t * bins + i] - 2 * s->m.k[t] * cell_laplacian(dims, N, p->dx, rho + t * bins, i);
    if(s->integrator == 0) {
        double M = s->mob[0];
        for(int t = 0; t < S; t++)
            for(int64_t i = 0; i < bins; i++)
                s->fder[t * bins + i] = M * cell_laplacian(dims, N, p->dx, s->mu + t * bins, i);
        for(int64_t i = 0; i < S * bins; i++) rho[i] += s->fder[i] * p->dt;
        return;
    }
    for(int t = 0; t < S; t++) {
        const double *mu = s->mu + t * bins, *mob = s->mob + t * bins;
        double *r = rho + t * bins, *upd = s->fder + t * bins;
        for(int64_t i = 0; i < bins; i++) {
            double res = 0;
            for(int d = 0; d < dims; d++) {
                int64_t ip = neighbour(dims, N, i, d, 1), im = neighbour(dims, N, i, d, -1);
                double Jq = 0.5 * (mob[i] + mob[ip]) * (mu[ip] - mu[i]) / p->dx;
                double Jm = 0.5 * (mob[im] + mob[i]) * (mu[i] - mu[im]) / p->dx;
                res += (Jq - Jm) / p->dx;
            }
            upd[i] = res;
        }
        for(int64_t i = 0; i < bins; i++) r[i] += upd[i] * p->dt;
    }
}

/* CahnHilliard.cpp:285-319 over every species */
void och_ms_observables(const och_ms *s, double dcf, const double *rho, double *mass, double *fe) {
    const och_params *p = &s->p;
    const int S = s->S, dims = p->dims, N = p->N;
    double ms = 0., f = 0.;
    for(int64_t i = 0; i < s->bins; i++) {
        double rhos[OCH_MAX_SPECIES], cell = 0.;
        for(int t = 0; t < S; t++) {
            rhos[t] = rho[t * s->bins + i];
            cell += rhos[t];
        }
        ms += cell;
        double inter = 0.;
        int c[3], cp[3];
        fill_coords(dims, N, i, c);
        for(int t = 0; t < S; t++)
            for(int d = 0; d < dims; d++) {
                memcpy(cp, c, sizeof(c));
                cp[d] = (c[d] + 1) & (N - 1);
                double g = (rho[t * s->bins + cell_idx(dims, N, cp)] - rhos[t]) / p->dx;
                inter += s->m.k[t] * g * g;
            }
        f += och_ms_bulk(&s->m, dcf, rhos) + inter;
    }
    double V_bin = p->dx * p->dx * p->dx;
    *mass = ms * V_bin / s->bins;
    *fe = f * V_bin / s->bins;
}
