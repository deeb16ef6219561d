/* Exercises the oracle (oracle/ch_oracle.c) on small grids so that an
 * AddressSanitizer / UndefinedBehaviorSanitizer build reports any invalid
 * access: every integrator and model, the spectral transforms, S(q), the
 * adjoints, conserved noise, pseudospectral_mobility (both schemes) and a
 * multi-species system.  Test infrastructure: built and run by
 * tests/test_sanitizers.py with gcc. */
#include <stdio.h>
#include <stdlib.h>

#include "ch_oracle.h"

int main(void) {
    och_model lan = {0};
    lan.model = 0;
    lan.epsilon = 0.9;
    och_model wer = {0};
    wer.model = 1;
    wer.B2 = 1.0;
    wer.delta = 1e4;
    wer.valence = 4;
    wer.regularisation_delta = 0.01;
    double sum = 0;
    for(int dims = 1; dims <= 3; dims++) {
        const int N = dims == 1 ? 32 : (dims == 2 ? 16 : 8);
        int64_t n = 1;
        for(int d = 0; d < dims; d++) n *= N;
        double *rho = malloc(n * sizeof(double)), *lam = malloc(n * sizeof(double));
        double *out = malloc(n * sizeof(double)), grad[8] = {0};
        for(int integ = 0; integ < 3; integ++)
            for(int mob = 0; mob < 4; mob++) {
                if(integ != 1 && mob != 0) continue;
                och_params p = {dims, N, integ == 2 ? 0.05 : 1e-3, 0.9, 1.1, mob, 1.0, 0.1, 0.05, 0.3, 1};
                const och_model *m = mob == 1 ? &wer : &lan;
                och_init_density(dims, N, mob == 1 ? 0.05 : 0.3, 0.1, 0, 7, rho);
                och_run(integ, &p, m, rho, 5);
                sum += och_average_mass(&p, rho) + och_average_free_energy(&p, m, rho);
                for(int64_t i = 0; i < n; i++) lam[i] = rho[i];
                och_adjoint(integ, &p, m, rho, 3, lam, out, grad);
                sum += out[0] + grad[0];
            }
        och_params pn = {dims, N, 1e-3, 1.0, 1.0, 2, 1.0, 0.2, 0.05, 0.0, 0};
        och_init_density(dims, N, 0.4, 0.1, 0, 9, rho);
        och_euler_mobility_noise_step(&pn, &lan, 0.1, 42, 3, rho);
        for(int gm = 0; gm < 2; gm++) {
            och_psm *s = och_psm_create(&pn, &lan, rho, gm, 10, 50, 1e-10);
            och_psm_step(s, rho);
            och_psm_destroy(s);
        }
        sum += rho[n - 1];
        free(rho);
        free(lam);
        free(out);
    }
    {
        double f[64 * 64], Sq[64], g[64 * 64];
        for(int i = 0; i < 64 * 64; i++) f[i] = (i % 7) * 0.01;
        const int ns = och_sq_nshells(64);
        if(ns > 64) return 2;
        och_radial_sq(64, 1.0, f, Sq);
        sum += och_sq_loss(64, 1.0, f, Sq, g);
        double in[2 * 12], o[2 * 12];
        for(int i = 0; i < 24; i++) in[i] = i * 0.1;
        och_dft_brute(12, 0, in, o);
        sum += o[3];
    }
    {
        och_species_model sm = {0};
        sm.model = 2;
        sm.n_species = 3;
        for(int s = 0; s < 3; s++) sm.k[s] = 1.0;
        sm.saleh_valence[0] = 4;
        sm.saleh_valence[1] = 4;
        sm.saleh_valence[2] = 2;
        sm.saleh_B2 = 0.7;
        sm.saleh_B3 = 0.0;
        sm.saleh_delta_AA = 1e3;
        sm.saleh_delta_BB = 2e3;
        och_params p = {2, 16, 1e-4, 1.0, 1.0, 0, 1.0, 0.1, 0.05, 0.0, 0};
        double rho[3 * 256], mass[3], fe;
        for(int i = 0; i < 3 * 256; i++) rho[i] = 0.03 * (1.0 + 0.05 * ((i * 37) % 11 - 5) / 5.0);
        for(int integ = 0; integ < 3; integ++) {
            och_ms *s = och_ms_create(integ, &p, &sm, rho);
            och_ms_step(s, rho);
            och_ms_observables(s, 1.0, rho, mass, &fe);
            och_ms_destroy(s);
            sum += mass[0] + fe;
        }
    }
    printf("checksum %.17g\n", sum);
    return 0;
}
