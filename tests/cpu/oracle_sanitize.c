/* Exercises the FP64 oracle's entry points on a small intense case, in both
 * arithmetic profiles, for a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  Prints a checksum;
 * exits 1 if any result is not finite. */
#include "../../oracle/oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int finite_all(const ocplx *A, long n) {
    for (long i = 0; i < n; i++)
        if (!isfinite(A[i].re) || !isfinite(A[i].im)) return 0;
    return 1;
}

int main(void) {
    const int nx = 8, ny = 8, nt = 64;
    const long nxy = (long)nx * ny, N = nxy * nt;
    const double c0 = 299792458.0, lambda = 1030e-9, pi = 3.14159265358979323846;
    const double w0 = 2 * pi * c0 / lambda, k0 = w0 / c0;
    const double win_t = 1e-12, win_x = 2e-3, dt = win_t / nt, dx = win_x / nx;
    double *betas = malloc(sizeof(double) * nt), *w = malloc(sizeof(double) * nt);
    double *wr = malloc(sizeof(double) * ny), *r = malloc(sizeof(double) * ny);
    for (int k = 0; k < nt; k++) {
        w[k] = 2 * pi * (k - nt / 2) / win_t;
        betas[k] = 0.5 * 40e-30 * w[k] * w[k];
    }
    for (int j = 0; j < ny; j++) {
        wr[j] = 2 * pi * (j - ny / 2) / win_x;
        r[j] = (j - ny / 2) * dx;
    }
    oracle_params p;
    memset(&p, 0, sizeof p);
    p.dims[0] = nx; p.dims[1] = ny; p.dims[2] = nt;
    p.sst_on = 1; p.plasm = 1; p.steps_per_pass = 2; p.npasses = 3;
    p.dzz = 1e-3; p.Cav = 2e-3; p.dist = 6e-3; p.k0 = k0; p.w0 = w0; p.deltat = dt;
    p.gamma2 = 2e-23 * k0 * 1.054560652926899e-34 * w0;
    p.rho_c = 8.854e-12 * 9.1094e-31 * w0 * w0 / (1.602e-19 * 1.602e-19);
    p.rho_nt = 5e25; p.n0 = 1.0; p.puls_e = 1e-3; p.Ab_M = 0.9; p.dr = dx;
    p.def_err = 1e-3; p.dzmin = 1e-7; p.Esig = 1.0; p.mirror_R = 0.3; p.aperture = 3 * dx;
    p.betas = betas; p.wr = wr; p.w = w; p.r = r;
    snprintf(p.gas, sizeof p.gas, "Argon");

    ocplx *A0 = malloc(sizeof(ocplx) * N), *A = malloc(sizeof(ocplx) * N), *B = malloc(sizeof(ocplx) * N);
    for (int k = 0; k < nt; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double t = (k - nt / 2) * dt, x = r[i], y = r[j];
                const double a = 3e18 * exp(-(x * x + y * y) / (4 * dx * dx) - t * t / (2 * 60e-15 * 60e-15));
                A0[k * nxy + j * nx + i].re = a;
                A0[k * nxy + j * nx + i].im = 0.1 * a;
            }
    double sum = 0.0;
    int ok = 1;
    const unsigned profiles[2] = {OQ_PHYSICAL, OQ_LITERAL};
    for (int q = 0; q < 2; q++) {
        p.quirks = profiles[q];
        /* fixed run across the mirrors, one isolated step with the diagnostics */
        memcpy(A, A0, sizeof(ocplx) * N);
        double rho = 0.0;
        oracle_run_fixed(A, &p, 0, 5, &rho);
        /* the LITERAL profile's real diffraction factor diverges (a reference
         * defect): only the PHYSICAL results enter the checks */
        if (q == 0) {
            ok &= finite_all(A, N);
            sum += rho;
        }
        memcpy(B, A0, sizeof(ocplx) * N);
        int sm = 1;
        double pe = p.puls_e, bp1[8], bp2[8], y1[4], y2[4], b1 = 0, b2 = 0;
        oracle_ssfmprop(B, &p, p.dzz, &sm, 0.5 * p.Cav, &pe, bp1, bp2, y1, y2, &b1, &b2, 1);
        if (q == 0) sum += b1 + b2 + y1[1] + y2[1];
        /* adaptive steps */
        memcpy(A, A0, sizeof(ocplx) * N);
        double z = 0.0, dz = p.dzz, errs[3], dzs[3];
        sm = 1;
        const int acc = oracle_run_adaptive(A, &p, &z, &dz, &sm, 3, errs, dzs);
        if (q == 0) sum += z + acc;
        /* plasma and diagnostics on their own */
        double *tau = malloc(sizeof(double) * N), *F = malloc(sizeof(double) * N), *rf = malloc(sizeof(double) * N);
        oracle_ioniz(A0, p.dims, w0, dt, "Argon", p.rho_c, w, p.rho_nt, p.n0, tau, F, rf, p.quirks);
        if (q == 0)
            for (long i = 0; i < N; i++) sum += rf[i];
        free(tau); free(F); free(rf);
        double temp[64], spec[64];
        oracle_temp_spec(A0, p.dims, temp, spec);
        oracle_beam_profiles(A0, p.dims, bp1, bp2);
        double v = 0.0;
        oracle_calc1oes(bp1, p.dims, r, &v, y1, p.quirks);
        if (q == 0) sum += v + temp[nt / 2] / spec[0];
        int bad = 0;
        double err = 0.0;
        dz = p.dzz;
        oracle_setstep(p.dims, &dz, A0, B, p.def_err, p.dzmin, y1, y2, &bad, &err);
        oracle_richardson(B, A0, B, N);
        oracle_calibrate(B, N, p.puls_e, 1.0, p.quirks);
        if (q == 0) sum += err + dz;
    }
    printf("checksum %.6e finite %d\n", sum, ok && isfinite(sum));
    free(A0); free(A); free(B); free(betas); free(w); free(wr); free(r);
    return (ok && isfinite(sum)) ? 0 : 1;
}
