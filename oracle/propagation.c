/*
 * propagation.c -- oracle restatement of the reference split-step operator
 * (propagation.c:1-677) and of MPC.c's main-loop arithmetic (MPC.c:206-216,
 * 269-483), without MPI.  FP64, test infrastructure only.
 *
 * The field is A[t][y][x] (x fastest), m = k*Nx*Ny + j*Nx + i, 16-byte
 * interleaved complex, as in the reference (settings.h:9-17).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_sizeof_params(void) { return (int)sizeof(oracle_params); }

static inline ocplx cm(ocplx a, ocplx b) {
    ocplx c = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return c;
}
static inline ocplx cconj(ocplx a) { ocplx c = { a.re, -a.im }; return c; }
static inline double abs2(ocplx a) { const double m = hypot(a.re, a.im); return pow(m, 2.0); }

/* propagation.c:378 */
int oracle_pass_sign(double z, double Cav) {
    return ((int)(floor(z / Cav) + 1) % 2) * 2 - 1;
}

/* propagation.c:150-153 (exp of dzz/2 (i beta - alpha/2)) and fftshift axis 0 (:393-402) */
void oracle_exp_D0(ocplx *e, const double *betas, double alpha, double dzz, int nt) {
    ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * (size_t)nt);
    for (int k = 0; k < nt; k++) {
        const double re = -alpha / 2.0 * dzz / 2.0;
        const double im = betas[k] * dzz / 2.0;
        tmp[k].re = exp(re) * cos(im);
        tmp[k].im = exp(re) * sin(im);
    }
    for (int k = 0; k < nt / 2; k++) {
        e[k] = tmp[nt / 2 + k];
        e[nt / 2 + k] = tmp[k];
    }
    free(tmp);
}

/*
 * calcexpR (propagation.c:375-391).  dzz here is the half step.
 * LITERAL (Q2+Q3+Q4): real exp(s dzz/(2k0) (wr[i]-wc)^2), wc = (wr[Ny/2-1]+wr[Ny/2])/2.
 * PHYSICAL: exp(+i s dzz/(2k0) kx_i^2) with kx in FFT order (the intent at :389).
 * The caller multiplies by conj(expR).  signmem is set to s (:380-382).
 */
void oracle_calcexpR(ocplx *expR, int *signmem, double z, double Cav, double k0,
                     double dzz, const double *wr, const int *dims, unsigned q) {
    const int signum = oracle_pass_sign(z, Cav);
    if (-*signmem == signum) *signmem = -*signmem;
    const int ny = dims[1];
    const double wc = (q & OQ_Q4_WC_CENTRE) ? (wr[ny / 2 - 1] + wr[ny / 2]) / 2.0 : 0.0;
    for (int i = 0; i < dims[0]; i++) {
        const int src = (q & OQ_Q3_EXPR_UNSHIFTED) ? i : (i + ny / 2) % ny;
        const double u = pow(wr[src] - wc, 2);
        const double part = signum * (-1.0 / (2.0 * k0) * dzz * (-1.0)) * u;
        if (q & OQ_Q2_REAL_EXPR) {
            expR[i].re = exp(part) * cos(0.0);
            expR[i].im = exp(part) * sin(0.0);
        } else {
            expR[i].re = cos(part);
            expR[i].im = sin(part);
        }
    }
}

/* as calcexpR, with an explicit sign (curved-mirror model: s = +1) */
static void expR_with_sign(ocplx *expR, int signum, double k0, double dzz,
                           const double *wr, const int *dims, unsigned q) {
    const int ny = dims[1];
    const double wc = (q & OQ_Q4_WC_CENTRE) ? (wr[ny / 2 - 1] + wr[ny / 2]) / 2.0 : 0.0;
    for (int i = 0; i < dims[0]; i++) {
        const int src = (q & OQ_Q3_EXPR_UNSHIFTED) ? i : (i + ny / 2) % ny;
        const double u = pow(wr[src] - wc, 2);
        const double part = signum * (-1.0 / (2.0 * k0) * dzz * (-1.0)) * u;
        if (q & OQ_Q2_REAL_EXPR) { expR[i].re = exp(part); expR[i].im = 0.0; }
        else { expR[i].re = cos(part); expR[i].im = sin(part); }
    }
}

/* linearstep, CPU branch (propagation.c:65-107) */
void oracle_linearstep(ocplx *A, const ocplx *exp_D0, const ocplx *expR, const int *dims) {
    const long nx = dims[0], ny = dims[1], nt = dims[2], nxy = nx * ny, N = nxy * nt;
    oracle_fft_axis(A, dims, 2, -1);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++)
        for (long i = 0; i < nxy; i++) A[k * nxy + i] = cm(A[k * nxy + i], exp_D0[k]);
    oracle_fft_axis(A, dims, 2, +1);
    oracle_fft_axis(A, dims, 1, -1);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                const long m = k * nxy + j * ny + i; /* j*dims[1]: square grid (Q19) */
                A[m] = cm(A[m], cconj(expR[j]));
            }
    oracle_fft_axis(A, dims, 1, +1);
    oracle_fft_axis(A, dims, 0, -1);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                const long m = k * nxy + j * ny + i;
                A[m] = cm(A[m], cconj(expR[i]));
            }
    oracle_fft_axis(A, dims, 0, +1);
    const double inv = (double)(nt * ny * nx);
    #pragma omp parallel for schedule(static)
    for (long m = 0; m < N; m++) { A[m].re = A[m].re / inv; A[m].im = A[m].im / inv; }
}

/* self-steepening, CPU branch (propagation.c:164-177, :215-240) */
void oracle_sst(const ocplx *A, ocplx *S, const int *dims, const double *w,
                double gamma2, double w0, int sst_on) {
    const long nxy = (long)dims[0] * dims[1], nt = dims[2], N = nxy * nt;
    if (sst_on != 1) { memset(S, 0, sizeof(ocplx) * (size_t)N); return; }
    double *wbuff = (double *)malloc(sizeof(double) * (size_t)nt);
    memcpy(wbuff, w, sizeof(double) * (size_t)nt);
    for (long k = 0; k < nt / 2; k++) { wbuff[k] = w[nt / 2 + k]; wbuff[nt / 2 + k] = w[k]; }
    ocplx *buf = (ocplx *)malloc(sizeof(ocplx) * (size_t)N);
    memcpy(buf, A, sizeof(ocplx) * (size_t)N);
    oracle_fft_axis(buf, dims, 2, -1);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++)
        for (long i = 0; i < nxy; i++) {
            const ocplx b = buf[k * nxy + i];
            /* -1.0 * I * buffer * wbuff[k] */
            ocplx d = { b.im * wbuff[k], -b.re * wbuff[k] };
            buf[k * nxy + i] = d;
        }
    oracle_fft_axis(buf, dims, 2, +1);
    #pragma omp parallel for schedule(static)
    for (long m = 0; m < N; m++) {
        ocplx dA = { buf[m].re / (double)nt, buf[m].im / (double)nt };
        ocplx t1 = cm(dA, cconj(A[m]));
        ocplx t2 = cm(A[m], cconj(dA));
        S[m].re = gamma2 / w0 * (2.0 * t1.re + t2.re);
        S[m].im = gamma2 / w0 * (2.0 * t1.im + t2.im);
    }
    free(buf);
    free(wbuff);
}

/* peak rho_nt * (ions1 + 2 ions2) seen by the most recent nonlinear step */
static double g_last_rho_peak = 0.0;
double oracle_last_rho_peak(void) { return g_last_rho_peak; }

/* nonlinear half (propagation.c:164-327): A <- A exp(dzz A_nl) */
void oracle_nonlinear(ocplx *A, const oracle_params *p, double dzz, double *rho_peak) {
    const unsigned q = p->quirks;
    const long N = (long)p->dims[0] * p->dims[1] * p->dims[2];
    ocplx *S = (ocplx *)malloc(sizeof(ocplx) * (size_t)N);
    oracle_sst(A, S, p->dims, p->w, p->gamma2, p->w0, p->sst_on);
    double *tau_c = NULL, *F = NULL, *rf = NULL;
    if (p->plasm == 1) {
        tau_c = (double *)malloc(sizeof(double) * (size_t)N);
        F = (double *)malloc(sizeof(double) * (size_t)N);
        rf = (double *)malloc(sizeof(double) * (size_t)N);
        oracle_ioniz(A, p->dims, p->w0, p->deltat, p->gas, p->rho_c, p->w, p->rho_nt, p->n0,
                     tau_c, F, rf, q);
        double mx = 0.0;
        for (long m = 0; m < N; m++) if (rf[m] > mx) mx = rf[m];
        g_last_rho_peak = mx * p->rho_nt;
        if (rho_peak) *rho_peak = g_last_rho_peak;
    } else {
        g_last_rho_peak = 0.0;
        if (rho_peak) *rho_peak = 0.0;
    }
    #pragma omp parallel for schedule(static)
    for (long m = 0; m < N; m++) {
        const double a2 = abs2(A[m]);
        /* z1 = i gamma2 |A|^2 (propagation.c:287-288, :306-308) */
        double nre = 0.0 * p->gamma2 * a2 - S[m].re;
        double nim = 1.0 * p->gamma2 * a2 - S[m].im;
        if (p->plasm == 1) {
            /* z2 = (1 + i w0 tau_c) rho_nt F / 2; the reference's Im part is
             * (1 + w0 tau_c) rho_nt F / 2 (Q21, propagation.c:289-290) */
            const double base = p->rho_nt * F[m] / 2.0;
            const double z2re = (1.0 + 0.0 * p->w0 * tau_c[m]) * base;
            const double z2im = (q & OQ_Q21_PLASMA_IM) ? (1.0 + 1.0 * p->w0 * tau_c[m]) * base
                                                       : (p->w0 * tau_c[m]) * base;
            nre -= z2re;
            nim -= z2im;
        }
        /* propagation.c:317-327 */
        const double er = exp(dzz * nre);
        const double ph = dzz * nim;
        const double c = cos(ph), s = sin(ph);
        if (q & OQ_Q1_NONUNITARY_UPDATE) {
            const double x = A[m].re * er * c - A[m].im * er * s;
            const double y = A[m].im * er * s + x * er * c;
            A[m].re = x;
            A[m].im = y;
        } else {
            const double x = A[m].re * c - A[m].im * s;
            const double y = A[m].re * s + A[m].im * c;
            A[m].re = er * x;
            A[m].im = er * y;
        }
    }
    free(S); free(tau_c); free(F); free(rf);
}

/* BP profiles (propagation.c:343-354) in a fixed summation order */
void oracle_beam_profiles(const ocplx *A, const int *dims, double *BP1, double *BP2) {
    const long nx = dims[0], ny = dims[1], nt = dims[2];
    memset(BP1, 0, sizeof(double) * (size_t)nx);
    memset(BP2, 0, sizeof(double) * (size_t)ny);
    for (long k = 0; k < nt; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                const double v = abs2(A[i + nx * j + k * nx * ny]);
                BP1[i] += v;
                BP2[j] += v;
            }
}

/* calc1oes (propagation.c:460-504), serial.  Q15: posi2 starts at 0. */
void oracle_calc1oes(const double *BP, const int *dims, const double *r, double *vald1oes,
                     double *y, unsigned q) {
    const int n = dims[0];
    double bpex = 0;
    int posi1 = 0, posi2 = (q & OQ_Q15_POSI2_INIT) ? 0 : n / 2;
    double *nb = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) if (BP[k] > bpex) bpex = BP[k];
    for (int k = 0; k < n; k++) nb[k] = BP[k] / bpex;
    const double target = 1.0 / pow(exp(1.0), 2.0);
    for (int k = 0; k < n / 2; k++) {
        const double h = fabs(nb[k] - target);
        if (k == 0) bpex = h;
        else if (h < bpex) { bpex = h; posi1 = k; }
    }
    for (int k = n / 2; k < n; k++) {
        const double h = fabs(nb[k] - target);
        if (k == n / 2) bpex = h;
        else if (h < bpex) { bpex = h; posi2 = k; }
    }
    *vald1oes = -r[posi1] + r[posi2];
    y[0] = r[posi1];
    y[1] = r[posi2];
    y[2] = BP[posi1];
    y[3] = BP[posi2];
    free(nb);
}

/* one split step (propagation.c:111-372), CPU-branch arithmetic */
void oracle_ssfmprop(ocplx *A, const oracle_params *p, double dzz, int *signmem, double z,
                     double *puls_e, double *BP1, double *BP2, double *y1, double *y2,
                     double *bps1, double *bps2, int change) {
    const int nt = p->dims[2];
    ocplx *eD0 = (ocplx *)malloc(sizeof(ocplx) * (size_t)nt);
    ocplx *eR = (ocplx *)malloc(sizeof(ocplx) * (size_t)p->dims[0]);
    oracle_exp_D0(eD0, p->betas, p->alpha, dzz, nt);
    if (p->mirror_R > 0.0) {
        /* curved-mirror model: diffraction sign stays +1 */
        *signmem = 1;
        expR_with_sign(eR, 1, p->k0, dzz / 2, p->wr, p->dims, p->quirks);
    } else {
        oracle_calcexpR(eR, signmem, z, p->Cav, p->k0, dzz / 2, p->wr, p->dims, p->quirks);
    }
    oracle_linearstep(A, eD0, eR, p->dims);
    oracle_nonlinear(A, p, dzz, NULL);
    oracle_linearstep(A, eD0, eR, p->dims);
    if (change == 1) {
        /* propagation.c:334-341 (dead: calcexpR already set signmem) */
        const int signum = oracle_pass_sign(z, p->Cav);
        if (p->mirror_R <= 0.0 && *signmem == -signum) {
            *puls_e = *puls_e * p->Ab_M;
            *signmem *= -1;
        }
        oracle_beam_profiles(A, p->dims, BP1, BP2);
        oracle_calc1oes(BP1, p->dims, p->r, bps1, y1, p->quirks);
        oracle_calc1oes(BP2, p->dims, p->r, bps2, y2, p->quirks);
    }
    free(eD0);
    free(eR);
}

/*
 * Mirror at a pass boundary.  The reference's only mirror effect is the
 * diffraction sign flip inside calcexpR; its absorption branch is dead
 * (Q16).  PHYSICAL applies sqrt(Ab_M) to the field; with mirror_R > 0 it
 * also applies the thin-lens phase exp(-i k0 (x^2+y^2)/R) of a mirror of
 * radius R, and a hard circular aperture when aperture > 0.
 */
void oracle_mirror(ocplx *A, const oracle_params *p) {
    const long nx = p->dims[0], ny = p->dims[1], nt = p->dims[2], nxy = nx * ny;
    const double amp = (p->quirks & OQ_Q16_NO_MIRROR_LOSS) ? 1.0 : sqrt(p->Ab_M);
    ocplx *M = (ocplx *)malloc(sizeof(ocplx) * (size_t)nxy);
    for (long j = 0; j < ny; j++)
        for (long i = 0; i < nx; i++) {
            const double x = p->r[i], y = p->r[j];
            const double rr = x * x + y * y;
            double a = amp;
            if (p->aperture > 0.0 && rr > p->aperture * p->aperture) a = 0.0;
            double ph = 0.0;
            if (p->mirror_R > 0.0) ph = -p->k0 * rr / p->mirror_R;
            M[j * nx + i].re = a * cos(ph);
            M[j * nx + i].im = a * sin(ph);
        }
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++)
        for (long i = 0; i < nxy; i++) A[k * nxy + i] = cm(A[k * nxy + i], M[i]);
    free(M);
}

/* temp/spec as the GPU branch computes them (MPC.c:389-449, kernels.cu:13-24) */
void oracle_temp_spec(const ocplx *A, const int *dims, double *temp, double *spec) {
    const long nxy = (long)dims[0] * dims[1], nt = dims[2], N = nxy * nt;
    ocplx *C = (ocplx *)malloc(sizeof(ocplx) * (size_t)N);
    memcpy(C, A, sizeof(ocplx) * (size_t)N);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++) {
        double s = 0;
        for (long i = 0; i < nxy; i++) s += A[k * nxy + i].re * A[k * nxy + i].re + A[k * nxy + i].im * A[k * nxy + i].im;
        temp[k] = s;
    }
    oracle_fft_axis(C, dims, 2, -1);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < nt; k++) {
        double s = 0;
        for (long i = 0; i < nxy; i++) s += C[k * nxy + i].re * C[k * nxy + i].re + C[k * nxy + i].im * C[k * nxy + i].im;
        spec[k] = s;
    }
    free(C);
}

/* setStep decision table (propagation.c:636-676) */
void oracle_setstep_decide(double *dzz, double err, double def_err, double dzmin,
                           const double *y1, const double *y2, int *bad) {
    *bad = 0;
    const double err_fact = pow(2.0, (1.0 / 3.0));
    double ujdz;
    if (err > 2.0 * def_err) {
        ujdz = *dzz / 2.0;
        if (ujdz > dzmin) { *dzz = ujdz; *bad = 1; return; }
    } else if (err > def_err && err < 2.0 * def_err) {
        ujdz = *dzz / err_fact;
        if (ujdz > dzmin) *dzz = ujdz;
    } else if (err > 0.5 * def_err && err <= def_err) {
        /* keep */
    } else {
        ujdz = *dzz * err_fact;
        if (!(ujdz > 100 * (y2[1] + y1[1]) / 2.0)) *dzz = ujdz;
    }
}

/* setStep (propagation.c:616-677) */
void oracle_setstep(const int *dims, double *dzz, const ocplx *C1, const ocplx *C2,
                    double def_err, double dzmin, const double *y1, const double *y2,
                    int *bad, double *err) {
    const long N = (long)dims[0] * dims[1] * dims[2];
    double sum1 = 0, sum2 = 0;
    for (long i = 0; i < N; i++) {
        ocplx cc = { C2[i].re - C1[i].re, C2[i].im - C1[i].im };
        sum1 += abs2(cc);
        sum2 += abs2(C2[i]);
    }
    *err = sqrt(sum1) / sqrt(sum2);
    oracle_setstep_decide(dzz, *err, def_err, dzmin, y1, y2, bad);
}

/* MPC.c:377-381 */
void oracle_richardson(ocplx *A, const ocplx *C1, const ocplx *C2, long n) {
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        A[i].re = (4.0 / 3.0) * C2[i].re - (1.0 / 3.0) * C1[i].re;
        A[i].im = (4.0 / 3.0) * C2[i].im - (1.0 / 3.0) * C1[i].im;
    }
}

/* MPC.c:206-216 (Q9: only the real part is scaled) */
void oracle_calibrate(ocplx *A, long n, double puls_e, double Esig, unsigned q) {
    const double Ksn = sqrt(puls_e / Esig);
    for (long i = 0; i < n; i++) {
        A[i].re = Ksn * A[i].re;
        if (!(q & OQ_Q9_CALIB_RE_ONLY)) A[i].im = Ksn * A[i].im;
    }
}

/*
 * Fixed-step run.  Step n sits in pass n / steps_per_pass; it is evaluated at
 * z = (n + 1/2) dzz so floor(z/Cav) is the pass index without rounding
 * ambiguity at the boundary.  The mirror acts between passes.
 */
void oracle_run_fixed(ocplx *A, const oracle_params *p, int step0, int nsteps, double *rho_peak) {
    int signmem = 1;
    double puls_e = p->puls_e, bps1 = 0, bps2 = 0, y1[4], y2[4];
    double *BP1 = (double *)malloc(sizeof(double) * (size_t)p->dims[0]);
    double *BP2 = (double *)malloc(sizeof(double) * (size_t)p->dims[1]);
    const int S = p->steps_per_pass > 0 ? p->steps_per_pass : 1;
    const int npass = p->npasses > 0 ? p->npasses : 1;
    for (int n = step0; n < step0 + nsteps; n++) {
        const double z = ((double)n + 0.5) * p->dzz;
        oracle_ssfmprop(A, p, p->dzz, &signmem, z, &puls_e, BP1, BP2, y1, y2, &bps1, &bps2, 0);
        if ((n + 1) % S == 0 && (n + 1) / S < npass) oracle_mirror(A, p);
    }
    /* peak density inside the last step's nonlinear operator */
    if (rho_peak) *rho_peak = g_last_rho_peak;
    free(BP1);
    free(BP2);
}

/*
 * Adaptive run (MPC.c:269-383 sequencing with the MPI-mode change flags):
 * C1 = one 2dz step, C2 = two dz steps (the second with change=1), setStep,
 * Richardson.  Q17: both C2 sub-steps use the same z.  Q22: z advances by
 * 2 * (the dzz setStep returns) instead of the dzz used.
 */
int oracle_run_adaptive(ocplx *A, const oracle_params *p, double *z, double *dzz,
                        int *signmem, int max_steps, double *errs, double *dzs) {
    const long N = (long)p->dims[0] * p->dims[1] * p->dims[2];
    ocplx *C1 = (ocplx *)malloc(sizeof(ocplx) * (size_t)N);
    ocplx *C2 = (ocplx *)malloc(sizeof(ocplx) * (size_t)N);
    double *BP1 = (double *)malloc(sizeof(double) * (size_t)p->dims[0]);
    double *BP2 = (double *)malloc(sizeof(double) * (size_t)p->dims[1]);
    double y1[4] = {0, 0, 0, 0}, y2[4] = {0, 0, 0, 0}, bps1 = 0, bps2 = 0, puls_e = p->puls_e;
    int accepted = 0, guard = 0;
    while (*z < p->dist && accepted < max_steps && guard < 64 * max_steps + 64) {
        guard++;
        memcpy(C1, A, sizeof(ocplx) * (size_t)N);
        memcpy(C2, A, sizeof(ocplx) * (size_t)N);
        int sm1 = *signmem;
        const double dz = *dzz;
        oracle_ssfmprop(C1, p, 2 * dz, &sm1, *z, &puls_e, BP1, BP2, y1, y2, &bps1, &bps2, 0);
        oracle_ssfmprop(C2, p, dz, signmem, *z, &puls_e, BP1, BP2, y1, y2, &bps1, &bps2, 0);
        const double z2 = (p->quirks & OQ_Q17_SAME_Z) ? *z : *z + dz;
        oracle_ssfmprop(C2, p, dz, signmem, z2, &puls_e, BP1, BP2, y1, y2, &bps1, &bps2, 1);
        int bad = 0;
        double err = 0;
        oracle_setstep(p->dims, dzz, C1, C2, p->def_err, p->dzmin, y1, y2, &bad, &err);
        if (bad) continue;
        oracle_richardson(A, C1, C2, N);
        const double zold = *z;
        *z = *z + 2.0 * ((p->quirks & OQ_Q22_Z_ADVANCE_NEW_DZ) ? *dzz : dz);
        if (*z > p->dist) *z = p->dist;
        /* mirror between passes when the accepted step crossed a cell end */
        if (floor(*z / p->Cav) > floor(zold / p->Cav) && *z < p->dist) oracle_mirror(A, p);
        if (errs) errs[accepted] = err;
        if (dzs) dzs[accepted] = *dzz;
        accepted++;
    }
    free(C1); free(C2); free(BP1); free(BP2);
    return accepted;
}
