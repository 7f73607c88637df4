/*
 * host.c -- host-side pieces of the library that touch no device memory:
 * the synthetic-input generator, param.bin I/O, the checkpoint file helpers
 * (utils.h) and the small 1-D table helpers of propagation.h / plasma.h.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/constants.h"
#include "../../include/mpc.h"
#include "../../include/plasma.h"
#include "../../include/propagation.h"
#include "../../include/utils.h"

int mpc_sizeof_params(void) { return (int)sizeof(mpc_params); }
int mpc_sizeof_input_spec(void) { return (int)sizeof(mpc_input_spec); }

/* ------------------------------------------------------------------------ */
/* synthetic input generator                                                 */
/* ------------------------------------------------------------------------ */
/* SplitMix64 in counter form: the n-th output of the generator seeded with
 * `seed` (state advanced n+1 times), so elements can be drawn in parallel */
static uint64_t splitmix64_at(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01_open(uint64_t v) { /* (0, 1] */
    return ((double)(v >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}
static double u01(uint64_t v) { /* [0, 1) */
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

int mpc_mkinput(const mpc_input_spec *s, mpc_params *p, double *betas, double *wr, double *w,
                double *r, cufftDoubleComplex *A) {
    if (!s || !p || s->nx <= 0 || s->ny <= 0 || s->nt <= 0) return MPC_EINVAL;
    if (s->nx != s->ny || s->steps_per_pass <= 0 || s->passes <= 0) return MPC_EINVAL;
    const int nx = s->nx, ny = s->ny, nt = s->nt;
    const double eps0 = 8.8541878128e-12;
    const double dx = s->window_xy / nx, dt = s->window_t / nt;
    const double w0 = 2.0 * M_PI * C0 / s->lambda0;
    const double n0 = 1.0;
    const double k0 = n0 * w0 / C0;
    memset(p, 0, sizeof(*p));
    p->dims[0] = nx; p->dims[1] = ny; p->dims[2] = nt;
    p->sst_on = s->sst_on;
    p->plasm = s->plasm;
    p->steps_per_pass = s->steps_per_pass;
    p->npasses = s->passes;
    p->Cav = s->cell_L;
    p->dist = s->passes * s->cell_L;
    p->dzz = s->cell_L / s->steps_per_pass;
    p->alpha = 0.0;
    p->k0 = k0;
    p->w0 = w0;
    p->gamma2 = s->n2_per_bar * s->pressure * k0 * HBAR * w0;
    p->deltat = dt;
    p->rho_c = eps0 * ME * w0 * w0 / (EL0 * EL0);
    p->rho_nt = s->rho_nt_per_bar * s->pressure;
    p->n0 = n0;
    p->puls_e = s->energy;
    p->Ab_M = s->Ab_M;
    p->dr = dx;
    p->def_err = s->def_err;
    p->dzmin = s->dzmin;
    p->mirror_R = s->mirror_R;
    p->aperture = s->aperture;
    p->quirks = 0;
    snprintf(p->gas, sizeof(p->gas), "%s", s->gas[0] ? s->gas : "Argon");
    for (int k = 0; k < nt; k++) {
        const double Om = (k - nt / 2) * 2.0 * M_PI / (nt * dt);
        w[k] = Om;
        betas[k] = s->beta2 / 2.0 * Om * Om;
    }
    for (int j = 0; j < ny; j++) {
        wr[j] = (j - ny / 2) * 2.0 * M_PI / (ny * dx);
        r[j] = (j - ny / 2) * dx;
    }
    p->betas = betas; p->wr = wr; p->w = w; p->r = r;
    /* Gaussian in x, y, t: exp(-(x^2+y^2)/waist^2) exp(-2 ln2 t^2/tau^2), plus
     * complex Gaussian noise (Box-Muller on draws 2m, 2m+1 for element m) */
    const double sig = s->noise_rel; /* max|A| = 1 at the grid centre */
    const uint64_t seed = s->seed;
    double *fx = (double *)malloc(sizeof(double) * (size_t)nx);
    double *ft = (double *)malloc(sizeof(double) * (size_t)nt);
    for (int i = 0; i < nx; i++) fx[i] = exp(-(r[i] * r[i]) / (s->waist * s->waist));
    for (int k = 0; k < nt; k++) {
        const double t = (k - nt / 2) * dt;
        ft[k] = exp(-2.0 * log(2.0) * t * t / (s->tau_fwhm * s->tau_fwhm));
    }
    double esum = 0.0;
    #pragma omp parallel for schedule(static) reduction(+ : esum)
    for (int k = 0; k < nt; k++) {
        double ek = 0.0;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long m = ((long)k * ny + j) * nx + i;
                const double a = fx[i] * fx[j] * ft[k];
                const double u1 = u01_open(splitmix64_at(seed, 2 * (uint64_t)m));
                const double u2 = u01(splitmix64_at(seed, 2 * (uint64_t)m + 1));
                const double rad = sqrt(-2.0 * log(u1));
                const double re = a + sig * rad * cos(2.0 * M_PI * u2);
                const double im = sig * rad * sin(2.0 * M_PI * u2);
                A[m].x = re;
                A[m].y = im;
                ek += re * re + im * im;
            }
        esum += ek;
    }
    free(fx);
    free(ft);
    p->Esig = esum * HBAR * w0 * dx * dx * dt;
    return MPC_OK;
}

/* ------------------------------------------------------------------------ */
/* param.bin (MPC.c:76-77, 180-203): packed, native endian, sequential        */
/* ------------------------------------------------------------------------ */
int mpc_write_param_bin(const char *path, const mpc_params *p) {
    FILE *f = fopen(path, "wb");
    if (!f) return MPC_EIO;
    const int nt = p->dims[2], ny = p->dims[1];
    char gas[9];
    memset(gas, 0, sizeof gas);
    strncpy(gas, p->gas, 8);
    int ok = 1;
    ok &= fwrite(p->dims, sizeof(int), 3, f) == 3;
    ok &= fwrite(&p->dist, 8, 1, f) == 1;
    ok &= fwrite(&p->dzz, 8, 1, f) == 1;
    ok &= fwrite(p->betas, 8, (size_t)nt, f) == (size_t)nt;
    ok &= fwrite(&p->alpha, 8, 1, f) == 1;
    ok &= fwrite(&p->Cav, 8, 1, f) == 1;
    ok &= fwrite(&p->k0, 8, 1, f) == 1;
    ok &= fwrite(p->wr, 8, (size_t)ny, f) == (size_t)ny;
    ok &= fwrite(&p->sst_on, sizeof(int), 1, f) == 1;
    ok &= fwrite(p->w, 8, (size_t)nt, f) == (size_t)nt;
    ok &= fwrite(&p->gamma2, 8, 1, f) == 1;
    ok &= fwrite(&p->w0, 8, 1, f) == 1;
    ok &= fwrite(&p->plasm, sizeof(int), 1, f) == 1;
    ok &= fwrite(&p->deltat, 8, 1, f) == 1;
    ok &= fwrite(&p->rho_c, 8, 1, f) == 1;
    ok &= fwrite(&p->rho_nt, 8, 1, f) == 1;
    ok &= fwrite(&p->n0, 8, 1, f) == 1;
    ok &= fwrite(&p->puls_e, 8, 1, f) == 1;
    ok &= fwrite(&p->Ab_M, 8, 1, f) == 1;
    ok &= fwrite(p->r, 8, (size_t)ny, f) == (size_t)ny;
    ok &= fwrite(&p->dr, 8, 1, f) == 1;
    ok &= fwrite(&p->def_err, 8, 1, f) == 1;
    ok &= fwrite(&p->dzmin, 8, 1, f) == 1;
    ok &= fwrite(&p->Esig, 8, 1, f) == 1;
    ok &= fwrite(gas, 1, 9, f) == 9;
    fclose(f);
    return ok ? MPC_OK : MPC_EIO;
}

int mpc_read_param_bin(const char *path, mpc_params *p) {
    FILE *f = fopen(path, "rb");
    if (!f) return MPC_EIO;
    memset(p, 0, sizeof(*p));
    int ok = fread(p->dims, sizeof(int), 3, f) == 3;
    if (!ok || p->dims[0] <= 0 || p->dims[1] <= 0 || p->dims[2] <= 0) { fclose(f); return MPC_EIO; }
    const int nt = p->dims[2], ny = p->dims[1];
    double *betas = (double *)malloc(8 * (size_t)nt), *w = (double *)malloc(8 * (size_t)nt);
    double *wr = (double *)malloc(8 * (size_t)ny), *r = (double *)malloc(8 * (size_t)ny);
    char gas[9] = {0};
    ok &= fread(&p->dist, 8, 1, f) == 1;
    ok &= fread(&p->dzz, 8, 1, f) == 1;
    ok &= fread(betas, 8, (size_t)nt, f) == (size_t)nt;
    ok &= fread(&p->alpha, 8, 1, f) == 1;
    ok &= fread(&p->Cav, 8, 1, f) == 1;
    ok &= fread(&p->k0, 8, 1, f) == 1;
    ok &= fread(wr, 8, (size_t)ny, f) == (size_t)ny;
    ok &= fread(&p->sst_on, sizeof(int), 1, f) == 1;
    ok &= fread(w, 8, (size_t)nt, f) == (size_t)nt;
    ok &= fread(&p->gamma2, 8, 1, f) == 1;
    ok &= fread(&p->w0, 8, 1, f) == 1;
    ok &= fread(&p->plasm, sizeof(int), 1, f) == 1;
    ok &= fread(&p->deltat, 8, 1, f) == 1;
    ok &= fread(&p->rho_c, 8, 1, f) == 1;
    ok &= fread(&p->rho_nt, 8, 1, f) == 1;
    ok &= fread(&p->n0, 8, 1, f) == 1;
    ok &= fread(&p->puls_e, 8, 1, f) == 1;
    ok &= fread(&p->Ab_M, 8, 1, f) == 1;
    ok &= fread(r, 8, (size_t)ny, f) == (size_t)ny;
    ok &= fread(&p->dr, 8, 1, f) == 1;
    ok &= fread(&p->def_err, 8, 1, f) == 1;
    ok &= fread(&p->dzmin, 8, 1, f) == 1;
    ok &= fread(&p->Esig, 8, 1, f) == 1;
    ok &= fread(gas, 1, 9, f) == 9;
    fclose(f);
    gas[8] = 0;
    snprintf(p->gas, sizeof(p->gas), "%s", gas);
    p->betas = betas; p->w = w; p->wr = wr; p->r = r;
    if (!ok) { mpc_free_params(p); return MPC_EIO; }
    return MPC_OK;
}

void mpc_free_params(mpc_params *p) {
    free((void *)p->betas); free((void *)p->w); free((void *)p->wr); free((void *)p->r);
    p->betas = p->w = p->wr = p->r = NULL;
}

/* ------------------------------------------------------------------------ */
/* utils.h (utils.c:4-69): .bak backup/restore.  The copy is done in C       */
/* rather than through system("cp ...").                                     */
/* ------------------------------------------------------------------------ */
static int copy_bytes(const char *from, const char *to) {
    FILE *a = fopen(from, "rb");
    if (!a) return -1;
    FILE *b = fopen(to, "wb");
    if (!b) { fclose(a); return -1; }
    char buf[1 << 16];
    size_t n;
    int rc = 0;
    while ((n = fread(buf, 1, sizeof buf, a)) > 0)
        if (fwrite(buf, 1, n, b) != n) { rc = -1; break; }
    fclose(a);
    if (fclose(b) != 0) rc = -1;
    return rc;
}

void copyfile(const char *path, int dorestore) {
    size_t L = strlen(path) + 5;
    char *bak = (char *)malloc(L);
    if (!bak) { fprintf(stderr, "Memory allocation failed in copyfile function"); exit(1); }
    snprintf(bak, L, "%s.bak", path);
    const int rc = dorestore ? copy_bytes(bak, path) : copy_bytes(path, bak);
    if (rc != 0) fprintf(stderr, "copyfile: could not copy %s (%s)\n", dorestore ? bak : path, strerror(errno));
    free(bak);
}

void restore(const char *path) { copyfile(path, 1); }

void backup(FILE *fp) {
    char link[64], path[PATH_MAX];
    fflush(fp);
    snprintf(link, sizeof link, "/proc/self/fd/%d", fileno(fp));
    memset(path, 0, sizeof path);
    if (readlink(link, path, sizeof(path) - 1) < 0) { fprintf(stderr, "backup: readlink failed\n"); return; }
    copyfile(path, 0);
}

FILE *opener(const char *cpath, const char *fname, const char *fopts, int dorestore) {
    size_t L = strlen(cpath) + strlen(fname) + 2;
    char *path = (char *)malloc(L);
    if (!path) { fprintf(stderr, "Memory allocation failed in opener function\n"); exit(1); }
    snprintf(path, L, "%s/%s", cpath, fname);
    if (dorestore) restore(path);
    FILE *fp = fopen(path, fopts);
    if (!fp) { fprintf(stderr, "Failed to open file: %s\n", path); exit(1); }
    free(path);
    return fp;
}

/* ------------------------------------------------------------------------ */
/* small table helpers (propagation.h / plasma.h); corrected arithmetic      */
/* ------------------------------------------------------------------------ */
static int pass_sign(double z, double Cav) { return ((int)(floor(z / Cav) + 1) % 2) * 2 - 1; }

/* calcexpR (propagation.c:375-391) with the phase the reference intends at
 * :389: expR[i] = exp(+i s dzz/(2 k0) kx_i^2), kx in FFT order. */
void calcexpR(cufftDoubleComplex *expR, int *signmem, double z, double Cav, double k0,
              double dzz, double *wr, int *dims) {
    const int signum = pass_sign(z, Cav);
    if (-*signmem == signum) *signmem = -*signmem;
    const int ny = dims[1];
    for (int i = 0; i < dims[0]; i++) {
        const double kx = wr[(i + ny / 2) % ny];
        const double ph = signum * dzz / (2.0 * k0) * kx * kx;
        expR[i].x = cos(ph);
        expR[i].y = sin(ph);
    }
}

/* fftshift (propagation.c:393-457): swap halves along one axis */
void fftshift(cufftDoubleComplex *in, int *dims, cufftDoubleComplex *buffer, int axis) {
    const long nx = dims[0], ny = dims[1], nt = dims[2];
    if (axis == 0) {
        memcpy(buffer, in, sizeof(cufftDoubleComplex) * (size_t)nt);
        for (long i = 0; i < nt / 2; i++) { in[i] = buffer[nt / 2 + i]; in[nt / 2 + i] = buffer[i]; }
        return;
    }
    const long N = nx * ny * nt;
    memcpy(buffer, in, sizeof(cufftDoubleComplex) * (size_t)N);
    for (long k = 0; k < nt; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                long si = i, sj = j, sk = k;
                if (axis == 1) si = (i + nx / 2) % nx;
                if (axis == 2) sj = (j + ny / 2) % ny;
                if (axis == 3) sk = (k + nt / 2) % nt;
                in[k * nx * ny + j * nx + i] = buffer[sk * nx * ny + sj * nx + si];
            }
}

/* calc1oes (propagation.c:460-504), serial, posi2 searched from Nx/2 */
void calc1oes(double *BP, int *dims, double *r, double *vald1oes, double *y) {
    const int n = dims[0];
    double mx = 0;
    for (int k = 0; k < n; k++) if (BP[k] > mx) mx = BP[k];
    const double target = 1.0 / pow(exp(1.0), 2.0);
    int p1 = 0, p2 = n / 2;
    double b1 = 0, b2 = 0;
    for (int k = 0; k < n / 2; k++) {
        const double h = fabs(BP[k] / mx - target);
        if (k == 0) b1 = h;
        else if (h < b1) { b1 = h; p1 = k; }
    }
    for (int k = n / 2; k < n; k++) {
        const double h = fabs(BP[k] / mx - target);
        if (k == n / 2) b2 = h;
        else if (h < b2) { b2 = h; p2 = k; }
    }
    *vald1oes = -r[p1] + r[p2];
    y[0] = r[p1]; y[1] = r[p2]; y[2] = BP[p1]; y[3] = BP[p2];
}

/* plasma.c:4-40 */
void ionizpot(const char *gas, double *I1, double *I2) {
    if (strcmp(gas, "Argon") == 0)    { *I1 = 15.75962; *I2 = 27.62967; }
    if (strcmp(gas, "Helium") == 0)   { *I1 = 24.58741; *I2 = 54.41778; }
    if (strcmp(gas, "Neon") == 0)     { *I1 = 21.5646;  *I2 = 40.96328; }
    if (strcmp(gas, "Air") == 0) {
        *I1 = 0.78084 * 14.53414 + 0.20946 * 13.61806 + 0.00934 * 15.75962;
        *I2 = 0.78084 * 29.6013 + 0.20946 * 35.11730 + 0.00934 * 27.62967;
    }
    if (strcmp(gas, "Nitrogen") == 0) { *I1 = 14.53414; *I2 = 29.6013; }
    if (strcmp(gas, "Xenon") == 0)    { *I1 = 12.1298;  *I2 = 21.20979; }
    if (strcmp(gas, "Krypton") == 0)  { *I1 = 13.99961; *I2 = 24.35985; }
}

/* plasma.c:61-115 */
void Natoms(const char *gas, double *Natm, double *sigm_coll) {
    if (strcmp(gas, "Argon") == 0)    { *Natm = 1.784e3 * NAVO / 39.941;    *sigm_coll = 1.57e-20; }
    if (strcmp(gas, "Helium") == 0)   { *Natm = 0.166e3 * NAVO / 4.002602; *sigm_coll = 6.11e-20; }
    if (strcmp(gas, "Neon") == 0)     { *Natm = 0.9002e3 * NAVO / 20.1797; *sigm_coll = 1.65e-20; }
    if (strcmp(gas, "Air") == 0)      { *Natm = 1.205e3 * NAVO / 28.97;    *sigm_coll = 10e-20; }
    if (strcmp(gas, "Nitrogen") == 0) { *Natm = 1.65e3 * NAVO / 14.0067;   *sigm_coll = 10.2e-20; }
    if (strcmp(gas, "Xenon") == 0)    { *Natm = 5.86e3 * NAVO / 5.761;     *sigm_coll = 3.55e-20; }
    if (strcmp(gas, "Krypton") == 0)  { *Natm = 3.749e3 * NAVO / 83.798;   *sigm_coll = 1.15e-20; }
}

/* the ADK constants (plasma.c:45-51 with the quantum numbers of :159-174),
 * corrected: n* of each ionisation from its own potential, C_nl with
 * 2^(2 n*), |C_nl| as a double.  kon[0..5] = Kon1..Kon6; returns 0 if the
 * gas is known. */
int mpc_adk_constants(const char *gas, double *kon) {
    double I1 = -1, I2 = -1;
    ionizpot(gas, &I1, &I2);
    if (I1 < 0) return MPC_EINVAL;
    const double Ei0 = I1 / IH, Ei1 = I2 / IH;
    const double nq0 = pow(2 * Ei0, -0.5), nq1 = 2.0 * pow(2 * Ei1, -0.5);
    const double c0 = sqrt(pow(2.0, 2.0 * nq0) / (nq0 * tgamma(nq0 + 1.0) * tgamma(nq0)));
    const double c1 = sqrt(pow(2.0, 2.0 * nq1) / (nq1 * tgamma(nq1 + 1.0) * tgamma(nq1)));
    kon[0] = pow(fabs(c0), 2.0) * sqrt(6.0 / M_PI) * 1.0 * Ei0;
    kon[1] = pow(fabs(c1), 2.0) * sqrt(6.0 / M_PI) * 1.0 * Ei1;
    kon[2] = 2.0 * pow(2.0 * Ei0, 1.5);
    kon[3] = 2.0 * pow(2.0 * Ei1, 1.5);
    kon[4] = 2.0 * nq0 - 0.0 - 1.5;
    kon[5] = 2.0 * nq1 - 0.0 - 1.5;
    return MPC_OK;
}

/* ------------------------------------------------------------------------ */
/* solver guard verdict (mpc.h)                                              */
/* ------------------------------------------------------------------------ */
int mpc_guard_verdict(double e_before, double e_after, double w_before, double w_after, double gain_tol,
                      double energy_tol, double width_tol) {
    if (!isfinite(e_after)) return MPC_ENONFINITE;
    if (width_tol >= 0.0 && !isfinite(w_after)) return MPC_ENONFINITE;
    if (gain_tol >= 0.0 && e_after > e_before * (1.0 + gain_tol)) return MPC_EGUARD;
    if (energy_tol >= 0.0 && !(fabs(e_after / e_before - 1.0) <= energy_tol)) return MPC_EGUARD;
    if (width_tol >= 0.0 && !(fabs(w_after / w_before - 1.0) <= width_tol)) return MPC_EGUARD;
    return 0;
}
