/*
 * plasma.c -- oracle restatement of the reference plasma model
 * (plasma.c:4-246): ADK photo-ionisation, collision time, inverse
 * bremsstrahlung cross-section, avalanche rate and the reverse-t density
 * recurrence.  FP64, test infrastructure only.
 *
 * Constants are the reference's numeric data (constants.h:1-12,
 * plasma.c:6-39,67-114).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZW0 377.0
#define NC 1.0
#define FCTR 0.89
#define IH 27.21
#define EF 5.14220642e11
#define ME 9.1094e-31
#define EL0 1.602e-19
#define NAVO 6.02214e+23
#define T0 2.4188843265e-17
#define C0 299792458.0
#define HBAR 1.054560652926899e-34

/* plasma.c:4-40 -- ionisation potentials in eV (unknown gas: left untouched) */
void oracle_ionizpot(const char *gas, double *I1, double *I2) {
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

/* plasma.c:61-115 -- number density at 1 bar and collision cross-section */
void oracle_natoms(const char *gas, double *Natm, double *sigm_coll) {
    if (strcmp(gas, "Argon") == 0)    { *Natm = 1.784e3 * NAVO / 39.941;    *sigm_coll = 1.57e-20; }
    if (strcmp(gas, "Helium") == 0)   { *Natm = 0.166e3 * NAVO / 4.002602; *sigm_coll = 6.11e-20; }
    if (strcmp(gas, "Neon") == 0)     { *Natm = 0.9002e3 * NAVO / 20.1797; *sigm_coll = 1.65e-20; }
    if (strcmp(gas, "Air") == 0)      { *Natm = 1.205e3 * NAVO / 28.97;    *sigm_coll = 10e-20; }
    if (strcmp(gas, "Nitrogen") == 0) { *Natm = 1.65e3 * NAVO / 14.0067;   *sigm_coll = 10.2e-20; }
    if (strcmp(gas, "Xenon") == 0)    { *Natm = 5.86e3 * NAVO / 5.761;     *sigm_coll = 3.55e-20; }
    if (strcmp(gas, "Krypton") == 0)  { *Natm = 3.749e3 * NAVO / 83.798;   *sigm_coll = 1.15e-20; }
}

/*
 * ADK constants (plasma.c:45-51 using the quantum numbers of :159-174).
 * kon[0..5] = Kon1..Kon6.  With Q6 the second-ionisation rate is defined as
 * zero (the reference reads Iion[2] out of bounds), signalled by kon[1] = 0.
 */
void oracle_adk_consts(const char *gas, unsigned q, double *kon) {
    double I1 = 0, I2 = 0;
    oracle_ionizpot(gas, &I1, &I2);
    const double Iion0 = I1 / IH, Iion1 = I2 / IH;
    /* plasma.c:162 (Q5: uses Iion[1] for the first ionisation) */
    const double nq0 = (q & OQ_Q5_NSTAR_I2) ? pow(2 * Iion1, -0.5) : pow(2 * Iion0, -0.5);
    /* plasma.c:163 (Q6: Iion[2] is out of bounds) */
    const double nq1 = 2.0 * pow(2 * Iion1, -0.5);
    const double lq = 0.0, mq = 0.0, f_lm = 1.0;
    /* plasma.c:173-174 */
    const double C0n = sqrt(pow(2.0, 2.0 * nq0) * pow(nq0 * tgamma(nq0 + lq + 1.0) * tgamma(nq0 - lq), -1.0));
    const double C1n = sqrt(pow(2.0, (q & OQ_Q7_CNL1_2N) ? nq1 : 2.0 * nq1) *
                            pow(nq1 * tgamma(nq1 + lq + 1.0) * tgamma(nq1 - lq), -1.0));
    /* plasma.c:46-47: C's int abs() on a double truncates toward zero (Q10) */
    const double a0 = (q & OQ_Q10_INT_ABS) ? (double)abs((int)C0n) : fabs(C0n);
    const double a1 = (q & OQ_Q10_INT_ABS) ? (double)abs((int)C1n) : fabs(C1n);
    kon[0] = pow(a0, 2.0) * sqrt(6.0 / M_PI) * f_lm * Iion0;
    kon[1] = (q & OQ_Q6_W2_ZERO) ? 0.0 : pow(a1, 2.0) * sqrt(6.0 / M_PI) * f_lm * Iion1;
    kon[2] = 2.0 * pow(2.0 * Iion0, 3.0 / 2.0);
    kon[3] = 2.0 * pow(2.0 * Iion1, 3.0 / 2.0);
    kon[4] = 2.0 * nq0 - mq - 3.0 / 2.0;
    kon[5] = 2.0 * nq1 - mq - 3.0 / 2.0;
}

/* plasma.c:53-57 and the NaN -> 0 sweep at :184-192 */
void oracle_adk(double *W1, double *W2, long n, const double *kon, const double *Ea) {
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double w1 = kon[0] * pow(kon[2] / Ea[i], kon[4]) * exp(-kon[2] / (3.0 * Ea[i]));
        double w2 = (kon[1] == 0.0) ? 0.0
                  : kon[1] * pow(kon[3] / Ea[i], kon[5]) * exp(-kon[3] / (3.0 * Ea[i]));
        if (isnan(w1)) w1 = 0.0;
        if (isnan(w2)) w2 = 0.0;
        W1[i] = w1;
        W2[i] = w2;
    }
}

/*
 * The reverse-t recurrence for one (x, y) line (plasma.c:218-241).
 * LITERAL (Q8): ions1/ions2 arrays of nt+1 with the reference's read-before-
 *   write order; consequence: ions2 == 0 and F[k] uses R through sample k-1.
 * PHYSICAL: consistent two-species recurrence in the same descending order,
 *   rho_k from samples >= k, with 1-exp(-R) evaluated as -expm1(-R) so tiny
 *   ionisation fractions do not cancel to zero.
 */
void oracle_density_line(const double *b, const double *a, const double *c,
                         const double *sig, int nt, double *F, double *rho_frac, unsigned q) {
    if (q & OQ_Q8_RECURRENCE_ORDER) {
        double *ions1 = (double *)calloc((size_t)nt + 1, sizeof(double));
        double *ions2 = (double *)calloc((size_t)nt + 1, sizeof(double));
        double R = 0.0;
        for (int k = nt - 1; k >= 0; k--) {
            R = R + b[k];
            R = R + (1.0 - R) * a[k];
            ions1[k + 1] = 1.0 - exp(-R) - ions2[k];
            ions2[k + 1] = ions2[k] + c[k] * ions1[k];
        }
        for (int k = 0; k < nt; k++) {
            const double fr = ions1[k] * 1.0 + ions2[k] * 2.0;
            F[k] = fr * sig[k];
            if (rho_frac) rho_frac[k] = fr;
        }
        free(ions1);
        free(ions2);
    } else {
        double R = 0.0, i1 = 0.0, i2 = 0.0;
        for (int k = nt - 1; k >= 0; k--) {
            R = R + b[k];
            R = R + (1.0 - R) * a[k];
            const double u = -expm1(-R);
            const double i2n = i2 + c[k] * i1;
            const double i1n = u - i2n;
            const double fr = i1n + 2.0 * i2n;
            F[k] = fr * sig[k];
            if (rho_frac) rho_frac[k] = fr;
            i1 = i1n;
            i2 = i2n;
        }
    }
}

/* plasma.c:118-246 */
void oracle_ioniz(const ocplx *A, const int *dims, double w0, double deltat,
                  const char *gas, double rho_c, const double *w, double rho_nt,
                  double n0, double *tau_c, double *F, double *rho_frac, unsigned q) {
    const long nxy = (long)dims[0] * dims[1];
    const int nt = dims[2];
    const long N = nxy * nt;
    double *Ip = (double *)malloc(sizeof(double) * (size_t)N);
    double *E0 = (double *)malloc(sizeof(double) * (size_t)N);
    double *Ea = (double *)malloc(sizeof(double) * (size_t)N);
    double *W1 = (double *)malloc(sizeof(double) * (size_t)N);
    double *W2 = (double *)malloc(sizeof(double) * (size_t)N);
    double *sig = (double *)malloc(sizeof(double) * (size_t)N);
    double *Wava = (double *)malloc(sizeof(double) * (size_t)N);

    /* plasma.c:133-149 */
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        const double m = hypot(A[i].re, A[i].im);
        double ip = pow(m, 2) * HBAR * w0;
        if (ip == 0) ip = 1e-30;
        Ip[i] = ip;
        E0[i] = sqrt(2.0 * ZW0 * ip / NC) * FCTR;
        Ea[i] = fabs(E0[i]) / EF;
    }
    const double Kon = deltat / T0;
    const double k0 = w0 / C0;

    /* plasma.c:153-192 */
    double kon[6];
    oracle_adk_consts(gas, q, kon);
    oracle_adk(W1, W2, N, kon, Ea);

    /* plasma.c:195-213 */
    double I1 = 0, I2 = 0, Natm = 0, sigm_coll = 0;
    oracle_ionizpot(gas, &I1, &I2);
    oracle_natoms(gas, &Natm, &sigm_coll);
    const double K1 = 2.0 * pow(EL0, 2.0);
    const double K2 = 4.0 * pow(ME, 2.0);
    const double K3 = sigm_coll * rho_nt;
    const double K4 = k0 / (n0 * rho_c) * w0;
    const double K5 = pow(w0, 2);
    #pragma omp parallel for schedule(static)
    for (long m = 0; m < N; m++) {
        const long k = m / nxy;
        const double wk = (q & OQ_Q12_VE_GRID) ? w[k] : 0.0;
        const double ve = sqrt(K1 * pow(E0[m], 2.0) / (K2 * pow(wk + w0, 2.0)));
        const double tc = 1.0 / (K3 * ve);
        tau_c[m] = tc;
        sig[m] = K4 * tc / (1.0 + K5 * pow(tc, 2.0));
        /* Q11: the reference divides by I1 in eV and later multiplies by
         * Kon = dt/T0; PHYSICAL converts eV to J and uses dt in seconds */
        Wava[m] = (q & OQ_Q11_AVALANCHE_UNITS) ? sig[m] * Ip[m] / I1 : sig[m] * Ip[m] / (I1 * EL0);
    }
    const double acoef = (q & OQ_Q11_AVALANCHE_UNITS) ? Kon : deltat;

    /* plasma.c:218-241, line by line over (x, y) */
    #pragma omp parallel
    {
        double *bb = (double *)malloc(sizeof(double) * (size_t)nt * 6);
        double *aa = bb + nt, *cc = aa + nt, *ss = cc + nt, *ff = ss + nt, *rr = ff + nt;
        #pragma omp for schedule(static)
        for (long i = 0; i < nxy; i++) {
            for (int k = 0; k < nt; k++) {
                const long m = (long)k * nxy + i;
                bb[k] = W1[m] * Kon;
                aa[k] = Wava[m] * acoef;
                cc[k] = W2[m] * Kon;
                ss[k] = sig[m];
            }
            oracle_density_line(bb, aa, cc, ss, nt, ff, rr, q);
            for (int k = 0; k < nt; k++) {
                const long m = (long)k * nxy + i;
                F[m] = ff[k];
                if (rho_frac) rho_frac[m] = rr[k];
            }
        }
        free(bb);
    }
    free(Ip); free(E0); free(Ea); free(W1); free(W2); free(sig); free(Wava);
}
