/*
 * fft.c -- oracle FFT: batched 1-D complex DFT in FP64 (test infrastructure).
 *
 * Restates the transforms the reference obtains from FFTW guru plans
 * (propagation.c:563-603; executed at :67,76,77,88,89,100,216,223 and
 * MPC.c:470): unnormalised, FFTW_FORWARD = exp(-2 pi i jk/n), FFTW_BACKWARD
 * = exp(+2 pi i jk/n).  FFTW itself is not in this image and is not
 * vendored by the reference; the transform is fixed by the DFT definition,
 * which tests/ check by brute force and against numpy.fft.
 *
 * Algorithm: mixed-radix (2,3,4,5) Stockham autosort on a contiguous copy
 * of each line, lines distributed over OpenMP threads.  Each line is
 * independent, so results do not depend on the thread count.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int n, nrad, rad[40];
    ocplx *tw;   /* tw[m] = exp(-2 pi i m / n) */
} oplan;

static int build_plan(oplan *P, int n) {
    P->n = n;
    P->nrad = 0;
    int m = n;
    while (m % 4 == 0 && m > 4) { P->rad[P->nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { P->rad[P->nrad++] = 2; m /= 2; }
    while (m % 3 == 0) { P->rad[P->nrad++] = 3; m /= 3; }
    while (m % 5 == 0) { P->rad[P->nrad++] = 5; m /= 5; }
    if (m != 1) return -1;
    P->tw = (ocplx *)malloc(sizeof(ocplx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        /* exact reduction of the angle fraction before calling sin/cos */
        double a = -2.0 * M_PI * (double)k / (double)n;
        P->tw[k].re = cos(a);
        P->tw[k].im = sin(a);
    }
    return 0;
}

static inline ocplx cmul(ocplx a, ocplx b) {
    ocplx c = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return c;
}

/* one line, contiguous, in place via scratch (len n) */
static void fft_line(const oplan *P, ocplx *x, ocplx *scr, int dir) {
    const int n = P->n;
    ocplx *in = x, *out = scr;
    int Ns = 1;
    const double s = (double)dir; /* -1 forward, +1 backward */
    for (int st = 0; st < P->nrad; st++) {
        const int R = P->rad[st];
        const int nR = n / R;
        const int tstep = n / (Ns * R);
        for (int j = 0; j < nR; j++) {
            const int k = j % Ns;
            ocplx v[5];
            for (int r = 0; r < R; r++) {
                ocplx a = in[j + r * nR];
                if (r && k) {
                    /* table holds exp(-i..): conjugate it for dir = +1 */
                    ocplx t = P->tw[(k * r * tstep) % n];
                    if (dir > 0) t.im = -t.im;
                    a = cmul(a, t);
                }
                v[r] = a;
            }
            ocplx o[5];
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                ocplx a0 = { v[0].re + v[2].re, v[0].im + v[2].im };
                ocplx a1 = { v[0].re - v[2].re, v[0].im - v[2].im };
                ocplx a2 = { v[1].re + v[3].re, v[1].im + v[3].im };
                ocplx a3 = { v[1].re - v[3].re, v[1].im - v[3].im };
                /* a3 * (-i) for the forward transform, a3 * (+i) for backward */
                ocplx a3r;
                a3r.re = (dir < 0) ? a3.im : -a3.im;
                a3r.im = (dir < 0) ? -a3.re : a3.re;
                o[0].re = a0.re + a2.re;  o[0].im = a0.im + a2.im;
                o[2].re = a0.re - a2.re;  o[2].im = a0.im - a2.im;
                o[1].re = a1.re + a3r.re; o[1].im = a1.im + a3r.im;
                o[3].re = a1.re - a3r.re; o[3].im = a1.im - a3r.im;
            } else {
                /* generic small-R DFT (R = 3, 5) */
                for (int q = 0; q < R; q++) {
                    double re = 0, im = 0;
                    for (int r = 0; r < R; r++) {
                        double ang = s * 2.0 * M_PI * (double)((q * r) % R) / (double)R;
                        double c = cos(ang), sn = sin(ang);
                        re += v[r].re * c - v[r].im * sn;
                        im += v[r].re * sn + v[r].im * c;
                    }
                    o[q].re = re; o[q].im = im;
                }
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) out[base + r * Ns] = o[r];
        }
        ocplx *t = in; in = out; out = t;
        Ns *= R;
    }
    if (in != x) memcpy(x, in, sizeof(ocplx) * (size_t)n);
}

void oracle_fft_lines(ocplx *a, int n, long es, long c0, long s0, long c1, long s1, int dir) {
    oplan P;
    if (build_plan(&P, n) != 0) abort();
    /* lines adjacent in memory (s0 == 1) are gathered in blocks of B so each
     * cache line fetched from a strided row is used B times */
    const long B = (s0 == 1 && es != 1) ? 16 : 1;
    const long nb0 = (c0 + B - 1) / B;
    const long nblocks = nb0 * c1;
    #pragma omp parallel
    {
        ocplx *buf = (ocplx *)malloc(sizeof(ocplx) * (size_t)n * (size_t)(B + 1));
        ocplx *scr = buf + (size_t)n * (size_t)B;
        #pragma omp for schedule(static)
        for (long blk = 0; blk < nblocks; blk++) {
            const long ib = blk % nb0, i1 = blk / nb0;
            const long i0 = ib * B;
            const long nb = (c0 - i0 < B) ? (c0 - i0) : B;
            ocplx *base = a + i0 * s0 + i1 * s1;
            for (int t = 0; t < n; t++)
                for (long b = 0; b < nb; b++) buf[b * n + t] = base[b * s0 + (long)t * es];
            for (long b = 0; b < nb; b++) fft_line(&P, buf + b * n, scr, dir);
            for (int t = 0; t < n; t++)
                for (long b = 0; b < nb; b++) base[b * s0 + (long)t * es] = buf[b * n + t];
        }
        free(buf);
    }
    free(P.tw);
}

void oracle_fft_axis(ocplx *A, const int *dims, int axis, int dir) {
    const long nx = dims[0], ny = dims[1], nt = dims[2];
    if (axis == 0)      /* propagation.c:563-575: n=Nx, stride 1; {Ny:Nx},{Nt:NxNy} */
        oracle_fft_lines(A, (int)nx, 1, ny, nx, nt, nx * ny, dir);
    else if (axis == 1) /* propagation.c:577-589: n=Ny, stride Nx; {Nx:1},{Nt:NxNy} */
        oracle_fft_lines(A, (int)ny, nx, nx, 1, nt, nx * ny, dir);
    else                /* propagation.c:591-603: n=Nt, stride NxNy; {Nx:1},{Ny:Nx} */
        oracle_fft_lines(A, (int)nt, nx * ny, nx, 1, ny, nx, dir);
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
