/*
 * propagation.h -- the propagator entry points the reference driver calls
 * (reference propagation.h:7-14).  Same names, argument order and meaning.
 * All field pointers are HOST pointers in the reference layout A[t][y][x]
 * (x fastest); each call uploads, runs on the GPU, and downloads.  The
 * device-resident path for production runs is in mpc.h.
 *
 * sendmpi/recmpi (reference propagation.h:15-16) are not provided: this
 * build runs on one GPU with no MPI task pipeline.
 *
 * Errors: as in the reference these return void and print to stderr; in
 * addition the library keeps a sticky error flag, read with mpc_last_error().
 */
#ifndef MPC_PROPAGATION_H
#define MPC_PROPAGATION_H
#include "settings.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces reference propagation.h:7 (propagation.c:375-391): transverse
 * propagator table for a half step dzz; sets *signmem to the pass sign. */
void calcexpR(cufftDoubleComplex *expR, int *signmem, double z, double Cav, double k0,
              double dzz, double *wr, int *dims);
/* replaces reference propagation.h:8 (propagation.c:393-457): swap halves
 * along axis 0 (a length-dims[2] table), 1 (x), 2 (y) or 3 (t) */
void fftshift(cufftDoubleComplex *in, int *dims, cufftDoubleComplex *buffer, int axis);
/* replaces reference propagation.h:9 (propagation.c:460-504): 1/e^2 width */
void calc1oes(double *BP, int *dims, double *r, double *vald1oes, double *y);
/* replaces reference propagation.h:10 (propagation.c:111-372): one split
 * step L(dzz/2) N(dzz) L(dzz/2) on A in place */
void ssfmprop(cufftDoubleComplex *A, int *dims, int sst_on, double dzz, double *betas,
              double alpha, int *signmem, double z, double Cav, double k0, double *wr,
              fftw_plan pR1fft, fftw_plan pR1ifft, fftw_plan pR2fft, fftw_plan pR2ifft,
              fftw_plan pTfft, fftw_plan pTifft, double *w, double gamma2, double w0,
              int plasm, double deltat, char *gas, double rho_c, double rho_nt, double n0,
              double *puls_e, double *r, double *BP1, double *BP2, double *y1, double *y2,
              double *bps1, double *bps2, double Ab_M, int change, int mpirank);
/* replaces reference propagation.h:11 (propagation.c:3-108): t-FFT x exp_D0
 * x t-IFFT, y-FFT x conj(expR[j]) x y-IFFT, x-FFT x conj(expR[i]) x x-IFFT, /N */
void linearstep(cufftDoubleComplex *A, cuDoubleComplex *exp_D0, cuDoubleComplex *expR,
                fftw_plan pTfft, fftw_plan pR2fft, fftw_plan pR1fft, int *dims);
/* replaces reference propagation.h:13 (propagation.c:508-613): creates the
 * six plan handles for dims; readin != 0 reads the autotuner cache
 * "<cpath><dims[0]>" (the reference's FFTW wisdom file) and writes it when
 * absent. */
void createPlans(int readin, int *dims, const char *cpath, fftw_plan *pR1fft,
                 fftw_plan *pR1ifft, fftw_plan *pR2fft, fftw_plan *pR2ifft,
                 fftw_plan *pTfft, fftw_plan *pTifft);
/* replaces reference propagation.h:14 (propagation.c:616-677) */
void setStep(int *dims, double *dzz, cufftDoubleComplex *C1, cufftDoubleComplex *C2,
             double def_err, double dzmin, double *y1, double *y2, int *bad, double *err);

#ifdef __cplusplus
}
#endif
#endif
