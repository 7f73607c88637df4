/*
 * oracle.h -- FP64 CPU restatement of kmecseki/MPC's split-step propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (mpc_amd/, include/)
 * includes, links or calls this code.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load liboracle.so, and only as the
 * checker / CPU baseline, never as the thing measured or shipped.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * The reference cannot be compiled here (its headers pull <cufft.h>,
 * <cuda_runtime.h>, <mpi.h>, <fftw3.h>; none exist in this image), so the
 * oracle is pinned by closed forms, brute force and numpy (tests/).
 *
 * Two arithmetic profiles, selected by a quirks bitmask:
 *   OQ_LITERAL  = every bit set: the reference CPU branch's arithmetic as
 *                 written, defects included;
 *   OQ_PHYSICAL = 0: each defect fixed (one bit per defect).
 *
 * Parity status per function (what pins it) is listed next to each
 * declaration.  No function here is "parity unpinned".
 */
#ifndef MPC_ORACLE_H
#define MPC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ocplx;

/* ---- quirk bits (reference defects reproduced when set) ---------------- */
#define OQ_Q1_NONUNITARY_UPDATE   (1u << 0)  /* propagation.c:325-326 */
#define OQ_Q2_REAL_EXPR           (1u << 1)  /* propagation.c:386-388 */
#define OQ_Q3_EXPR_UNSHIFTED      (1u << 2)  /* propagation.c:158 vs :156 */
#define OQ_Q4_WC_CENTRE           (1u << 3)  /* propagation.c:386 */
#define OQ_Q5_NSTAR_I2            (1u << 4)  /* plasma.c:162 */
#define OQ_Q6_W2_ZERO             (1u << 5)  /* plasma.c:163 (reads Iion[2], OOB) */
#define OQ_Q7_CNL1_2N             (1u << 6)  /* plasma.c:174 */
#define OQ_Q8_RECURRENCE_ORDER    (1u << 7)  /* plasma.c:223-231,238-241 */
#define OQ_Q9_CALIB_RE_ONLY       (1u << 8)  /* MPC.c:213 */
#define OQ_Q10_INT_ABS            (1u << 9)  /* plasma.c:46-47 */
#define OQ_Q11_AVALANCHE_UNITS    (1u << 10) /* plasma.c:211,227 */
#define OQ_Q12_VE_GRID            (1u << 11) /* plasma.c:207 */
#define OQ_Q15_POSI2_INIT         (1u << 12) /* propagation.c:464,478-497 */
#define OQ_Q16_NO_MIRROR_LOSS     (1u << 13) /* propagation.c:334-341 vs :380-382 */
#define OQ_Q17_SAME_Z             (1u << 14) /* MPC.c:326-327,346-347 */
#define OQ_Q21_PLASMA_IM          (1u << 15) /* propagation.c:289-290: Im uses 1+w0*tau_c */
#define OQ_Q22_Z_ADVANCE_NEW_DZ   (1u << 16) /* MPC.c:352,383: z += 2*dzz after setStep changed dzz */

#define OQ_PHYSICAL 0u
#define OQ_LITERAL  0x1FFFFu

/* ---- parameters: the param.bin content (MPC.c:180-203) plus runtime
 *      options that param.bin does not hold ------------------------------ */
typedef struct {
    int    dims[3];          /* Nx, Ny, Nt (x fastest) */
    int    sst_on, plasm;
    int    steps_per_pass;   /* fixed-step mode: steps per cell pass */
    int    npasses;
    double dist, dzz, alpha, Cav, k0, gamma2, w0, deltat;
    double rho_c, rho_nt, n0, puls_e, Ab_M, dr, def_err, dzmin, Esig;
    double mirror_R;         /* 0: reference sign-flip mirror; >0: curved */
    double aperture;         /* 0: none; >0: hard circular aperture radius */
    const double *betas;     /* [Nt] ascending centred */
    const double *wr;        /* [Ny] ascending centred kx */
    const double *w;         /* [Nt] ascending centred Omega */
    const double *r;         /* [Ny] x grid */
    char   gas[16];
    unsigned quirks;
    int    pad_;
} oracle_params;

int oracle_sizeof_params(void);

/* ---- FFT (replaces FFTW guru plans, propagation.c:563-603) -------------
 * Batched 1-D complex DFT, unnormalised, dir=-1: exp(-2 pi i jk/n) (FFTW_FORWARD),
 * dir=+1: exp(+...).  Line of length n with element stride es; batch is a
 * 2-level loop (c0 lines at stride s0) x (c1 at stride s1).  n = 2^a 3^b 5^c.
 * Pinned by: brute-force DFT, numpy.fft.                                    */
void oracle_fft_lines(ocplx *a, int n, long es, long c0, long s0, long c1, long s1, int dir);
/* axis: 0 = x (dims[0]), 1 = y, 2 = t; in place over the whole field */
void oracle_fft_axis(ocplx *A, const int *dims, int axis, int dir);

/* ---- tables --------------------------------------------------------------
 * exp_D0[k] = exp(dzz/2 (i betas[k] - alpha/2)), halves swapped into FFT
 * order (propagation.c:150-156, :393-402).  Pinned by closed form.          */
void oracle_exp_D0(ocplx *exp_D0, const double *betas, double alpha, double dzz, int nt);
/* calcexpR (propagation.c:375-391); dzz is the half step.  Pinned: closed form. */
void oracle_calcexpR(ocplx *expR, int *signmem, double z, double Cav, double k0,
                     double dzz, const double *wr, const int *dims, unsigned q);
/* pass sign s = ((floor(z/Cav)+1) % 2)*2 - 1 (propagation.c:378) */
int  oracle_pass_sign(double z, double Cav);

/* ---- operator pieces ------------------------------------------------------ */
/* linearstep (propagation.c:65-107): t-FFT x exp_D0 t-IFFT, y-FFT x conj(expR[j])
 * y-IFFT, x-FFT x conj(expR[i]) x-IFFT, / N.  Pinned: Gaussian diffraction,
 * dispersion chirp closed forms, brute force 3-D DFT.                       */
void oracle_linearstep(ocplx *A, const ocplx *exp_D0, const ocplx *expR, const int *dims);
/* self-steepening S (propagation.c:164-240).  Pinned: analytic d/dt of a Gaussian. */
void oracle_sst(const ocplx *A, ocplx *S, const int *dims, const double *w,
                double gamma2, double w0, int sst_on);

/* ---- plasma (plasma.c) ---------------------------------------------------- */
void oracle_ionizpot(const char *gas, double *I1, double *I2);        /* plasma.c:4-40 */
void oracle_natoms(const char *gas, double *Natm, double *sigm_coll); /* plasma.c:61-115 */
/* ADK constants Kon1..Kon6 (plasma.c:45-51, with :153-174), k[0..5]       */
void oracle_adk_consts(const char *gas, unsigned q, double *kon);
/* ADK rate sweep (plasma.c:43-58) + NaN->0 (plasma.c:184-192) */
void oracle_adk(double *W1, double *W2, long n, const double *kon, const double *Ea);
/* ioniz (plasma.c:118-246).  Outputs tau_c[N], F[N] (freeelectrons_sp) and
 * optionally rho_frac[N] = ions1 + 2 ions2 (may be NULL).
 * Pinned: ADK/tau_c/sigma closed forms; recurrence brute force + special cases. */
void oracle_ioniz(const ocplx *A, const int *dims, double w0, double deltat,
                  const char *gas, double rho_c, const double *w, double rho_nt,
                  double n0, double *tau_c, double *F, double *rho_frac, unsigned q);
/* the reverse-t recurrence alone on per-sample coefficients (plasma.c:218-241):
 * b = W1*Kon, a = avalanche coefficient, c = W2*Kon, sig = sigma.  Output
 * F (len nt) and rho_frac (len nt, may be NULL).  One line, stride 1.       */
void oracle_density_line(const double *b, const double *a, const double *c,
                         const double *sig, int nt, double *F, double *rho_frac, unsigned q);

/* ---- one split step (propagation.c:111-372) -------------------------------- */
void oracle_ssfmprop(ocplx *A, const oracle_params *p, double dzz, int *signmem, double z,
                     double *puls_e, double *BP1, double *BP2, double *y1, double *y2,
                     double *bps1, double *bps2, int change);
/* the nonlinear half alone: A <- A exp(dzz A_nl) (propagation.c:164-327) */
void oracle_nonlinear(ocplx *A, const oracle_params *p, double dzz, double *rho_peak);

/* rho_nt * max(ions1 + 2 ions2) seen by the most recent nonlinear step */
double oracle_last_rho_peak(void);

/* mirror at a pass boundary (curvature phase x aperture x sqrt(Ab_M)). */
void oracle_mirror(ocplx *A, const oracle_params *p);

/* ---- diagnostics / control ------------------------------------------------- */
/* calc1oes (propagation.c:460-504).  Pinned: brute force + tie rule.       */
void oracle_calc1oes(const double *BP, const int *dims, const double *r, double *vald1oes,
                     double *y, unsigned q);
/* BP1[i] = sum_{j,k}|A|^2, BP2[j] = sum_{i,k}|A|^2 (propagation.c:343-354) */
void oracle_beam_profiles(const ocplx *A, const int *dims, double *BP1, double *BP2);
/* temp[k] = sum_xy |A|^2 ; spec[k] = sum_xy |FFT_t A|^2 (MPC.c:384-449, kernels.cu:13-24) */
void oracle_temp_spec(const ocplx *A, const int *dims, double *temp, double *spec);
/* setStep (propagation.c:616-677): returns via pointers; err printed if verbose */
void oracle_setstep(const int *dims, double *dzz, const ocplx *C1, const ocplx *C2,
                    double def_err, double dzmin, const double *y1, const double *y2,
                    int *bad, double *err);
/* setStep decision on a given err (no reduction) */
void oracle_setstep_decide(double *dzz, double err, double def_err, double dzmin,
                           const double *y1, const double *y2, int *bad);
/* Richardson A = 4/3 C2 - 1/3 C1 (MPC.c:377-381) */
void oracle_richardson(ocplx *A, const ocplx *C1, const ocplx *C2, long n);
/* energy calibration (MPC.c:206-216) */
void oracle_calibrate(ocplx *A, long n, double puls_e, double Esig, unsigned q);

/* ---- drivers (MPC.c:269-555 without MPI) ----------------------------------- */
/* fixed-step run: nsteps split steps starting at step index step0; mirror at
 * pass boundaries (between passes).  rho_peak (may be NULL) receives
 * rho_nt*max(ions1+2 ions2) inside the last step's nonlinear operator.     */
void oracle_run_fixed(ocplx *A, const oracle_params *p, int step0, int nsteps, double *rho_peak);
/* adaptive run (step doubling + Richardson) for at most max_steps accepted
 * steps or until z >= dist.  Returns accepted steps; *z, *dzz updated.     */
int  oracle_run_adaptive(ocplx *A, const oracle_params *p, double *z, double *dzz,
                         int *signmem, int max_steps, double *errs, double *dzs);

void oracle_set_threads(int n);
int  oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
