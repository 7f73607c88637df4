/*
 * mpc.h -- device-resident API of libmpc_hip.so.
 *
 * The reference has no counterpart for these entry points: its propagator
 * copies the whole field host<->device on every call (propagation.c:6-8,59,
 * 180-210).  Here the field stays in HBM for a whole run; the host sees it
 * only at upload, dumps and checkpoints.  The legacy entry points in
 * propagation.h / plasma.h are thin wrappers over this API.
 *
 * All functions return 0 on success and a negative MPC_E* code on error;
 * mpc_strerror() gives a message.  Nothing here falls back to the CPU.
 */
#ifndef MPC_MPC_H
#define MPC_MPC_H
#include "settings.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPC_OK          0
#define MPC_EINVAL    (-1)
#define MPC_ENOMEM    (-2)
#define MPC_EHIP      (-3)
#define MPC_ENODEV    (-4)
#define MPC_EUNSUPP   (-5)
#define MPC_EIO       (-6)
#define MPC_ENONFINITE (-7)  /* the guard found a NaN or Inf in the field */
#define MPC_EGUARD    (-8)   /* the guard's energy or spectral-width check failed */

#define MPC_F64 0   /* plain FP64 path */
#define MPC_F32 1   /* FP32 complex with FP64 islands */

/*
 * Run parameters: the content of param.bin (reference MPC.c:180-203, same
 * meaning and units) plus runtime options param.bin does not hold.
 * Arrays are host pointers, read at mpc_create().
 */
typedef struct {
    int    dims[3];          /* Nx, Ny, Nt; Nx == Ny (reference propagation.c:83,95,386) */
    int    sst_on, plasm;
    int    steps_per_pass;   /* fixed-step schedule: z-steps per cell pass */
    int    npasses;
    double dist, dzz, alpha, Cav, k0, gamma2, w0, deltat;
    double rho_c, rho_nt, n0, puls_e, Ab_M, dr, def_err, dzmin, Esig;
    double mirror_R;         /* 0: reference sign-flip mirror; >0: curved mirror radius (m) */
    double aperture;         /* 0: none; >0: hard circular aperture radius (m) */
    const double *betas;     /* [Nt] dispersion on the ascending centred Omega grid */
    const double *wr;        /* [Ny] ascending centred kx grid */
    const double *w;         /* [Nt] ascending centred Omega grid */
    const double *r;         /* [Ny] transverse x grid */
    char   gas[16];          /* "Argon", "Helium", "Neon", "Air", "Nitrogen", "Xenon", "Krypton" */
    unsigned quirks;         /* must be 0 on the device (the corrected arithmetic) */
    int    pad_;
} mpc_params;

/* Diagnostics of the resident field (reference MPC.c:384-512 outputs plus
 * per-pass observables).  Arrays are host pointers the caller sizes; any may
 * be NULL. */
typedef struct {
    double energy_sum;       /* sum |A|^2 */
    double rho_peak;         /* rho_nt * max(ions1 + 2 ions2) of the last nonlinear step */
    double spec_rms_width;   /* rms width of spec over the FFT-order Omega grid (rad/s) */
    double *temp;            /* [Nt] sum_xy |A|^2 */
    double *spec;            /* [Nt] sum_xy |FFT_t A|^2, FFT order, unnormalised */
    double *BP1;             /* [Nx] sum_{y,t} |A|^2 */
    double *BP2;             /* [Ny] sum_{x,t} |A|^2 */
} mpc_diag;

typedef struct mpc_ctx_s mpc_ctx;

const char *mpc_strerror(int code);
int  mpc_last_error(void);               /* sticky error of the legacy void entry points */
int  mpc_sizeof_params(void);
int  mpc_device_count(void);

int  mpc_create(const mpc_params *p, int precision, mpc_ctx **out);
void mpc_destroy(mpc_ctx *ctx);
/* field transfer, reference layout A[t][y][x] (x fastest) */
int  mpc_upload(mpc_ctx *ctx, const cufftDoubleComplex *A);
int  mpc_download(mpc_ctx *ctx, cufftDoubleComplex *A);
/*
 * Asynchronous checkpoint snapshot (the reference downloads and writes its
 * checkpoint synchronously, MPC.c:521-544).  mpc_snapshot_begin queues the
 * resident field's conversion to the reference layout (a t-spectral field is
 * first brought back to the time domain in place, so a run resumed from the
 * snapshot and the uninterrupted run go on from the same field) and its
 * copy into a pinned host buffer the context owns, then returns at once;
 * runs started afterwards overlap the copy and do not change the snapshot.
 * mpc_snapshot_poll returns 1 once the copy is complete (or none is
 * pending), 0 while it runs.  mpc_snapshot_wait blocks until it is complete
 * and gives the pinned buffer (N points, A[t][y][x]); it stays valid and
 * unchanged until the next mpc_snapshot_begin or mpc_destroy, so the caller
 * must be done reading it before it begins another snapshot.
 */
int  mpc_snapshot_begin(mpc_ctx *ctx);
int  mpc_snapshot_poll(mpc_ctx *ctx);
int  mpc_snapshot_wait(mpc_ctx *ctx, const cufftDoubleComplex **A);
/* scale the resident field by sqrt(puls_e/Esig) (reference MPC.c:206-216,
 * both parts: the reference scales Re only) */
int  mpc_calibrate(mpc_ctx *ctx);
/*
 * Fixed-step run of nsteps z-steps starting at global step index step0:
 * each step is L(dzz/2) N(dzz) L(dzz/2) evaluated at z = (n + 1/2) dzz,
 * with the mirror applied between passes.  Consecutive half steps are merged
 * (3 HBM sweeps per interior step).  On return the field is t-spectral (the
 * state it keeps between steps), so a following run or mpc_diagnostics
 * starts without a transform.  mpc_download and mpc_export_planar read a
 * transformed copy and leave the field as it is; mpc_snapshot_begin,
 * mpc_step and mpc_linear first bring it back to the time domain in place.
 */
int  mpc_run(mpc_ctx *ctx, int step0, int nsteps);
/* One isolated split step at z with step dzz (reference ssfmprop semantics,
 * signmem updated as calcexpR does). */
int  mpc_step(mpc_ctx *ctx, double dzz, double z, int *signmem);
/* mirror (curvature x aperture x sqrt(Ab_M)) applied once to the field */
int  mpc_mirror(mpc_ctx *ctx);
/* Linear propagation only, over `length` (m) in one go: dispersion and
 * absorption on the t-lines, diffraction with coefficient sign * length on x
 * and y, then the mirror if `mirror` != 0.  Equals the fixed-step schedule of
 * mpc_run with the nonlinear step removed (its linear factors commute and
 * merge).  Time-domain field in (a t-spectral one is brought back first)
 * and out.  No reference counterpart: the pass surrogate takes it as the
 * physics prior that its network corrects. */
int  mpc_linear(mpc_ctx *ctx, double length, int sign, int mirror);
/* temp, spec, BP1/BP2, energy, rms spectral width and peak rho of the resident
 * field, in the time or the t-spectral domain as it is */
int  mpc_diagnostics(mpc_ctx *ctx, mpc_diag *d);
/* the resident field, in the time domain (a t-spectral one is read through a
 * transformed copy and left as it is), as two float planes (re then im, N points
 * each, point order [y][x][t], t fastest) in a device buffer, times scale;
 * and the reverse (the pass surrogate hands its prediction back this way) */
int  mpc_export_planar(mpc_ctx *ctx, float *planes, double scale);
int  mpc_import_planar(mpc_ctx *ctx, const float *planes, double scale);
/* adaptive step doubling + Richardson (reference MPC.c:343-383,
 * propagation.c:616-677) for at most max_steps accepted steps.  z, dzz and
 * signmem are updated; errs/dzs (may be NULL) receive per accepted step.
 * The attempts run on the t-spectral field and the call leaves it so, as
 * mpc_run does. */
int  mpc_run_adaptive(mpc_ctx *ctx, double *z, double *dzz, int *signmem, int max_steps,
                      int *accepted, double *errs, double *dzs);

/*
 * setStep (propagation.h:14, propagation.c:616-677) on device-resident
 * fields, allocating nothing: no counterpart in the reference; it is for
 * callers that cannot allocate (code inside stream capture, serving loops,
 * harnesses that forbid hipMalloc).  C1 and C2 are DEVICE pointers to the
 * dims[0]*dims[1]*dims[2] complex points of the two fields (any point order,
 * the same in both); the other arguments are host pointers with setStep's
 * meaning.  The partial sums live in a fixed 64 KB device array the library
 * holds per GPU until the process ends, which bounds a field to 2^28 points
 * (MPC_EINVAL beyond).  Runs on the null stream of the current device and
 * returns once *err is known; nothing is printed.  Calls that overlap on
 * one GPU share that array and race.
 */
int  mpc_setstep_noalloc(const int *dims, double *dzz, const cufftDoubleComplex *C1, const cufftDoubleComplex *C2,
                         double def_err, double dzmin, const double *y1, const double *y2, int *bad, double *err);

/*
 * Solver guard (no counterpart in the reference, whose only guards are step
 * rejection and NaN -> 0 in the ionisation rates, MPC.c:370-373,
 * plasma.c:184-192).  mpc_guard_verdict judges a stretch of propagation from
 * the energies and rms spectral widths before and after it:
 *   MPC_ENONFINITE  e_after is NaN or infinite (a single non-finite field
 *                   value makes the energy so), or w_after is when checked;
 *   MPC_EGUARD      the energy grew by more than gain_tol (relative), or
 *                   moved by more than energy_tol either way, or the width
 *                   by more than width_tol;
 *   0               otherwise.  A tolerance < 0 skips its check.
 * mpc_set_guard(ctx, gain_tol) makes mpc_run apply it to the resident
 * field's energy (one extra read of the field before and after each call;
 * gain_tol < 0, the default, turns it off).  mpc_run then returns the
 * verdict's code and leaves the field as the run left it.  The pass
 * surrogate's fallback uses the same verdict with energy_tol and width_tol.
 */
int  mpc_guard_verdict(double e_before, double e_after, double w_before, double w_after, double gain_tol,
                       double energy_tol, double width_tol);
int  mpc_set_guard(mpc_ctx *ctx, double gain_tol);
/* sum |A|^2 of the resident field (time or t-spectral domain as it is) */
int  mpc_energy(mpc_ctx *ctx, double *energy_sum);

/* FFT/tile autotuner: times every kernel variant (FFT radix plan, tile
 * width, elements per thread) on a scratch copy of the resident field and
 * keeps the fastest in the context (call after mpc_upload).  The variants
 * compute the same transform; the choice only changes speed.  The
 * MPC_TLINE_VARIANT / MPC_AXIS_VARIANT environment variables pin them. */
int  mpc_autotune(mpc_ctx *ctx);
int  mpc_get_variants(mpc_ctx *ctx, int *tline_variant, int *axis_variant);   /* axis: the x-line's */
int  mpc_set_variants(mpc_ctx *ctx, int tline_variant, int axis_variant);    /* same variant for x and y */
/* per-axis form: the x- and y-line sweeps are tuned separately (the y lines'
 * stride is Nx times the x lines', so their best tile order differs) */
int  mpc_get_variants_xy(mpc_ctx *ctx, int *tline_variant, int *x_variant, int *y_variant);
int  mpc_set_variants_xy(mpc_ctx *ctx, int tline_variant, int x_variant, int y_variant);
/* number of variants of each kind: t-line, and x/y-line (plan + 4 * tile order) */
int  mpc_variant_counts(int *tline_variants, int *axis_variants);

/*
 * Batched runs (no counterpart in the reference): pulses of uneven grids,
 * energies and media, one context each (same precision, steps_per_pass and
 * npasses), advance together.  Pulses of one size class are split into
 * launch groups of at most 256 MB of fields, and every sweep of a step is one
 * launch per group.  Within a pass each group runs all its sweeps before the
 * next group starts, so a pulse's sweep finds part of what its previous
 * sweep wrote still in the last-level cache (when the sweep kinds group the
 * pulses differently, all groups' launches of one sweep run first instead).
 * With use_graph the launches of each pass are captured once into a HIP
 * graph and replayed for every pass with the same launch list.  Each launch
 * takes its group's first pulse's kernel variants (mpc_set_variants_xy /
 * mpc_autotune): the persistent variants walk the line groups or tiles of
 * all pulses of the group with the workgroups that fit on the device at
 * once, the others take one workgroup per group or tile.  Each pulse's result is bit-identical to mpc_run on its
 * context with those variants.  The batch does not own the contexts; destroy
 * it before them.
 */
typedef struct mpc_batch_s mpc_batch;
int  mpc_batch_create(mpc_ctx *const *ctxs, int n, mpc_batch **out);
void mpc_batch_destroy(mpc_batch *b);
int  mpc_batch_run(mpc_batch *b, int step0, int nsteps, int use_graph);
int  mpc_batch_graphs(mpc_batch *b);   /* distinct pass graphs captured so far */
/* Batch-aware tuning (no counterpart in the reference, whose FFTW plans are
 * per size): times the batch's own launches and sets every pulse of a launch
 * class to the fastest variant: the t-line form (persistent or one workgroup
 * per line group) per t length and plasma switch, and the x/y variants per
 * line length where a launch covers several pulses (a launch over one pulse
 * runs the per-field kernel, which the context's own tuning chose).  A
 * variant faster on one pulse alone can be slower over a group of many.
 * Runs on scratch copies of the fields; fields and diagnostics are left as
 * they were.  Call after the contexts' own mpc_autotune. */
int  mpc_batch_autotune(mpc_batch *b);

/* Timing of the most recent mpc_run: average device time per launch of each
 * sweep kernel, measured with HIP events on the library's stream.
 * which: 0 = t-line, 1 = x-line, 2 = y-line.  Returns ms, or <0. */
int  mpc_set_timing(mpc_ctx *ctx, int enable);
double mpc_kernel_ms(mpc_ctx *ctx, int which);
long long mpc_field_bytes(mpc_ctx *ctx);

/*
 * Input generator (synthetic workloads): fills the param.bin content and the
 * unscaled input field for a Gaussian pulse in x, y, t with seeded complex
 * Gaussian noise (SplitMix64 + Box-Muller).  Arrays are caller-allocated:
 * betas[Nt], wr[Ny], w[Nt], r[Ny], A[N] in the reference layout A[t][y][x].
 */
typedef struct {
    int    nx, ny, nt;
    int    passes, steps_per_pass;
    int    sst_on, plasm;
    int    pad_;
    double window_xy;        /* m, full transverse window */
    double window_t;         /* s, full time window */
    double lambda0;          /* m */
    double waist;            /* m, 1/e field radius exp(-(x^2+y^2)/waist^2) */
    double tau_fwhm;         /* s, intensity FWHM */
    double energy;           /* J */
    double noise_rel;        /* noise sigma per component / max|A| */
    double n2_per_bar;       /* m^2/W at 1 bar */
    double beta2;            /* s^2/m */
    double pressure;         /* bar */
    double rho_nt_per_bar;   /* m^-3 at 1 bar */
    double cell_L;           /* m */
    double mirror_R, aperture, Ab_M, def_err, dzmin;
    unsigned long long seed;
    char   gas[16];
} mpc_input_spec;

/* ADK constants Kon1..Kon6 of a gas (plasma.c:45-51 with the quantum
 * numbers of :159-174, corrected arithmetic); 0 if the gas is known */
int mpc_adk_constants(const char *gas, double *kon);

int mpc_mkinput(const mpc_input_spec *s, mpc_params *p, double *betas, double *wr,
                double *w, double *r, cufftDoubleComplex *A);
int mpc_sizeof_input_spec(void);

/* param.bin / field file I/O (reference MPC.c:76-77,159-168,180-203 formats) */
int mpc_write_param_bin(const char *path, const mpc_params *p);
/* reads param.bin; arrays are malloc'd here and returned through p (free
 * with mpc_free_params) */
int mpc_read_param_bin(const char *path, mpc_params *p);
void mpc_free_params(mpc_params *p);

#ifdef __cplusplus
}
#endif
#endif
