/*
 * internal.h -- device context and kernel launchers shared by the library's
 * translation units.  Not part of the public ABI (the headers under include/ are).
 *
 * HBM layout of a resident field: A[y][x][t] (t fastest, "line-major"),
 * complex<T> interleaved, T = double (MPC_F64) or float (MPC_F32).  Every
 * t-line is contiguous (16 KB FP32 / 32 KB FP64 at Nt = 2048); x- and
 * y-lines are read as tiles of G consecutive t values (256 B per row).
 * The reference layout A[t][y][x] is converted at upload/download only.
 */
#ifndef MPC_INTERNAL_H
#define MPC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <vector>
#include "../../include/mpc.h"
#include "fft_core.h"

namespace mpc {

/* FP64 constants of the plasma model (plasma.c:141-212), evaluated once.
 * With s = sqrt(2 ZW0 I / NC) (so E0 = FCTR s) every per-element quantity
 * is a product with s or 1/s:
 *   tau_c = kt / s,  sigma = K4 tau_c / (1 + K5 tau_c^2),
 *   ADK exponent Kon3 / (3 Ea) = ke1 / s,  log(Kon3 / Ea) = L1 + log(1/s). */
struct PlasmaK {
    double zf;           /* 2 ZW0 / NC */
    double cs;           /* zf * HBAR w0 (scaled by the field scale^2 at launch) */
    double kt;           /* 1 / (K3 FCTR sqrt(K1/K2) / w0) */
    double ke1, ke2;     /* Kon3 / (3 FCTR / EF), Kon4 / (3 FCTR / EF) */
    double L1, L2;       /* log(Kon3 EF / FCTR), log(Kon4 EF / FCTR) */
    double hbar_w0;      /* HBAR * w0 */
    double inv_ef;       /* 1 / EF */
    double kon1, kon2, kon3, kon4, kon5, kon6;
    double K1, K2, K3, K4, K5;
    double w0;
    double ve_fac;       /* sqrt(K1 / K2) / w0: quiver velocity per unit field */
    double I1e;          /* I1[eV] * EL0 */
    double Kon;          /* deltat / T0 */
    double deltat;
    double dt_I1e;       /* deltat / I1e: the avalanche rate's factor */
    double rho_nt;
    double sig_bound;    /* K4 / (2 sqrt K5) + w0 K4 / K5 >= sigma (1 + w0 tau) for every tau */
};

template <class T> struct TArgs {
    cx<T> *field;
    const cx<T> *tw;     /* Nt twiddles exp(-2 pi i k / Nt) */
    const cx<T> *P;      /* pre table, FFT order, includes 1/Nt; nullptr: already applied */
    const cx<T> *Q;      /* post table (only with out_time) or nullptr */
    const T *wfft;       /* FFT-order Omega grid */
    double *rho_block;   /* per-block max(ions1 + 2 ions2), may be nullptr */
    PlasmaK pk;
    T dz, gamma, sst_coef; /* sst_coef = gamma2 / w0 */
    int in_freq, do_nl, out_time, sst_on;
    long nlines;
};

template <class T> struct AArgs {
    cx<T> *field;
    const cx<T> *src;    /* read from here instead of field (same layout), or nullptr: in place */
    const cx<T> *tw;     /* n twiddles */
    const cx<T> *D;      /* n-entry propagator (FFT order), includes 1/n */
    const cx<T> *M;      /* mirror table M[y*Nx + x] or nullptr */
    long s_line;         /* element stride along the line */
    long s_outer;        /* stride between tiles' outer index */
    long m_line, m_outer;/* mirror index = i*m_line + outer*m_outer */
    const cx<T> *Pt;     /* per-t factor applied before the store (t-spectral field), or nullptr */
    int nt;              /* contiguous t extent */
    int n_outer;         /* number of lines per t value */
    int tg;              /* tile order: runs of tg t-chunks, then the outer index */
};

/* launchers (kernels.hip) */
/* nblocks (may be nullptr) receives the grid size, i.e. how many rho_block
 * entries the launch writes */
template <class T> int launch_tline(const TArgs<T> &a, int nt, bool plasma, int variant, hipStream_t s, int *nblocks);
template <class T> int launch_axis(const AArgs<T> &a, int n, int outer, int variant, hipStream_t s);
int tline_variants(void);
int axis_variants(void);
template <class T> int launch_to_internal(const double *src, cx<T> *dst, int nx, int ny, int nt, double scale,
                                          hipStream_t s);
template <class T> int launch_to_reference(const cx<T> *src, double *dst, int nx, int ny, int nt, double scale,
                                           hipStream_t s);
int launch_maxabs(const double *v, long n, double *part, int nb, hipStream_t s);
template <class T> int launch_scale(cx<T> *f, long n, T s, hipStream_t st);
template <class T> int launch_planar(cx<T> *f, float *planes, long n, double s, int to_planes, hipStream_t st);
template <class T> int launch_diag_lines(const cx<T> *f, const cx<T> *tw, int nt, long nlines,
                                         double *line_power, double *temp_part, double *spec_part,
                                         int nblocks, bool spectral, hipStream_t s);
int launch_reduce_cols(const double *part, int nrows, int ncols, double *out, hipStream_t s);
template <class T> int launch_err_partials(const cx<T> *c1, const cx<T> *c2, long n, double *part,
                                           int nblocks, hipStream_t s);
/* the last y sweep of an adaptive fine branch fused with the error norms,
 * per-line powers and the Richardson combination (axis.hip); n = Ny = Nx */
template <class T> int launch_axis_richardson(const AArgs<T> &a, int n, const cx<T> *c1, double *err_part,
                                              double *lp_part, hipStream_t s);
int axis_richardson_tiles(int precision, int nt, int nx);
/* sums of the n (a, b) pairs part[2i], part[2i+1] in a fixed order into out[0..1] */
int launch_sum_pairs(const double *part, long n, double *out, hipStream_t s);
template <class T> int launch_energy_partials(const cx<T> *f, long n, double *part, int nblocks, hipStream_t s);
template <class T> int launch_ioniz(const cx<T> *f, int nt, long nlines, const PlasmaK &pk,
                                    cx<double> *tau_F, hipStream_t s);
int launch_adk(const double *E, long n, const double *kon, double *W1, double *W2, hipStream_t s);

/* setStep's step-size decision on a given error norm (propagation.c:633-677):
 * halve and reject above 2 tol, shrink by 2^(1/3) above tol, keep above tol/2,
 * else grow by 2^(1/3) up to 100x the mean beam radius y1[1], y2[1] */
void setstep_decide(double *dzz, double err, double def_err, double dzmin, const double *y1, const double *y2,
                    int *bad);

bool tline_supported(int nt);
bool axis_supported(int n);
int axis_tile(int precision);   /* G: t values per axis tile */
int axis_tile_g(int precision, int variant);   /* t values per tile of an axis variant */
int axis_tile_tg(int variant, int nchunks);    /* tile-order run length of an axis variant */
int tline_lines_per_block(int nt, bool plasma, int prec);

/* batched sweeps over several fields of one size class (one workgroup per
 * line group / tile of every field; descriptors and start offsets in device
 * memory) */
template <class T>
int launch_tline_batch(const TArgs<T> *d, const long *gstart, int n, long total, int nt, bool plasma, int variant,
                       hipStream_t s);
/* workgroups a per-field t-line launch over nlines lines runs (its plasma slots) */
template <class T> long tline_grid(long nlines, int nt, bool plasma, int variant);
/* workgroups a batched t-line launch over `total` line groups runs (variant 0: persistent) */
template <class T> long tline_batch_grid(long total, int nt, bool plasma, int variant);
template <class T>
int launch_axis_batch(const AArgs<T> *d, const long *tstart, int n, long nblocks, int len, int variant, hipStream_t s);

} // namespace mpc

/* the context behind mpc_ctx* */
struct mpc_ctx_s {
    mpc_params p;
    std::vector<double> betas, wr, w, r;
    int prec = MPC_F64;
    int nx = 0, ny = 0, nt = 0;
    long nxy = 0, N = 0;
    size_t esz = 16;                 /* bytes per complex element */
    hipStream_t stream = nullptr;
    int device = 0;
    void *field = nullptr;           /* resident field, internal layout */
    void *aux1 = nullptr, *aux2 = nullptr; /* adaptive C1/C2 (lazily) */
    double *staging = nullptr;       /* FP64 reference-layout staging (N complex) */
    /* checkpoint snapshot: staging -> pinned host copy on copy_stream;
     * anything that writes staging on `stream` first waits for snap_done */
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    cufftDoubleComplex *pinned = nullptr;
    int snap_pending = 0;
    double guard_tol = -1.0;         /* mpc_set_guard: energy-gain tolerance of mpc_run, < 0 off */
    void *tw_t = nullptr, *tw_x = nullptr;   /* twiddles */
    void *wfft = nullptr;            /* FFT-order Omega, precision T */
    void *mirror = nullptr;          /* mirror table [y][x] */
    double *rho_block = nullptr;     /* per t-line block */
    int rho_nvalid = 0;              /* entries written by the last plasma t-line launch */
    long rho_nblocks = 0;
    double *diag_part = nullptr;     /* diagnostics partials */
    double *diag_line = nullptr;     /* per-line power */
    double *diag_out = nullptr;      /* reduced temp/spec */
    double *err_part = nullptr;
    double *rich_err = nullptr;      /* adaptive: per-tile error norms of the fused sweep, then their sums */
    double *rich_lp = nullptr;       /* adaptive: per-chunk line powers of C2 */
    int diag_nblocks = 0;
    mpc::PlasmaK pk{};
    std::map<std::string, void *> tables;   /* table cache keyed by kind+coef */
    int domain_freq = 0;             /* resident field is F_t(A) (unnormalised) */
    int tvar = 0, avar = 0, avar_y = 0; /* kernel variants: t-line, x-line, y-line (autotuner / MPC_*_VARIANT) */
    int tuned = 0;
    /* the resident field holds A / ascale; ascale is a power of two chosen
     * at upload so |A / ascale| <= 1 (photon-flux amplitudes reach 1e18 and
     * their SST products would overflow FP32).  Nonlinear coefficients are
     * multiplied by ascale^2 instead. */
    double ascale = 1.0;
    double rho_peak = 0.0;
    /* per-launch timing: event pairs recorded on the library stream around
     * each sweep launch, resolved once the stream has drained */
    struct EvRec { hipEvent_t a = nullptr, b = nullptr; int which = 0; };
    int timing = 0;
    std::vector<EvRec> evs;
    size_t evn = 0;
    double kms[3] = {0, 0, 0};
    long kcount[3] = {0, 0, 0};
};

#endif
