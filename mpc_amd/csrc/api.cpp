/*
 * api.cpp -- host orchestration of the device-resident propagator and the
 * legacy entry points of propagation.h / plasma.h.
 *
 * Schedule of a fixed-step run (mpc_run).  One z-step is
 * L(dz/2) N(dz) L(dz/2) with L = Lt Lx Ly (separable, mutually commuting).
 * Consecutive half steps merge, so an interior step costs three sweeps:
 *   x-line sweep  (Fx, exp(-i c kx^2/2k0)/Nx, Fx^-1)
 *   y-line sweep  (Fy, exp(-i c ky^2/2k0)/Ny, Fy^-1) [x mirror at pass ends]
 *   t-line sweep  (x Dt(omega) -> N(dz) -> Ft), field left t-spectral
 * where c accumulates s*dz/2 of the half steps merged (s = pass sign).  The
 * propagator phases are formed in FP64 and reduced mod 2 pi before rounding
 * to the field precision.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/constants.h"
#include "../../include/mpc.h"
#include "../../include/plasma.h"
#include "../../include/propagation.h"
#include "internal.h"

using namespace mpc;

extern "C" int mpc_adk_constants(const char *gas, double *kon);

static int g_last_error = 0;

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            fprintf(stderr, "mpc: HIP error %s at %s:%d\n", hipGetErrorString(e_),  \
                    __FILE__, __LINE__);                                            \
            return MPC_EHIP;                                                        \
        }                                                                           \
    } while (0)
#define RCHK(x)                       \
    do {                              \
        int r_ = (x);                 \
        if (r_ != 0) return r_;       \
    } while (0)

extern "C" const char *mpc_strerror(int code) {
    switch (code) {
    case MPC_OK: return "ok";
    case MPC_EINVAL: return "invalid argument";
    case MPC_ENOMEM: return "out of device memory";
    case MPC_EHIP: return "HIP runtime error";
    case MPC_ENODEV: return "no GPU device";
    case MPC_EUNSUPP: return "unsupported size or option";
    case MPC_EIO: return "I/O error";
    case MPC_ENONFINITE: return "non-finite field (guard)";
    case MPC_EGUARD: return "energy or spectral-width guard failed";
    default: return "unknown error";
    }
}
extern "C" int mpc_last_error(void) { return g_last_error; }
extern "C" int mpc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
static double reduce_2pi(double ph) { return std::remainder(ph, 2.0 * M_PI); }

static int build_pk(mpc_ctx *c) {
    PlasmaK &k = c->pk;
    const mpc_params &p = c->p;
    double kon[6];
    if (mpc_adk_constants(p.gas, kon) != 0) return MPC_EINVAL;
    double I1 = 0, I2 = 0, Natm = 0, sc = 0;
    ionizpot(p.gas, &I1, &I2);
    Natoms(p.gas, &Natm, &sc);
    const double k0 = p.w0 / C0; /* plasma.c:142 recomputes k0 from w0 */
    k.hbar_w0 = HBAR * p.w0;
    k.inv_ef = 1.0 / EF;
    k.kon1 = kon[0]; k.kon2 = kon[1]; k.kon3 = kon[2]; k.kon4 = kon[3]; k.kon5 = kon[4]; k.kon6 = kon[5];
    k.K1 = 2.0 * EL0 * EL0;
    k.K2 = 4.0 * ME * ME;
    k.K3 = sc * p.rho_nt;
    k.K4 = k0 / (p.n0 * p.rho_c) * p.w0;
    k.K5 = p.w0 * p.w0;
    k.w0 = p.w0;
    /* sigma = K4 tau / (1 + K5 tau^2) <= K4 / (2 sqrt K5) and sigma tau <= K4 / K5 */
    k.sig_bound = k.K4 / (2.0 * std::sqrt(k.K5)) + p.w0 * k.K4 / k.K5;
    k.ve_fac = std::sqrt(k.K1 / k.K2) / p.w0;
    k.I1e = I1 * EL0;
    k.Kon = p.deltat / T0;
    k.deltat = p.deltat;
    k.dt_I1e = p.deltat / k.I1e;
    k.rho_nt = p.rho_nt;
    k.zf = 2.0 * ZW0 / NC;
    k.cs = k.zf * k.hbar_w0;
    k.kt = 1.0 / (k.K3 * FCTR * k.ve_fac);
    k.ke1 = k.kon3 / (3.0 * FCTR * k.inv_ef);
    k.ke2 = k.kon4 / (3.0 * FCTR * k.inv_ef);
    k.L1 = std::log(k.kon3 / (FCTR * k.inv_ef));
    k.L2 = std::log(k.kon4 / (FCTR * k.inv_ef));
    return 0;
}

template <class T> static std::vector<cx<T>> twiddles(int n) {
    std::vector<cx<T>> t(n);
    for (int k = 0; k < n; k++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * k / n;
        t[k] = mk((T)cosl(a), (T)sinl(a));
    }
    return t;
}

template <class T> static int upload_vec(void **dst, const std::vector<T> &v) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    HIPCHK(hipMalloc(dst, sizeof(T) * v.size()));
    HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
}

static void free_tables(mpc_ctx *c) {
    for (auto &kv : c->tables) (void)hipFree(kv.second);
    c->tables.clear();
}

/* precision-dependent setup of twiddles, the FFT-order Omega grid and the mirror */
template <class T> static int setup_tables(mpc_ctx *c) {
    RCHK(upload_vec(&c->tw_t, twiddles<T>(c->nt)));
    RCHK(upload_vec(&c->tw_x, twiddles<T>(c->nx)));
    std::vector<T> wf(c->nt);
    for (int k = 0; k < c->nt; k++) wf[k] = (T)c->w[(k + c->nt / 2) % c->nt]; /* propagation.c:170-177 */
    RCHK(upload_vec(&c->wfft, wf));
    /* mirror: curvature phase (FP64, reduced) x hard aperture x sqrt(Ab_M) */
    std::vector<cx<T>> M((size_t)c->nx * c->ny);
    const double amp = std::sqrt(c->p.Ab_M);
    for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++) {
            const double x = c->r[i], y = c->r[j], rr = x * x + y * y;
            double a = amp;
            if (c->p.aperture > 0.0 && rr > c->p.aperture * c->p.aperture) a = 0.0;
            const double ph = c->p.mirror_R > 0.0 ? reduce_2pi(-c->p.k0 * rr / c->p.mirror_R) : 0.0;
            M[(size_t)j * c->nx + i] = mk((T)(a * std::cos(ph)), (T)(a * std::sin(ph)));
        }
    RCHK(upload_vec(&c->mirror, M));
    return 0;
}

static int check_params(const mpc_params *p) {
    if (!p || !p->betas || !p->wr || !p->w || !p->r) return MPC_EINVAL;
    if (p->dims[0] != p->dims[1]) return MPC_EUNSUPP;   /* square grid (propagation.c:83,95,386) */
    if (!axis_supported(p->dims[0]) || !tline_supported(p->dims[2])) return MPC_EUNSUPP;
    if (p->quirks != 0) return MPC_EUNSUPP;              /* the device computes the corrected arithmetic */
    double kon[6];
    if (p->plasm == 1 && mpc_adk_constants(p->gas, kon) != 0) return MPC_EINVAL;
    return 0;
}

static int ctx_set_params(mpc_ctx *c, const mpc_params *p) {
    c->p = *p;
    c->betas.assign(p->betas, p->betas + p->dims[2]);
    c->w.assign(p->w, p->w + p->dims[2]);
    c->wr.assign(p->wr, p->wr + p->dims[1]);
    c->r.assign(p->r, p->r + p->dims[1]);
    c->p.betas = c->betas.data(); c->p.w = c->w.data(); c->p.wr = c->wr.data(); c->p.r = c->r.data();
    free_tables(c);
    if (p->plasm == 1) RCHK(build_pk(c));
    return c->prec == MPC_F32 ? setup_tables<float>(c) : setup_tables<double>(c);
}

extern "C" int mpc_create(const mpc_params *p, int precision, mpc_ctx **out) {
    if (!out) return MPC_EINVAL;
    *out = nullptr;
    RCHK(check_params(p));
    if (precision != MPC_F32 && precision != MPC_F64) return MPC_EINVAL;
    if (mpc_device_count() < 1) return MPC_ENODEV;
    mpc_ctx *c = new mpc_ctx();
    c->prec = precision;
    c->nx = p->dims[0]; c->ny = p->dims[1]; c->nt = p->dims[2];
    c->nxy = (long)c->nx * c->ny;
    c->N = c->nxy * c->nt;
    c->esz = precision == MPC_F32 ? 8 : 16;
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int rc = 0;
    if (hipMalloc(&c->field, c->esz * (size_t)c->N) != hipSuccess) rc = MPC_ENOMEM;
    if (!rc && hipMalloc((void **)&c->staging, 16 * (size_t)c->N) != hipSuccess) rc = MPC_ENOMEM;
    if (!rc) {
        const int lpb = tline_lines_per_block(c->nt, true, precision);
        c->rho_nblocks = c->nxy / lpb;
        if (hipMalloc((void **)&c->rho_block, sizeof(double) * (size_t)c->rho_nblocks) != hipSuccess) rc = MPC_ENOMEM;
    }
    if (!rc) {
        /* blocks of the diagnostics pass: one per line group up to 1024 (a
         * group is max(1, 2048 / nt) lines) */
        const long dgroups = c->nxy / std::max(1, 2048 / c->nt);
        c->diag_nblocks = (int)std::max(1L, std::min(1024L, dgroups));
        if (hipMalloc((void **)&c->diag_part, sizeof(double) * 2 * (size_t)c->diag_nblocks * c->nt) != hipSuccess ||
            hipMalloc((void **)&c->diag_line, sizeof(double) * (size_t)c->nxy) != hipSuccess ||
            hipMalloc((void **)&c->diag_out, sizeof(double) * 2 * (size_t)c->nt) != hipSuccess ||
            hipMalloc((void **)&c->err_part, sizeof(double) * 2 * 1024) != hipSuccess)
            rc = MPC_ENOMEM;
    }
    if (!rc) rc = ctx_set_params(c, p);
    if (rc) { mpc_destroy(c); return rc; }
    if (const char *e = getenv("MPC_TLINE_VARIANT")) c->tvar = std::max(0, std::min(tline_variants() - 1, atoi(e)));
    if (const char *e = getenv("MPC_AXIS_VARIANT")) c->avar = c->avar_y = std::max(0, std::min(axis_variants() - 1, atoi(e)));
    *out = c;
    return 0;
}

extern "C" void mpc_destroy(mpc_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->pinned) (void)hipHostFree(c->pinned);
    free_tables(c);
    void *bufs[] = {c->field, c->aux1, c->aux2, c->staging, c->tw_t, c->tw_x, c->wfft, c->mirror,
                    c->rho_block, c->diag_part, c->diag_line, c->diag_out, c->err_part, c->rich_err, c->rich_lp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &e : c->evs) {
        if (e.a) (void)hipEventDestroy(e.a);
        if (e.b) (void)hipEventDestroy(e.b);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" long long mpc_field_bytes(mpc_ctx *c) { return c ? (long long)c->esz * c->N : 0; }

/* ------------------------------------------------------------------------ */
/* tables                                                                    */
/* ------------------------------------------------------------------------ */
/* kind 't': exp(ct (i beta_k - alpha/2)) / Nt, beta in FFT order
 *           (propagation.c:150-156 for one half step, ct = dz/2);
 * kind 'x'/'y': exp(-i c k^2 / (2 k0)) / n, k in FFT order, c = sum of
 *           s * dz/2 over the merged half steps (conj of calcexpR). */
template <class T> static const cx<T> *get_table(mpc_ctx *c, char kind, double coef) {
    char key[64];
    snprintf(key, sizeof key, "%c:%.17g", kind, coef);
    auto it = c->tables.find(key);
    if (it != c->tables.end()) return (const cx<T> *)it->second;
    std::vector<cx<T>> tab;
    if (kind == 't' || kind == 'n') {
        /* 't': carries the 1/nt of one forward/inverse t-FFT pair; 'n': the
         * same factor without it */
        const int nt = c->nt;
        tab.resize(nt);
        for (int k = 0; k < nt; k++) {
            const double beta = c->betas[(k + nt / 2) % nt];
            const double mag = std::exp(-c->p.alpha / 2.0 * coef) / (kind == 't' ? nt : 1);
            const double ph = reduce_2pi(beta * coef);
            tab[k] = mk((T)(mag * std::cos(ph)), (T)(mag * std::sin(ph)));
        }
    } else {
        const int n = c->nx;
        tab.resize(n);
        for (int i = 0; i < n; i++) {
            const double kx = c->wr[(i + c->ny / 2) % c->ny];
            const double ph = reduce_2pi(-coef / (2.0 * c->p.k0) * kx * kx);
            tab[i] = mk((T)(std::cos(ph) / n), (T)(std::sin(ph) / n));
        }
    }
    void *d = nullptr;
    if (upload_vec(&d, tab) != 0) return nullptr;
    c->tables[key] = d;
    return (const cx<T> *)d;
}

/* a caller-given table (legacy linearstep) uploaded with a scale */
template <class T> static const cx<T> *user_table(mpc_ctx *c, const char *tag, const cufftDoubleComplex *src,
                                                  int n, double scale, bool conj) {
    std::vector<cx<T>> tab(n);
    for (int i = 0; i < n; i++) tab[i] = mk((T)(src[i].x * scale), (T)((conj ? -src[i].y : src[i].y) * scale));
    std::string key = std::string("u:") + tag;
    auto it = c->tables.find(key);
    if (it != c->tables.end()) { (void)hipFree(it->second); c->tables.erase(it); }
    void *d = nullptr;
    if (upload_vec(&d, tab) != 0) return nullptr;
    c->tables[key] = d;
    return (const cx<T> *)d;
}

/* ------------------------------------------------------------------------ */
/* sweeps                                                                    */
/* ------------------------------------------------------------------------ */
static void tic(mpc_ctx *c, int which) {
    if (!c->timing) return;
    if (c->evn == c->evs.size()) {
        mpc_ctx::EvRec r;
        if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
        c->evs.push_back(r);
    }
    c->evs[c->evn].which = which;
    (void)hipEventRecord(c->evs[c->evn].a, c->stream);
}
static void toc(mpc_ctx *c, int which) {
    if (!c->timing || c->evn >= c->evs.size() || c->evs[c->evn].which != which) return;
    (void)hipEventRecord(c->evs[c->evn].b, c->stream);
    c->evn++;
}
/* after the stream has drained: fold recorded launch times into the averages */
static void resolve_timing(mpc_ctx *c) {
    for (size_t i = 0; i < c->evn; i++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->evs[i].a, c->evs[i].b) == hipSuccess) {
            c->kms[c->evs[i].which] += ms;
            c->kcount[c->evs[i].which] += 1;
        }
    }
    c->evn = 0;
}

/* kernel arguments of one x (axis 0) or y (axis 1) sweep over c's field */
template <class T> static AArgs<T> axis_args(const mpc_ctx *c, int axis, const cx<T> *D, bool mirror,
                                             const cx<T> *Pt) {
    AArgs<T> a{};
    a.field = (cx<T> *)c->field;
    a.tw = (const cx<T> *)c->tw_x;
    a.D = D;
    a.nt = c->nt;
    a.M = nullptr;
    a.Pt = Pt;
    if (axis == 0) {        /* x-lines: tile at fixed y */
        a.s_line = c->nt;
        a.s_outer = (long)c->nx * c->nt;
        if (mirror) { a.M = (const cx<T> *)c->mirror; a.m_line = 1; a.m_outer = c->nx; }
    } else {                /* y-lines: tile at fixed x */
        a.s_line = (long)c->nx * c->nt;
        a.s_outer = c->nt;
        if (mirror) { a.M = (const cx<T> *)c->mirror; a.m_line = c->nx; a.m_outer = 1; }
    }
    return a;
}

template <class T> static int sweep_axis(mpc_ctx *c, int axis, const cx<T> *D, bool mirror,
                                         const cx<T> *Pt = nullptr, const void *src = nullptr) {
    AArgs<T> a = axis_args<T>(c, axis, D, mirror, Pt);
    a.src = (const cx<T> *)src;
    tic(c, 1 + axis);
    const int rc = launch_axis<T>(a, axis == 0 ? c->nx : c->ny, axis == 0 ? c->ny : c->nx,
                                  axis == 0 ? c->avar : c->avar_y, c->stream);
    toc(c, 1 + axis);
    return rc;
}

/* x then y sweep with the merged diffraction factor; Pt (t-spectral field
 * only) is a per-t factor the y sweep applies before its store */
template <class T> static int run_xy(mpc_ctx *c, double coef, bool mirror, const cx<T> *Pt = nullptr,
                                     const void *src = nullptr) {
    if (coef == 0.0 && !mirror && !Pt && !src) return 0;
    const cx<T> *D = get_table<T>(c, 'x', coef);
    if (!D) return MPC_ENOMEM;
    RCHK(sweep_axis<T>(c, 0, D, false, nullptr, src));      /* src: the x sweep reads it, writes c's field */
    return sweep_axis<T>(c, 1, D, mirror, Pt);
}

/* kernel arguments of one t-line sweep over c's field */
template <class T> static TArgs<T> tline_args(const mpc_ctx *c, const cx<T> *P, const cx<T> *Q, bool in_freq,
                                              bool do_nl, bool out_time, double dz) {
    TArgs<T> a{};
    a.field = (cx<T> *)c->field;
    a.tw = (const cx<T> *)c->tw_t;
    a.P = P;
    a.Q = Q;
    a.wfft = (const T *)c->wfft;
    const double s2 = c->ascale * c->ascale;   /* the field holds A / ascale */
    a.pk = c->pk;
    a.pk.hbar_w0 = c->pk.hbar_w0 * s2;
    a.pk.cs = c->pk.cs * s2;
    a.dz = (T)dz;
    a.gamma = (T)(c->p.gamma2 * s2);
    a.sst_coef = (T)(c->p.gamma2 / c->p.w0 * s2);
    a.in_freq = in_freq;
    a.do_nl = do_nl;
    a.out_time = out_time;
    a.sst_on = c->p.sst_on == 1;
    a.nlines = c->nxy;
    const bool plasma = do_nl && c->p.plasm == 1;
    a.rho_block = plasma ? c->rho_block : nullptr;
    return a;
}

template <class T> static int sweep_t(mpc_ctx *c, const cx<T> *P, const cx<T> *Q, bool in_freq, bool do_nl,
                                      bool out_time, double dz) {
    const TArgs<T> a = tline_args<T>(c, P, Q, in_freq, do_nl, out_time, dz);
    const bool plasma = do_nl && c->p.plasm == 1;
    tic(c, 0);
    int nb = 0;
    const int rc = launch_tline<T>(a, c->nt, plasma, c->tvar, c->stream, &nb);
    if (!rc && plasma) c->rho_nvalid = nb;
    toc(c, 0);
    return rc;
}

static int pass_sign(const mpc_ctx *c, double z) {
    if (c->p.mirror_R > 0.0) return 1;
    return ((int)(std::floor(z / c->p.Cav) + 1) % 2) * 2 - 1;
}

/*
 * One sweep of the fixed-step schedule.  kind 'x' / 'y': diffraction with the
 * merged coefficient coef ('y' may carry the mirror and, on a t-spectral
 * field, the t table of coefficient pcoef); 't': the fused t-line sweep with
 * optional P (pcoef) and Q (dz/2) tables.
 */
struct SweepOp {
    char kind = 'x';
    double coef = 0.0;
    bool mirror = false;
    bool has_p = false, has_q = false;
    bool q_unnorm = false;        /* Q without the 1/nt: the field stays t-spectral after the t-line */
    double pcoef = 0.0;
    bool in_freq = false, out_time = false;
    double dz = 0.0;
    int step = 0;
};

/*
 * The schedule of mpc_run(step0, nsteps) on c as a list of sweeps: step n is
 * L(dz/2) N(dz) L(dz/2) at z = (n + 1/2) dz; consecutive half steps merge,
 * the field stays t-spectral between steps (the dispersion factor then rides
 * on the y sweep), and the mirror is applied after the last step of each
 * pass but the last.  The run also ends t-spectral (the last t-line applies
 * its closing half step with an 'n' table and no inverse transform), so the
 * next run, or the next diagnostics, starts without a transform; the calls
 * that need the time domain convert on their own.  *freq_end receives the
 * domain the run leaves.
 */
static void plan_fixed(const mpc_ctx *c, int step0, int nsteps, std::vector<SweepOp> &ops, bool *freq_end) {
    const double dz = c->p.dzz;
    const int S = c->p.steps_per_pass > 0 ? c->p.steps_per_pass : 1;
    const int npass = c->p.npasses > 0 ? c->p.npasses : 1;
    double ct = 0.0, cxy = 0.0;
    bool freq = c->domain_freq != 0;
    auto xy = [&](double coef, bool mirror, bool pfused, double pcoef, int n) {
        if (coef == 0.0 && !mirror && !pfused) return;
        SweepOp x;
        x.kind = 'x'; x.coef = coef; x.step = n;
        ops.push_back(x);
        SweepOp y = x;
        y.kind = 'y'; y.mirror = mirror; y.has_p = pfused; y.pcoef = pcoef;
        ops.push_back(y);
    };
    for (int n = step0; n < step0 + nsteps; n++) {
        const double z = ((double)n + 0.5) * dz;
        const int s = pass_sign(c, z);
        ct += dz / 2;
        cxy += s * dz / 2;
        /* on a t-spectral field the dispersion factor rides on the y sweep's
         * store (the linear factors commute); the t-line kernel then skips it */
        const bool p_in_y = freq && cxy != 0.0;
        xy(cxy, false, p_in_y, ct, n);
        cxy = 0.0;
        const bool last = (n == step0 + nsteps - 1);
        const bool mirror_after = ((n + 1) % S == 0) && ((n + 1) / S < npass);
        SweepOp t;
        t.kind = 't'; t.step = n; t.dz = dz; t.in_freq = freq;
        t.has_p = !p_in_y; t.pcoef = ct;
        if (last) {
            t.has_q = true; t.q_unnorm = true;
            ct = 0.0;
        } else {
            ct = dz / 2;
        }
        freq = true;
        ops.push_back(t);
        cxy = s * dz / 2;
        if (mirror_after) {
            xy(cxy, true, false, 0.0, n);
            cxy = 0.0;
        }
    }
    if (cxy != 0.0) xy(cxy, false, false, 0.0, step0 + nsteps - 1);
    *freq_end = freq;
}

/* one planned sweep on c's own stream */
template <class T> static int exec_op(mpc_ctx *c, const SweepOp &o) {
    if (o.kind == 't') {
        const cx<T> *P = o.has_p ? get_table<T>(c, 't', o.pcoef) : nullptr;
        const cx<T> *Q = o.has_q ? get_table<T>(c, o.q_unnorm ? 'n' : 't', o.dz / 2) : nullptr;
        if ((o.has_p && !P) || (o.has_q && !Q)) return MPC_ENOMEM;
        return sweep_t<T>(c, P, Q, o.in_freq, true, o.out_time, o.dz);
    }
    const cx<T> *D = get_table<T>(c, 'x', o.coef);
    const cx<T> *Pt = o.has_p ? get_table<T>(c, 't', o.pcoef) : nullptr;
    if (!D || (o.has_p && !Pt)) return MPC_ENOMEM;
    return sweep_axis<T>(c, o.kind == 'x' ? 0 : 1, D, o.mirror, Pt);
}

template <class T> static int run_fixed(mpc_ctx *c, int step0, int nsteps) {
    std::vector<SweepOp> ops;
    bool freq = false;
    plan_fixed(c, step0, nsteps, ops, &freq);
    for (const SweepOp &o : ops) RCHK(exec_op<T>(c, o));
    c->domain_freq = freq ? 1 : 0;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* batched runs: many pulses, one launch per sweep and size class, a HIP     */
/* graph per pass                                                            */
/* ------------------------------------------------------------------------ */
struct mpc_batch_s {
    std::vector<mpc_ctx *> ctx;
    int prec = MPC_F64;
    int S = 1;                                   /* steps per pass, common to all pulses */
    hipStream_t stream = nullptr;
    std::map<std::string, void *> pool;          /* device copies of descriptor arrays, by content */
    struct Graph { hipGraph_t g = nullptr; hipGraphExec_t e = nullptr; };
    std::map<std::string, Graph> graphs;         /* per-pass graphs, by launch list */
};

namespace {
/* one batched launch: sweep kind, size class, kernel variant and the device
 * descriptor / start arrays it reads */
struct BLaunch {
    char kind;
    int len;            /* t length ('t') or line length ('x', 'y') */
    int plasma;
    int variant;
    int n;              /* fields in the launch */
    long nblocks;       /* line groups ('t') or tiles ('x', 'y') over all fields */
    long grid;          /* workgroups launched */
    const void *desc;
    const long *starts;
    int step;
    std::vector<int> members;   /* batch indices, for the per-field bookkeeping */
    /* a launch group of one field runs the per-field kernel with this
     * descriptor (the same arithmetic per line and tile, without the batched
     * kernels' per-field lookups) */
    std::vector<unsigned char> single;
    int outer = 0;              /* 'x' / 'y': lines per t value of that field */
};
}

/* pulses of one size class share a launch up to this many field bytes (the
 * last-level cache, 256 MB on MI355X): a group's sweeps can then run back to
 * back and each finds part of what the previous one wrote still cached */
constexpr size_t BATCH_GROUP_BYTES = (size_t)256 << 20;

/* a device copy of bytes, shared by identical contents */
static const void *pool_put(mpc_batch *b, const void *src, size_t bytes) {
    std::string key((const char *)src, bytes);
    auto it = b->pool.find(key);
    if (it != b->pool.end()) return it->second;
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(d); return nullptr; }
    b->pool[key] = d;
    return d;
}

/* the launches of op index i of every pulse's plan, grouped by size class */
template <class T>
static int batch_group(mpc_batch *b, const std::vector<std::vector<SweepOp>> &plans, size_t i,
                       std::vector<BLaunch> &out) {
    const char kind = plans[0][i].kind;
    struct Cls { int len, plasma; std::vector<int> m; };
    std::vector<Cls> cls;
    for (int p = 0; p < (int)b->ctx.size(); p++) {
        const mpc_ctx *c = b->ctx[p];
        const int len = kind == 't' ? c->nt : (kind == 'x' ? c->nx : c->ny);
        const int pl = kind == 't' ? (c->p.plasm == 1) : 0;
        auto f = std::find_if(cls.begin(), cls.end(), [&](const Cls &k) { return k.len == len && k.plasma == pl; });
        if (f == cls.end()) { cls.push_back({len, pl, {p}}); } else f->m.push_back(p);
    }
    /* each class in groups of at most BATCH_GROUP_BYTES of fields (at least
     * one field per group) */
    std::vector<Cls> grp;
    for (const Cls &k : cls) {
        size_t bytes = 0;
        for (size_t j = 0; j < k.m.size(); j++) {
            const mpc_ctx *c = b->ctx[k.m[j]];
            const size_t fb = c->esz * (size_t)c->N;
            if (j == 0 || bytes + fb > BATCH_GROUP_BYTES) {
                grp.push_back({k.len, k.plasma, {}});
                bytes = 0;
            }
            grp.back().m.push_back(k.m[j]);
            bytes += fb;
        }
    }
    for (const Cls &k : grp) {
        BLaunch L{};
        L.kind = kind; L.len = k.len; L.plasma = k.plasma; L.n = (int)k.m.size(); L.step = plans[0][i].step;
        L.members = k.m;
        std::vector<long> starts;
        long tot = 0;
        if (kind == 't') {
            std::vector<TArgs<T>> d;
            for (int p : k.m) {
                mpc_ctx *c = b->ctx[p];
                const SweepOp &o = plans[p][i];
                const cx<T> *P = o.has_p ? get_table<T>(c, 't', o.pcoef) : nullptr;
                const cx<T> *Q = o.has_q ? get_table<T>(c, o.q_unnorm ? 'n' : 't', o.dz / 2) : nullptr;
                if ((o.has_p && !P) || (o.has_q && !Q)) return MPC_ENOMEM;
                d.push_back(tline_args<T>(c, P, Q, o.in_freq, true, o.out_time, o.dz));
                starts.push_back(tot);
                tot += c->nxy / tline_lines_per_block(c->nt, true, b->prec);
            }
            starts.push_back(tot);                           /* starts[n] = total */
            L.variant = b->ctx[k.m[0]]->tvar;                 /* persistent (0) or per group (1) */
            if (d.size() == 1) {
                L.single.resize(sizeof(TArgs<T>));
                std::memcpy(L.single.data(), &d[0], sizeof(TArgs<T>));
            }
            L.desc = pool_put(b, d.data(), d.size() * sizeof(TArgs<T>));
        } else {
            const int axis = kind == 'x' ? 0 : 1;
            const mpc_ctx *c0 = b->ctx[k.m[0]];
            L.variant = axis == 0 ? c0->avar : c0->avar_y;            /* plan, tile order and persistence */
            const int G = axis_tile_g(b->prec, L.variant);
            std::vector<AArgs<T>> d;
            for (int p : k.m) {
                mpc_ctx *c = b->ctx[p];
                const SweepOp &o = plans[p][i];
                const cx<T> *D = get_table<T>(c, 'x', o.coef);
                const cx<T> *Pt = o.has_p ? get_table<T>(c, 't', o.pcoef) : nullptr;
                if (!D || (o.has_p && !Pt)) return MPC_ENOMEM;
                if (c->nt % G) return MPC_EINVAL;
                AArgs<T> a = axis_args<T>(c, axis, D, o.mirror, Pt);
                const int nch = c->nt / G;
                a.n_outer = axis == 0 ? c->ny : c->nx;
                a.tg = axis_tile_tg(L.variant, nch);
                d.push_back(a);
                starts.push_back(tot);
                tot += (long)nch * a.n_outer;
            }
            starts.push_back(tot);                           /* starts[n] = total */
            if (d.size() == 1) {
                L.single.resize(sizeof(AArgs<T>));
                std::memcpy(L.single.data(), &d[0], sizeof(AArgs<T>));
                L.outer = d[0].n_outer;
            }
            L.desc = pool_put(b, d.data(), d.size() * sizeof(AArgs<T>));
        }
        L.starts = (const long *)pool_put(b, starts.data(), starts.size() * sizeof(long));
        L.nblocks = tot;
        if (!L.single.empty())
            L.grid = kind == 't' ? tline_grid<T>(b->ctx[k.m[0]]->nxy, k.len, k.plasma != 0, L.variant) : tot;
        else
            L.grid = kind == 't' ? tline_batch_grid<T>(tot, k.len, k.plasma != 0, L.variant) : tot;
        if (!L.desc || !L.starts) return MPC_ENOMEM;
        out.push_back(std::move(L));
    }
    return 0;
}

template <class T> static int batch_launch(mpc_batch *b, const BLaunch &L) {
    if (!L.single.empty()) {
        if (L.kind == 't') {
            TArgs<T> a;
            std::memcpy(&a, L.single.data(), sizeof a);
            return launch_tline<T>(a, L.len, L.plasma != 0, L.variant, b->stream, nullptr);
        }
        AArgs<T> a;
        std::memcpy(&a, L.single.data(), sizeof a);
        return launch_axis<T>(a, L.len, L.outer, L.variant, b->stream);
    }
    if (L.kind == 't')
        return launch_tline_batch<T>((const TArgs<T> *)L.desc, L.starts, L.n, L.nblocks, L.len, L.plasma != 0,
                                     L.variant, b->stream);
    return launch_axis_batch<T>((const AArgs<T> *)L.desc, L.starts, L.n, L.nblocks, L.len, L.variant, b->stream);
}

/* launches [i0, i1) of one pass in group-major order: each group's launches
 * in their own order, one group after another.  Applied only when every sweep
 * kind splits the pulses into the same groups (each pulse's launches then
 * stay in order); otherwise the sweep-major order is kept. */
static void group_major(std::vector<BLaunch> &ls, size_t i0, size_t i1, size_t np) {
    std::vector<const std::vector<int> *> owner(np, nullptr);
    for (size_t i = i0; i < i1; i++)
        for (int p : ls[i].members) {
            if (!owner[p]) owner[p] = &ls[i].members;
            else if (*owner[p] != ls[i].members) return;
        }
    std::stable_sort(ls.begin() + (long)i0, ls.begin() + (long)i1,
                     [](const BLaunch &a, const BLaunch &b) { return a.members[0] < b.members[0]; });
}

/* the launch list of a pass, as a graph key */
static std::string launch_key(const std::vector<BLaunch> &ls, size_t i0, size_t i1) {
    std::string k;
    for (size_t i = i0; i < i1; i++) {
        const BLaunch &L = ls[i];
        char buf[128];
        snprintf(buf, sizeof buf, "%c%d.%d.%d.%d.%ld.%ld.%p.%p;", L.kind, L.len, L.plasma, L.variant, L.n, L.nblocks,
                 L.grid, L.desc, (const void *)L.starts);
        k += buf;
    }
    return k;
}

template <class T> static int batch_run(mpc_batch *b, int step0, int nsteps, bool use_graph) {
    const size_t np = b->ctx.size();
    std::vector<std::vector<SweepOp>> plans(np);
    std::vector<char> fend(np);
    for (size_t p = 0; p < np; p++) {
        bool fe = false;
        plan_fixed(b->ctx[p], step0, nsteps, plans[p], &fe);
        fend[p] = fe;
        if (plans[p].size() != plans[0].size()) return MPC_EINVAL;
        for (size_t i = 0; i < plans[p].size(); i++)
            if (plans[p][i].kind != plans[0][i].kind || plans[p][i].step != plans[0][i].step) return MPC_EINVAL;
    }
    /* every table and descriptor exists before anything is launched or captured */
    std::vector<BLaunch> ls;
    for (size_t i = 0; i < plans[0].size(); i++) RCHK(batch_group<T>(b, plans, i, ls));
    for (size_t j0 = 0; j0 < ls.size();) {
        size_t j1 = j0;
        while (j1 < ls.size() && ls[j1].step / b->S == ls[j0].step / b->S) j1++;
        group_major(ls, j0, j1, np);
        j0 = j1;
    }
    for (mpc_ctx *c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    size_t i0 = 0;
    while (i0 < ls.size()) {
        /* launches of one pass */
        const int pass = ls[i0].step / b->S;
        size_t i1 = i0;
        while (i1 < ls.size() && ls[i1].step / b->S == pass) i1++;
        if (!use_graph) {
            for (size_t i = i0; i < i1; i++) RCHK(batch_launch<T>(b, ls[i]));
        } else {
            const std::string key = launch_key(ls, i0, i1);
            auto it = b->graphs.find(key);
            if (it == b->graphs.end()) {
                mpc_batch_s::Graph g;
                HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = 0;
                for (size_t i = i0; i < i1 && !rc; i++) rc = batch_launch<T>(b, ls[i]);
                const hipError_t ce = hipStreamEndCapture(b->stream, &g.g);
                if (rc) { if (g.g) (void)hipGraphDestroy(g.g); return rc; }
                HIPCHK(ce);
                HIPCHK(hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0));
                it = b->graphs.emplace(key, g).first;
            }
            HIPCHK(hipGraphLaunch(it->second.e, b->stream));
        }
        i0 = i1;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t p = 0; p < np; p++) b->ctx[p]->domain_freq = fend[p] ? 1 : 0;
    /* plasma maxima: a per-group launch wrote one slot per line group, a
     * persistent one the first min(grid, groups) slots of each field */
    for (const BLaunch &L : ls) {
        if (L.kind != 't' || !L.plasma) continue;
        for (int p : L.members) {
            mpc_ctx *c = b->ctx[p];
            const long groups = c->nxy / tline_lines_per_block(c->nt, true, b->prec);
            c->rho_nvalid = (int)(L.variant == 0 ? std::min(L.grid, groups) : groups);
        }
    }
    return 0;
}

/*
 * Re-tunes the kernel variants on the batch's own launches: the t-line form
 * (persistent or one workgroup per line group) of each launch class (t
 * length, plasma switch), and the x/y variants of each class whose launches
 * cover several pulses (the batched kernels; a launch over one pulse runs the
 * per-field kernel its context's own tuning chose).  A variant that is
 * faster on one pulse alone can be slower over a group of many.  Runs on
 * scratch copies of the fields (aux1), with the plasma slots detached, so the
 * pulses' fields and diagnostics are left as they were.
 */
template <class T> static int batch_autotune(mpc_batch *b) {
    const size_t np = b->ctx.size();
    std::vector<void *> keep_field(np);
    std::vector<double *> keep_rho(np);
    std::vector<int> keep_t(np), keep_x(np), keep_y(np);
    for (size_t p = 0; p < np; p++) {
        mpc_ctx *c = b->ctx[p];
        keep_x[p] = c->avar;
        keep_y[p] = c->avar_y;
        const size_t bytes = c->esz * (size_t)c->N;
        if (!c->aux1 && hipMalloc(&c->aux1, bytes) != hipSuccess) return MPC_ENOMEM;
        HIPCHK(hipMemcpyAsync(c->aux1, c->field, bytes, hipMemcpyDeviceToDevice, b->stream));
        keep_field[p] = c->field;
        keep_rho[p] = c->rho_block;
        keep_t[p] = c->tvar;
    }
    for (size_t p = 0; p < np; p++) {
        b->ctx[p]->field = b->ctx[p]->aux1;
        b->ctx[p]->rho_block = nullptr;
    }
    /* the scratch copies in the state a run keeps between steps (t-spectral,
     * unnormalised), so the timed t-lines see the run's intensities */
    (void)hipStreamSynchronize(b->stream);
    int rc = 0;
    for (size_t p = 0; p < np && !rc; p++) {
        mpc_ctx *c = b->ctx[p];
        if (!c->domain_freq) rc = sweep_t<T>(c, nullptr, nullptr, false, false, false, 0.0);
    }
    for (mpc_ctx *c : b->ctx) (void)hipStreamSynchronize(c->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = MPC_EHIP;
    /* the plan of three steps: its second t, x and y ops (t-spectral in and
     * out, as every interior step's) are the launch lists timed */
    std::vector<std::vector<SweepOp>> plans(np);
    for (size_t p = 0; p < np && !rc; p++) {
        bool fe = false;
        plan_fixed(b->ctx[p], 0, 3, plans[p], &fe);
        if (plans[p].size() != plans[0].size()) rc = MPC_EINVAL;
    }
    auto second_op = [&](char kind) -> size_t {
        int seen = 0;
        for (size_t i = 0; i < plans[0].size(); i++)
            if (plans[0][i].kind == kind && ++seen == 2) return i;
        return plans[0].size();
    };
    /* The timed t op applies the dispersion table itself (which carries the
     * 1/Nt of the transform pair), as a run's first step does: repeated on
     * the scratch copy it then keeps the state a run keeps between steps.
     * (In a run the y sweep before it applies the table instead.) */
    {
        const size_t oi = second_op('t');
        if (oi < plans[0].size())
            for (size_t p = 0; p < np; p++) {
                plans[p][oi].has_p = true;
                plans[p][oi].pcoef = b->ctx[p]->p.dzz;
            }
    }
    /* time of a launch: median of 5 samples of k back-to-back launches
     * (k: enough for about 0.2 ms) after one warm-up launch */
    auto time_launch = [&](const BLaunch &L) -> float {
        auto timed = [&](int k) -> float {
            (void)hipEventRecord(e0, b->stream);
            for (int i = 0; i < k; i++)
                if (batch_launch<T>(b, L) != 0) return 1e30f;
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 1e30f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return ms / (float)k;
        };
        const float t0 = timed(1);
        if (t0 >= 1e30f) return 1e30f;
        const int k = t0 > 0.0f && t0 < 0.2f ? std::min(64, (int)std::ceil(0.2f / t0)) : 1;
        float t[5];
        for (int r = 0; r < 5; r++) t[r] = timed(k);
        std::sort(t, t + 5);
        return t[2];
    };
    /* Tunes one sweep kind: every variant is set on every pulse (set_v), the
     * kind's launches of the op are built and timed, and each launch class
     * (line length, plasma switch) keeps its fastest variant in best_v.  With
     * multi_only, only launches over several pulses count (a launch over one
     * pulse runs the per-field kernel its context's own tuning chose), and a
     * class with none of them is left out. */
    using Key = std::pair<int, int>;
    auto tune_kind = [&](char kind, int nvar, bool multi_only, auto &&set_v, std::map<Key, int> &best_v) {
        const size_t oi = second_op(kind);
        if (rc || oi == plans[0].size()) return;
        (void)hipGetLastError();
        std::map<Key, float> best_ms;
        for (int v = 0; v < nvar && !rc; v++) {
            for (mpc_ctx *c : b->ctx) set_v(c, v);
            std::vector<BLaunch> ls;
            if (batch_group<T>(b, plans, oi, ls) != 0) continue;     /* a variant that does not fit: skipped */
            std::map<Key, float> sum;
            bool ok = true;
            for (const BLaunch &L : ls) {
                if (multi_only && !L.single.empty()) continue;
                const float ms = time_launch(L);
                if (ms >= 1e30f) { ok = false; break; }
                sum[{L.len, L.plasma}] += ms;
            }
            if (!ok) continue;
            for (const auto &kv : sum) {
                auto it = best_ms.find(kv.first);
                if (it == best_ms.end() || kv.second < it->second) {
                    best_ms[kv.first] = kv.second;
                    best_v[kv.first] = v;
                }
            }
        }
    };
    std::map<Key, int> bt, bx, by;
    tune_kind('t', tline_variants(), false, [](mpc_ctx *c, int v) { c->tvar = v; }, bt);
    for (size_t p = 0; p < np; p++) b->ctx[p]->tvar = keep_t[p];
    tune_kind('x', axis_variants(), true, [](mpc_ctx *c, int v) { c->avar = v; }, bx);
    for (size_t p = 0; p < np; p++) b->ctx[p]->avar = keep_x[p];
    tune_kind('y', axis_variants(), true, [](mpc_ctx *c, int v) { c->avar_y = v; }, by);
    for (size_t p = 0; p < np; p++) b->ctx[p]->avar_y = keep_y[p];
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(b->stream);
    for (size_t p = 0; p < np; p++) {
        mpc_ctx *c = b->ctx[p];
        c->field = keep_field[p];
        c->rho_block = keep_rho[p];
        if (rc) continue;
        auto it = bt.find({c->nt, c->p.plasm == 1 ? 1 : 0});
        if (it != bt.end()) c->tvar = it->second;
        auto ix = bx.find({c->nx, 0});
        if (ix != bx.end()) c->avar = ix->second;
        auto iy = by.find({c->ny, 0});
        if (iy != by.end()) c->avar_y = iy->second;
    }
    return rc;
}

extern "C" int mpc_batch_autotune(mpc_batch *b) {
    if (!b) return MPC_EINVAL;
    return b->prec == MPC_F32 ? batch_autotune<float>(b) : batch_autotune<double>(b);
}

extern "C" int mpc_batch_create(mpc_ctx *const *ctxs, int n, mpc_batch **out) {
    if (!ctxs || n <= 0 || !out) return MPC_EINVAL;
    *out = nullptr;
    auto *b = new mpc_batch_s();
    for (int i = 0; i < n; i++) {
        mpc_ctx *c = ctxs[i];
        if (!c) { delete b; return MPC_EINVAL; }
        const int S = c->p.steps_per_pass > 0 ? c->p.steps_per_pass : 1;
        if (i == 0) { b->prec = c->prec; b->S = S; }
        if (c->prec != b->prec || S != b->S || c->p.npasses != ctxs[0]->p.npasses || c->device != ctxs[0]->device) {
            delete b;
            return MPC_EINVAL;
        }
        b->ctx.push_back(c);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return MPC_EHIP; }
    *out = b;
    return 0;
}

extern "C" void mpc_batch_destroy(mpc_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->graphs) {
        if (g.second.e) (void)hipGraphExecDestroy(g.second.e);
        if (g.second.g) (void)hipGraphDestroy(g.second.g);
    }
    for (auto &d : b->pool) (void)hipFree(d.second);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int mpc_batch_run(mpc_batch *b, int step0, int nsteps, int use_graph) {
    if (!b || nsteps < 0) return MPC_EINVAL;
    if (nsteps == 0) return 0;
    return b->prec == MPC_F32 ? batch_run<float>(b, step0, nsteps, use_graph != 0)
                              : batch_run<double>(b, step0, nsteps, use_graph != 0);
}

extern "C" int mpc_batch_graphs(mpc_batch *b) { return b ? (int)b->graphs.size() : MPC_EINVAL; }

extern "C" int mpc_set_timing(mpc_ctx *c, int enable) {
    if (!c) return MPC_EINVAL;
    c->timing = enable;
    c->evn = 0;
    for (int i = 0; i < 3; i++) { c->kms[i] = 0; c->kcount[i] = 0; }
    return 0;
}
extern "C" double mpc_kernel_ms(mpc_ctx *c, int which) {
    if (!c || which < 0 || which > 2 || c->kcount[which] == 0) return -1.0;
    return c->kms[which] / (double)c->kcount[which];
}

static int finish(mpc_ctx *c, int rc) {
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->timing) resolve_timing(c);
    return 0;
}

/* sum |A|^2 of the resident field, in the field's own scale */
static int field_energy(mpc_ctx *c, double *e) {
    const int nb = 1024;
    RCHK(c->prec == MPC_F32 ? launch_energy_partials<float>((const cx<float> *)c->field, c->N, c->err_part, nb, c->stream)
                            : launch_energy_partials<double>((const cx<double> *)c->field, c->N, c->err_part, nb, c->stream));
    double part[1024];
    HIPCHK(hipMemcpyAsync(part, c->err_part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += part[i];
    *e = s;
    return 0;
}

extern "C" int mpc_energy(mpc_ctx *c, double *energy_sum) {
    if (!c || !energy_sum) return MPC_EINVAL;
    double e = 0.0;
    RCHK(field_energy(c, &e));
    /* the t-spectral field holds F_t(A), whose energy is Nt times larger */
    *energy_sum = e * c->ascale * c->ascale / (c->domain_freq ? (double)c->nt : 1.0);
    return 0;
}

extern "C" int mpc_set_guard(mpc_ctx *c, double gain_tol) {
    if (!c) return MPC_EINVAL;
    c->guard_tol = gain_tol;
    return 0;
}

extern "C" int mpc_run(mpc_ctx *c, int step0, int nsteps) {
    if (!c || nsteps < 0) return MPC_EINVAL;
    if (nsteps == 0) return 0;
    const bool guard = c->guard_tol >= 0.0;
    double e0 = 0.0, e1 = 0.0;
    if (guard) RCHK(mpc_energy(c, &e0));
    const int rc = c->prec == MPC_F32 ? run_fixed<float>(c, step0, nsteps) : run_fixed<double>(c, step0, nsteps);
    RCHK(finish(c, rc));
    if (!guard) return 0;
    RCHK(mpc_energy(c, &e1));
    return mpc_guard_verdict(e0, e1, 0.0, 0.0, c->guard_tol, -1.0, -1.0);
}

/* one isolated split step (propagation.c:111-372 semantics) */
/* src (nullptr: c's field) is the step's input; the step's first sweep reads
 * it and writes c's field */
template <class T> static int step_once(mpc_ctx *c, double dzz, double z, int *signmem, const void *src = nullptr) {
    int s;
    if (c->p.mirror_R > 0.0) {
        s = 1;
        if (signmem) *signmem = 1;
    } else {
        s = pass_sign(c, z);
        if (signmem && -*signmem == s) *signmem = -*signmem;   /* propagation.c:380-382 */
    }
    RCHK(run_xy<T>(c, s * dzz / 2, false, nullptr, src));
    const cx<T> *P = get_table<T>(c, 't', dzz / 2);
    if (!P) return MPC_ENOMEM;
    RCHK(sweep_t<T>(c, P, P, false, true, true, dzz));
    return run_xy<T>(c, s * dzz / 2, false);
}

/* the fused end of an adaptive attempt: C1's buffer and the partial sums
 * the last y sweep of C2 writes (launch_axis_richardson) */
struct RichOut {
    const void *c1 = nullptr;
    double *err_part = nullptr, *lp_part = nullptr;
};

/*
 * The two branches of an adaptive attempt on a t-spectral field (the
 * unnormalised t-FFT of A, the state the fixed-step schedule keeps between
 * steps), in and out: the dispersion half steps ride on the y sweeps, so each
 * t-line runs only its own transforms (no forward transform of its input, no
 * inverse one of its output).  The half step before each t-line takes a 't'
 * table (the 1/nt of the pair); the half step that ends a branch takes an 'n'
 * table, so the branch leaves the unnormalised scaling it started from.
 * C1 (step_once_f) is one step of 2dz; C2 (two_steps_f) two steps of dz with
 * the middle half steps merged (their linear factors commute): x y | t | x y
 * | t | x y, 8 sweeps instead of 10.  Signs and signmem as step_once gives
 * them, once per step.
 */
template <class T> static int step_once_f(mpc_ctx *c, double dzz, double z, int *signmem, const void *src) {
    int s;
    if (c->p.mirror_R > 0.0) {
        s = 1;
        if (signmem) *signmem = 1;
    } else {
        s = pass_sign(c, z);
        if (signmem && -*signmem == s) *signmem = -*signmem;
    }
    const cx<T> *P = get_table<T>(c, 't', dzz / 2), *Pn = get_table<T>(c, 'n', dzz / 2);
    if (!P || !Pn) return MPC_ENOMEM;
    RCHK(run_xy<T>(c, s * dzz / 2, false, P, src));
    RCHK(sweep_t<T>(c, nullptr, nullptr, true, true, false, dzz));
    return run_xy<T>(c, s * dzz / 2, false, Pn);
}
template <class T> static int two_steps_f(mpc_ctx *c, double dzz, double z, int *signmem, const void *src,
                                          const RichOut *rich) {
    int s[2];
    for (int i = 0; i < 2; i++) {
        const double zi = z + i * dzz;
        if (c->p.mirror_R > 0.0) {
            s[i] = 1;
            if (signmem) *signmem = 1;
        } else {
            s[i] = pass_sign(c, zi);
            if (signmem && -*signmem == s[i]) *signmem = -*signmem;
        }
    }
    const cx<T> *P = get_table<T>(c, 't', dzz / 2), *P2 = get_table<T>(c, 't', dzz);
    const cx<T> *Pn = get_table<T>(c, 'n', dzz / 2), *D = get_table<T>(c, 'x', s[1] * dzz / 2);
    if (!P || !P2 || !Pn || !D) return MPC_ENOMEM;
    RCHK(run_xy<T>(c, s[0] * dzz / 2, false, P, src));
    RCHK(sweep_t<T>(c, nullptr, nullptr, true, true, false, dzz));
    RCHK(run_xy<T>(c, (s[0] + s[1]) * dzz / 2, false, P2));
    RCHK(sweep_t<T>(c, nullptr, nullptr, true, true, false, dzz));
    /* last x sweep, then the y sweep (with the closing half step's t factor)
     * fused with the error norms, the per-line powers and the Richardson
     * combination */
    RCHK(sweep_axis<T>(c, 0, D, false));
    AArgs<T> a = axis_args<T>(c, 1, D, false, Pn);
    tic(c, 2);
    const int rc = launch_axis_richardson<T>(a, c->ny, (const cx<T> *)rich->c1, rich->err_part, rich->lp_part,
                                             c->stream);
    toc(c, 2);
    return rc;
}

/* a t-spectral field (left so by mpc_run or mpc_run_adaptive) back to the
 * time domain, in place: the 1/nt of the transform pair, then the inverse
 * transform.  Always with the per-group t-line form, whatever variant the
 * context is tuned to (in FP64 the forms' twiddle products round
 * differently), so the time-domain field does not depend on the tuning. */
template <class T> static int to_time(mpc_ctx *c) {
    if (!c->domain_freq) return 0;
    const cx<T> *Q0 = get_table<T>(c, 't', 0.0);
    if (!Q0) return MPC_ENOMEM;
    const int keep = c->tvar;
    c->tvar = 1;
    const int rc = sweep_t<T>(c, nullptr, Q0, true, false, true, 0.0);
    c->tvar = keep;
    RCHK(rc);
    c->domain_freq = 0;
    return 0;
}
static int to_time_any(mpc_ctx *c) { return c->prec == MPC_F32 ? to_time<float>(c) : to_time<double>(c); }
/* fn on a time-domain view of c's field that leaves c as it was (a read: a
 * t-spectral field is copied to aux1 and transformed there), so reading the
 * field never changes the path a later run takes */
template <class F> static int time_view(mpc_ctx *c, F &&fn) {
    if (!c->domain_freq) return fn();
    const size_t bytes = c->esz * (size_t)c->N;
    if (!c->aux1 && hipMalloc(&c->aux1, bytes) != hipSuccess) return MPC_ENOMEM;
    HIPCHK(hipMemcpyAsync(c->aux1, c->field, bytes, hipMemcpyDeviceToDevice, c->stream));
    void *keep = c->field;
    c->field = c->aux1;
    int rc = to_time_any(c);
    if (!rc) rc = fn();
    c->field = keep;
    c->domain_freq = 1;
    return rc;
}

extern "C" int mpc_step(mpc_ctx *c, double dzz, double z, int *signmem) {
    if (!c) return MPC_EINVAL;
    RCHK(to_time_any(c));
    const int rc = c->prec == MPC_F32 ? step_once<float>(c, dzz, z, signmem) : step_once<double>(c, dzz, z, signmem);
    return finish(c, rc);
}

/* the linear part of propagation over `length` in one go (dispersion on the
 * t-lines, diffraction with coefficient sign * length on x and y, then the
 * mirror if asked): the split step with the nonlinear step left out, whose
 * linear factors commute and so merge into one table each */
template <class T> static int linear_once(mpc_ctx *c, double length, int sign, bool mirror) {
    const cx<T> *P = get_table<T>(c, 't', length);
    if (!P) return MPC_ENOMEM;
    c->rho_nvalid = 0;     /* the field no longer is the one the plasma maxima came from */
    RCHK(sweep_t<T>(c, P, nullptr, false, false, true, length));
    return run_xy<T>(c, sign * length, mirror);
}

extern "C" int mpc_linear(mpc_ctx *c, double length, int sign, int mirror) {
    if (!c || !(length >= 0.0) || (sign != 1 && sign != -1)) return MPC_EINVAL;
    RCHK(to_time_any(c));
    const int rc = c->prec == MPC_F32 ? linear_once<float>(c, length, sign, mirror != 0)
                                      : linear_once<double>(c, length, sign, mirror != 0);
    return finish(c, rc);
}

extern "C" int mpc_mirror(mpc_ctx *c) {
    if (!c) return MPC_EINVAL;
    /* x/y sweeps with identity propagators, mirror applied in the y sweep */
    const int rc = c->prec == MPC_F32 ? run_xy<float>(c, 0.0, true) : run_xy<double>(c, 0.0, true);
    return finish(c, rc);
}

/* ------------------------------------------------------------------------ */
/* transfers                                                                 */
/* ------------------------------------------------------------------------ */
/* smallest power of two >= v (1 for v <= 0 or non-finite) */
static double pow2_above(double v) {
    if (!(v > 0.0) || !std::isfinite(v)) return 1.0;
    int e = 0;
    std::frexp(v, &e);
    return std::ldexp(1.0, e);
}

/* work on c->stream that writes the staging buffer waits for a snapshot's
 * copy out of it */
static int staging_guard(mpc_ctx *c) {
    if (c->snap_pending) HIPCHK(hipStreamWaitEvent(c->stream, c->snap_done, 0));
    return 0;
}

extern "C" int mpc_upload(mpc_ctx *c, const cufftDoubleComplex *A) {
    if (!c || !A) return MPC_EINVAL;
    RCHK(staging_guard(c));
    HIPCHK(hipMemcpyAsync(c->staging, A, 16 * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
    /* pick the power-of-two amplitude scale of the resident field */
    const int nb = 1024;
    RCHK(launch_maxabs(c->staging, 2 * c->N, c->err_part, nb, c->stream));
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->err_part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double mx = 0.0;
    for (double v : part) mx = std::max(mx, v);
    c->ascale = pow2_above(mx);
    const double inv = 1.0 / c->ascale;
    const int rc = c->prec == MPC_F32
        ? launch_to_internal<float>(c->staging, (cx<float> *)c->field, c->nx, c->ny, c->nt, inv, c->stream)
        : launch_to_internal<double>(c->staging, (cx<double> *)c->field, c->nx, c->ny, c->nt, inv, c->stream);
    c->domain_freq = 0;
    c->rho_nvalid = 0;     /* no nonlinear step has seen this field yet */
    return finish(c, rc);
}

extern "C" int mpc_download(mpc_ctx *c, cufftDoubleComplex *A) {
    if (!c || !A) return MPC_EINVAL;
    RCHK(staging_guard(c));
    RCHK(time_view(c, [&] {
        return c->prec == MPC_F32
            ? launch_to_reference<float>((const cx<float> *)c->field, c->staging, c->nx, c->ny, c->nt, c->ascale, c->stream)
            : launch_to_reference<double>((const cx<double> *)c->field, c->staging, c->nx, c->ny, c->nt, c->ascale, c->stream);
    }));
    HIPCHK(hipMemcpyAsync(A, c->staging, 16 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    return finish(c, 0);
}

/* Checkpoint snapshot (the reference writes resdata.bin synchronously,
 * MPC.c:521-544): the field in the reference layout goes to the staging
 * buffer on the context's stream, then to a pinned host buffer on a copy
 * stream, while the caller goes on stepping. */
extern "C" int mpc_snapshot_begin(mpc_ctx *c) {
    if (!c) return MPC_EINVAL;
    /* in place: an interrupted run and the run resumed from this snapshot
     * then go on from the same time-domain field */
    RCHK(to_time_any(c));
    if (!c->copy_stream) {
        HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
    }
    if (!c->pinned && hipHostMalloc((void **)&c->pinned, 16 * (size_t)c->N, hipHostMallocDefault) != hipSuccess) {
        c->pinned = nullptr;
        return MPC_ENOMEM;
    }
    /* the previous snapshot's copy must be out of the pinned buffer and of staging */
    if (c->snap_pending) HIPCHK(hipEventSynchronize(c->snap_done));
    c->snap_pending = 0;
    const int rc = c->prec == MPC_F32
        ? launch_to_reference<float>((const cx<float> *)c->field, c->staging, c->nx, c->ny, c->nt, c->ascale, c->stream)
        : launch_to_reference<double>((const cx<double> *)c->field, c->staging, c->nx, c->ny, c->nt, c->ascale, c->stream);
    RCHK(rc);
    HIPCHK(hipEventRecord(c->snap_ready, c->stream));
    HIPCHK(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
    HIPCHK(hipMemcpyAsync(c->pinned, c->staging, 16 * (size_t)c->N, hipMemcpyDeviceToHost, c->copy_stream));
    HIPCHK(hipEventRecord(c->snap_done, c->copy_stream));
    c->snap_pending = 1;
    return 0;
}

extern "C" int mpc_snapshot_poll(mpc_ctx *c) {
    if (!c) return MPC_EINVAL;
    if (!c->snap_pending) return 1;
    const hipError_t e = hipEventQuery(c->snap_done);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    fprintf(stderr, "mpc: HIP error %s in mpc_snapshot_poll\n", hipGetErrorString(e));
    return MPC_EHIP;
}

extern "C" int mpc_snapshot_wait(mpc_ctx *c, const cufftDoubleComplex **A) {
    if (!c || !A) return MPC_EINVAL;
    if (!c->pinned || !c->snap_done) return MPC_EINVAL;     /* no snapshot was taken */
    if (c->snap_pending) HIPCHK(hipEventSynchronize(c->snap_done));
    c->snap_pending = 0;
    *A = c->pinned;
    return 0;
}

/* the resident (time-domain) field as two float planes on the device, in the
 * internal point order [y][x][t], times scale; and back */
extern "C" int mpc_export_planar(mpc_ctx *c, float *planes, double scale) {
    if (!c || !planes) return MPC_EINVAL;
    const double s = scale * c->ascale;
    const int rc = time_view(c, [&] {
        return c->prec == MPC_F32 ? launch_planar<float>((cx<float> *)c->field, planes, c->N, s, 1, c->stream)
                                  : launch_planar<double>((cx<double> *)c->field, planes, c->N, s, 1, c->stream);
    });
    return finish(c, rc);
}
extern "C" int mpc_import_planar(mpc_ctx *c, const float *planes, double scale) {
    if (!c || !planes || !(scale != 0.0)) return MPC_EINVAL;
    const double s = 1.0 / (scale * c->ascale);
    const int rc = c->prec == MPC_F32
        ? launch_planar<float>((cx<float> *)c->field, (float *)planes, c->N, s, 0, c->stream)
        : launch_planar<double>((cx<double> *)c->field, (float *)planes, c->N, s, 0, c->stream);
    c->domain_freq = 0;
    c->rho_nvalid = 0;
    return finish(c, rc);
}

extern "C" int mpc_calibrate(mpc_ctx *c) {
    if (!c || !(c->p.Esig > 0)) return MPC_EINVAL;
    const double k = std::sqrt(c->p.puls_e / c->p.Esig);   /* MPC.c:209 */
    /* the power-of-two part of k goes into ascale, the rest into the data */
    const double k2 = pow2_above(k) / 2.0;
    const double kr = k / k2;
    c->ascale *= k2;
    const int rc = c->prec == MPC_F32 ? launch_scale<float>((cx<float> *)c->field, c->N, (float)kr, c->stream)
                                      : launch_scale<double>((cx<double> *)c->field, c->N, kr, c->stream);
    return finish(c, rc);
}

/* ------------------------------------------------------------------------ */
/* diagnostics                                                               */
/* ------------------------------------------------------------------------ */
extern "C" int mpc_diagnostics(mpc_ctx *c, mpc_diag *d) {
    if (!c || !d) return MPC_EINVAL;
    /* either domain: a t-spectral field gives spec directly and temp and the
     * line powers through the inverse transform */
    const int nb = c->diag_nblocks;
    double *tp = c->diag_part, *sp = c->diag_part + (size_t)nb * c->nt;
    const bool spec = c->domain_freq != 0;
    int rc = c->prec == MPC_F32
        ? launch_diag_lines<float>((const cx<float> *)c->field, (const cx<float> *)c->tw_t, c->nt, c->nxy, c->diag_line, tp, sp, nb, spec, c->stream)
        : launch_diag_lines<double>((const cx<double> *)c->field, (const cx<double> *)c->tw_t, c->nt, c->nxy, c->diag_line, tp, sp, nb, spec, c->stream);
    RCHK(rc);
    RCHK(launch_reduce_cols(tp, nb, c->nt, c->diag_out, c->stream));
    RCHK(launch_reduce_cols(sp, nb, c->nt, c->diag_out + c->nt, c->stream));
    std::vector<double> ts(2 * (size_t)c->nt), lp((size_t)c->nxy), rb;
    HIPCHK(hipMemcpyAsync(ts.data(), c->diag_out, sizeof(double) * 2 * c->nt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lp.data(), c->diag_line, sizeof(double) * c->nxy, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double sc2 = c->ascale * c->ascale;
    for (double &v : ts) v *= sc2;
    for (double &v : lp) v *= sc2;
    double e = 0.0;
    for (int k = 0; k < c->nt; k++) e += ts[k];
    d->energy_sum = e;
    if (d->temp) std::copy(ts.begin(), ts.begin() + c->nt, d->temp);
    if (d->spec) std::copy(ts.begin() + c->nt, ts.end(), d->spec);
    /* rms spectral width over the FFT-order Omega grid */
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < c->nt; k++) {
        const double om = c->w[(k + c->nt / 2) % c->nt], sv = ts[c->nt + k];
        s0 += sv; s1 += om * sv; s2 += om * om * sv;
    }
    d->spec_rms_width = s0 > 0 ? std::sqrt(std::max(0.0, s2 / s0 - (s1 / s0) * (s1 / s0))) : 0.0;
    /* BP1[x] = sum_{y,t}, BP2[y] = sum_{x,t} (propagation.c:343-354) */
    /* BP1 row by row (each BP1[i] still adds its lines in j order, as a
     * column walk would, with contiguous reads) */
    if (d->BP1) {
        std::fill(d->BP1, d->BP1 + c->nx, 0.0);
        for (int j = 0; j < c->ny; j++)
            for (int i = 0; i < c->nx; i++) d->BP1[i] += lp[(size_t)j * c->nx + i];
    }
    if (d->BP2) for (int j = 0; j < c->ny; j++) { double s = 0; for (int i = 0; i < c->nx; i++) s += lp[(size_t)j * c->nx + i]; d->BP2[j] = s; }
    d->rho_peak = 0.0;
    if (c->p.plasm == 1) {
        rb.resize((size_t)std::min<long>(c->rho_nvalid, c->rho_nblocks));
        if (!rb.empty())
            HIPCHK(hipMemcpy(rb.data(), c->rho_block, sizeof(double) * rb.size(), hipMemcpyDeviceToHost));
        double mx = 0.0;
        for (double v : rb) mx = std::max(mx, v);
        d->rho_peak = mx * c->p.rho_nt;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* adaptive step doubling (MPC.c:343-383, propagation.c:616-677)             */
/* ------------------------------------------------------------------------ */
void mpc::setstep_decide(double *dzz, double err, double def_err, double dzmin, const double *y1,
                         const double *y2, int *bad) {
    *bad = 0;
    const double f = std::pow(2.0, 1.0 / 3.0);
    if (err > 2.0 * def_err) {
        const double u = *dzz / 2.0;
        if (u > dzmin) { *dzz = u; *bad = 1; return; }
    } else if (err > def_err && err < 2.0 * def_err) {
        const double u = *dzz / f;
        if (u > dzmin) *dzz = u;
    } else if (err > 0.5 * def_err && err <= def_err) {
    } else {
        const double u = *dzz * f;
        if (!(u > 100 * (y2[1] + y1[1]) / 2.0)) *dzz = u;
    }
}

/*
 * Adaptive step doubling + Richardson (MPC.c:343-383): per attempt C1 = one
 * 2dz step and C2 = two dz steps, both from A.  C2's last sweep is fused
 * with the error norms, the per-line powers of C2 (BP1/BP2 -> calc1oes, the
 * step-size cap) and R = 4/3 C2 - 1/3 C1, which it leaves in C2's buffer;
 * an accepted step adopts that buffer as the field.  One host round trip per
 * attempt brings back the two norms and the line powers.  The attempts run
 * on the t-spectral field: the norms are sums of |.|^2 over whole lines, and
 * the line powers feed only calc1oes's normalised profile, so Parseval makes
 * both the same in either domain (the spectral line powers are nt times the
 * temporal ones).  The field enters in either domain and leaves t-spectral,
 * so consecutive calls (one accepted step each, as the driver makes them)
 * transform nothing between them.
 */
template <class T> static int adaptive(mpc_ctx *c, double *z, double *dzz, int *signmem, int max_steps,
                                       int *accepted, double *errs, double *dzs) {
    const size_t bytes = c->esz * (size_t)c->N;
    if (!c->aux1 && hipMalloc(&c->aux1, bytes) != hipSuccess) return MPC_ENOMEM;
    if (!c->aux2 && hipMalloc(&c->aux2, bytes) != hipSuccess) return MPC_ENOMEM;
    const long ntiles = axis_richardson_tiles(c->prec, c->nt, c->nx);
    const long nchunks = ntiles / c->nx;
    if (!c->rich_err && hipMalloc((void **)&c->rich_err, sizeof(double) * (2 * (size_t)ntiles + 2)) != hipSuccess)
        return MPC_ENOMEM;
    if (!c->rich_lp && hipMalloc((void **)&c->rich_lp, sizeof(double) * (size_t)nchunks * c->nxy) != hipSuccess)
        return MPC_ENOMEM;
    int acc = 0, guard = 0;
    std::vector<double> BP1(c->nx), BP2(c->ny), lp((size_t)c->nxy);
    double y1[4] = {0, 0, 0, 0}, y2[4] = {0, 0, 0, 0}, bps = 0, norms[2];
    RichOut ro;
    ro.err_part = c->rich_err;
    ro.lp_part = c->rich_lp;
    bool spectral = c->domain_freq != 0;
    while (*z < c->p.dist && acc < max_steps && guard < 64 * max_steps + 64) {
        guard++;
        const double dz = *dzz;
        void *A = c->field;
        /* the attempts run on the t-spectral field: one forward t-transform
         * before the first, one inverse after the last */
        if (!spectral) {
            RCHK(sweep_t<T>(c, nullptr, nullptr, false, false, false, 0.0));
            c->domain_freq = 1;
            spectral = true;
            A = c->field;
        }
        /* C1 and C2 start from A without copies: each one's first sweep reads
         * A and writes its own buffer */
        int sm1 = *signmem;
        c->field = c->aux1;
        int rc = step_once_f<T>(c, 2 * dz, *z, &sm1, A);
        c->field = c->aux2;
        ro.c1 = c->aux1;
        if (!rc) rc = two_steps_f<T>(c, dz, *z, signmem, A, &ro);   /* sub-steps at z and z + dz */
        c->field = A;
        RCHK(rc);
        RCHK(launch_sum_pairs(c->rich_err, ntiles, c->rich_err + 2 * ntiles, c->stream));
        RCHK(launch_reduce_cols(c->rich_lp, (int)nchunks, (int)c->nxy, c->diag_line, c->stream));
        HIPCHK(hipMemcpyAsync(norms, c->rich_err + 2 * ntiles, sizeof norms, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(lp.data(), c->diag_line, sizeof(double) * c->nxy, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->timing) resolve_timing(c);
        /* BP and calc1oes of C2 for the step-size cap (propagation.c:343-362) */
        std::fill(BP1.begin(), BP1.end(), 0.0);          /* row by row; each BP1[i] adds in j order */
        for (int j = 0; j < c->ny; j++)
            for (int i = 0; i < c->nx; i++) BP1[i] += lp[(size_t)j * c->nx + i];
        for (int j = 0; j < c->ny; j++) { double sm = 0; for (int i = 0; i < c->nx; i++) sm += lp[(size_t)j * c->nx + i]; BP2[j] = sm; }
        int dims[3] = {c->nx, c->ny, c->nt};
        calc1oes(BP1.data(), dims, c->r.data(), &bps, y1);
        calc1oes(BP2.data(), dims, c->r.data(), &bps, y2);
        const double err = std::sqrt(norms[0]) / std::sqrt(norms[1]);
        int bad = 0;
        setstep_decide(dzz, err, c->p.def_err, c->p.dzmin, y1, y2, &bad);
        if (bad) continue;
        /* accept: the Richardson field is in C2's buffer */
        c->field = c->aux2;
        c->aux2 = A;
        const double zold = *z;
        *z = *z + 2.0 * dz;
        if (*z > c->p.dist) *z = c->p.dist;
        if (std::floor(*z / c->p.Cav) > std::floor(zold / c->p.Cav) && *z < c->p.dist)
            RCHK(run_xy<T>(c, 0.0, true));
        if (errs) errs[acc] = err;
        if (dzs) dzs[acc] = *dzz;
        acc++;
    }
    /* the field stays t-spectral for the next call; the calls that need it
     * in time bring it back themselves (to_time) */
    if (accepted) *accepted = acc;
    return finish(c, 0);
}

extern "C" int mpc_run_adaptive(mpc_ctx *c, double *z, double *dzz, int *signmem, int max_steps,
                                int *accepted, double *errs, double *dzs) {
    if (!c || !z || !dzz || !signmem) return MPC_EINVAL;
    return c->prec == MPC_F32 ? adaptive<float>(c, z, dzz, signmem, max_steps, accepted, errs, dzs)
                              : adaptive<double>(c, z, dzz, signmem, max_steps, accepted, errs, dzs);
}

/* ------------------------------------------------------------------------ */
/* FFT / tile autotuner                                                      */
/* ------------------------------------------------------------------------ */
/* times every kernel variant (median of 5 samples after a warm-up launch)
 * on a scratch copy of the resident field and keeps the fastest; the choice
 * only changes speed.  A sample is enough back-to-back launches to last
 * about 0.2 ms (at most 64), so that launches of a few microseconds on small
 * fields are not ranked by timer noise; the t-line forms' samples
 * interleave. */
template <class T> static int autotune(mpc_ctx *c) {
    const size_t bytes = c->esz * (size_t)c->N;
    if (!c->aux1 && hipMalloc(&c->aux1, bytes) != hipSuccess) return MPC_ENOMEM;
    HIPCHK(hipMemcpyAsync(c->aux1, c->field, bytes, hipMemcpyDeviceToDevice, c->stream));
    void *keep = c->field;
    const int keep_t = c->tvar, keep_a = c->avar, keep_y = c->avar_y, keep_timing = c->timing;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return MPC_EHIP;
    }
    c->field = c->aux1;
    c->timing = 0;
    /* the scratch copy in the state a run keeps between steps (t-spectral,
     * unnormalised), so the timed t-lines see the run's intensities */
    int rc0 = c->domain_freq ? 0 : sweep_t<T>(c, nullptr, nullptr, false, false, false, 0.0);
    /* one warm-up launch (first-use costs; its time sizes the samples), then
     * the median of five samples of k launches each */
    auto time_median = [&](auto &&launch) -> float {
        auto timed = [&](int k) -> float {
            (void)hipEventRecord(e0, c->stream);
            for (int i = 0; i < k; i++)
                if (launch() != 0) return 1e30f;
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            float ms = 1e30f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return ms / (float)k;
        };
        const float t0 = timed(1);
        if (t0 >= 1e30f) return 1e30f;
        const int k = t0 > 0.0f && t0 < 0.2f ? std::min(64, (int)std::ceil(0.2f / t0)) : 1;
        float t[5];
        for (int r = 0; r < 5; r++) {
            t[r] = timed(k);
            if (t[r] >= 1e30f) return 1e30f;
        }
        std::sort(t, t + 5);
        return t[2];
    };
    const double dz = c->p.dzz;
    const cx<T> *D = get_table<T>(c, 'x', dz);
    const cx<T> *P = get_table<T>(c, 't', dz);
    int rc = rc0 ? rc0 : (!D || !P) ? MPC_ENOMEM : 0;
    float bestX = 1e30f, bestY = 1e30f, bestT = 1e30f;
    int vx = keep_a, vy = keep_y, vt = keep_t;
    for (int v = 0; !rc && v < axis_variants(); v++) {
        c->avar = c->avar_y = v;
        const float mx = time_median([&] { return sweep_axis<T>(c, 0, D, false); });
        const float my = time_median([&] { return sweep_axis<T>(c, 1, D, false); });
        if (mx < bestX) { bestX = mx; vx = v; }
        if (my < bestY) { bestY = my; vy = v; }
    }
    /* The t-line forms are timed over whole steps (x sweep, y sweep, t-line,
     * with the x/y variants just chosen), since what the sweeps before it
     * leave in the caches differs from back-to-back t-lines; their samples
     * interleave (each step advances the scratch field's nonlinear state,
     * which the plasma work depends on). */
    {
        const int nvt = tline_variants();
        std::vector<std::vector<float>> ts(nvt);
        c->avar = vx;
        c->avar_y = vy;
        auto launch_t = [&] {
            int r = sweep_axis<T>(c, 0, D, false);
            if (!r) r = sweep_axis<T>(c, 1, D, false);
            if (!r) r = sweep_t<T>(c, P, nullptr, true, true, false, dz);
            return r;
        };
        auto sample = [&](int k) -> float {
            (void)hipEventRecord(e0, c->stream);
            for (int i = 0; i < k; i++)
                if (launch_t() != 0) return 1e30f;
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            float ms = 1e30f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return ms / (float)k;
        };
        float tw = 1e30f;                                        /* warm-up; sizes the samples */
        for (int v = 0; !rc && v < nvt; v++) {
            c->tvar = v;
            tw = std::min(tw, sample(1));
        }
        const int k = tw > 0.0f && tw < 0.2f ? std::min(64, (int)std::ceil(0.2f / tw)) : 1;
        for (int r = 0; !rc && r < 5; r++)
            for (int v = 0; v < nvt; v++) {
                c->tvar = v;
                ts[v].push_back(sample(k));
            }
        for (int v = 0; !rc && v < nvt; v++) {
            std::sort(ts[v].begin(), ts[v].end());
            const float ms = ts[v][ts[v].size() / 2];
            if (ms < bestT) { bestT = ms; vt = v; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->field = keep;
    c->timing = keep_timing;
    c->avar = rc ? keep_a : vx;
    c->avar_y = rc ? keep_y : vy;
    c->tvar = rc ? keep_t : vt;
    if (rc) return rc;
    c->tuned = 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mpc_autotune(mpc_ctx *c) {
    if (!c) return MPC_EINVAL;
    return c->prec == MPC_F32 ? autotune<float>(c) : autotune<double>(c);
}
extern "C" int mpc_get_variants(mpc_ctx *c, int *tv, int *av) {
    if (!c) return MPC_EINVAL;
    if (tv) *tv = c->tvar;
    if (av) *av = c->avar;
    return 0;
}
extern "C" int mpc_variant_counts(int *tv, int *av) {
    if (tv) *tv = tline_variants();
    if (av) *av = axis_variants();
    return 0;
}
extern "C" int mpc_set_variants(mpc_ctx *c, int tv, int av) {
    return mpc_set_variants_xy(c, tv, av, av);
}
extern "C" int mpc_get_variants_xy(mpc_ctx *c, int *tv, int *xv, int *yv) {
    if (!c) return MPC_EINVAL;
    if (tv) *tv = c->tvar;
    if (xv) *xv = c->avar;
    if (yv) *yv = c->avar_y;
    return 0;
}
extern "C" int mpc_set_variants_xy(mpc_ctx *c, int tv, int xv, int yv) {
    const int na = axis_variants();
    if (!c || tv < 0 || tv >= tline_variants() || xv < 0 || xv >= na || yv < 0 || yv >= na) return MPC_EINVAL;
    c->tvar = tv;
    c->avar = xv;
    c->avar_y = yv;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* legacy entry points (propagation.h / plasma.h)                            */
/* ------------------------------------------------------------------------ */
struct mpc_plan_s {
    struct Shared {
        int dims[3];
        int refs;
        mpc_ctx *ctx;
        std::string wisdom;   /* "<cpath><Nx>": the kernel-variant cache */
        int readin;
        int tvar, avar, avar_y; /* from the cache, or -1 until tuned */
    } *sh;
    int axis, dir;
};

static std::mutex g_legacy_mu;

static void set_err(int rc) {
    if (rc) {
        g_last_error = rc;
        fprintf(stderr, "mpc: %s\n", mpc_strerror(rc));
    }
}

extern "C" void createPlans(int readin, int *dims, const char *cpath, fftw_plan *pR1fft, fftw_plan *pR1ifft,
                            fftw_plan *pR2fft, fftw_plan *pR2ifft, fftw_plan *pTfft, fftw_plan *pTifft) {
    auto *sh = new mpc_plan_s::Shared();
    std::memcpy(sh->dims, dims, sizeof sh->dims);
    sh->refs = 6;
    sh->ctx = nullptr;
    if (cpath) sh->wisdom = std::string(cpath) + std::to_string(dims[0]);
    fftw_plan *outs[6] = {pR1fft, pR1ifft, pR2fft, pR2ifft, pTfft, pTifft};
    const int ax[6] = {0, 0, 1, 1, 2, 2}, dr[6] = {-1, 1, -1, 1, -1, 1};
    for (int i = 0; i < 6; i++) {
        auto *p = new mpc_plan_s();
        p->sh = sh;
        p->axis = ax[i];
        p->dir = dr[i];
        *outs[i] = p;
    }
    /* the autotuner cache, read like FFTW wisdom (propagation.c:540-552);
     * it is written after the first tuning (:609-611) */
    sh->readin = readin;
    sh->tvar = sh->avar = sh->avar_y = -1;
    if (readin && !sh->wisdom.empty()) {
        FILE *f = fopen(sh->wisdom.c_str(), "r");
        if (f) {
            int t = -1, a = -1, y = -1;
            const int nr = fscanf(f, "mpc-kernel-variants tline=%d axis=%d yaxis=%d", &t, &a, &y);
            if (nr >= 2) { sh->tvar = t; sh->avar = a; sh->avar_y = nr == 3 ? y : a; }
            else printf("Importing wisdom failed, creating new one to save.\n");
            fclose(f);
        }
    }
}

/* apply the cached kernel variants, or tune on the uploaded field and cache */
static int legacy_tune(mpc_plan_s::Shared *sh) {
    mpc_ctx *c = sh->ctx;
    if (c->tuned) return 0;
    if (sh->tvar >= 0 && mpc_set_variants_xy(c, sh->tvar, sh->avar, sh->avar_y) == 0) { c->tuned = 1; return 0; }
    RCHK(mpc_autotune(c));
    sh->tvar = c->tvar;
    sh->avar = c->avar;
    sh->avar_y = c->avar_y;
    if (!sh->wisdom.empty()) {
        FILE *f = fopen(sh->wisdom.c_str(), "w");
        if (f) {
            fprintf(f, "mpc-kernel-variants tline=%d axis=%d yaxis=%d nx=%d ny=%d nt=%d\n", c->tvar, c->avar,
                    c->avar_y, sh->dims[0], sh->dims[1], sh->dims[2]);
            fclose(f);
        }
    }
    return 0;
}

extern "C" void mpc_destroy_plan(fftw_plan p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    auto *sh = p->sh;
    delete p;
    if (--sh->refs == 0) {
        mpc_destroy(sh->ctx);
        delete sh;
    }
}

/* a context for the plan's dims carrying the caller's parameters */
static int legacy_ctx(fftw_plan plan, const mpc_params &p, mpc_ctx **out) {
    auto *sh = plan->sh;
    if (!sh->ctx) RCHK(mpc_create(&p, MPC_F64, &sh->ctx));
    else {
        const mpc_params &o = sh->ctx->p;
        const bool same = o.dzz == p.dzz && o.alpha == p.alpha && o.Cav == p.Cav && o.k0 == p.k0 &&
                          o.gamma2 == p.gamma2 && o.w0 == p.w0 && o.sst_on == p.sst_on && o.plasm == p.plasm &&
                          o.deltat == p.deltat && o.rho_c == p.rho_c && o.rho_nt == p.rho_nt && o.n0 == p.n0 &&
                          o.Ab_M == p.Ab_M && std::strcmp(o.gas, p.gas) == 0 &&
                          std::equal(sh->ctx->betas.begin(), sh->ctx->betas.end(), p.betas) &&
                          std::equal(sh->ctx->w.begin(), sh->ctx->w.end(), p.w) &&
                          std::equal(sh->ctx->wr.begin(), sh->ctx->wr.end(), p.wr) &&
                          std::equal(sh->ctx->r.begin(), sh->ctx->r.end(), p.r);
        if (!same) {
            RCHK(check_params(&p));
            RCHK(ctx_set_params(sh->ctx, &p));
        }
    }
    *out = sh->ctx;
    return 0;
}

static std::vector<double> zeros_or(const double *v, int n) {
    return v ? std::vector<double>(v, v + n) : std::vector<double>(n, 0.0);
}

extern "C" void ssfmprop(cufftDoubleComplex *A, int *dims, int sst_on, double dzz, double *betas, double alpha,
                         int *signmem, double z, double Cav, double k0, double *wr, fftw_plan pR1fft,
                         fftw_plan pR1ifft, fftw_plan pR2fft, fftw_plan pR2ifft, fftw_plan pTfft, fftw_plan pTifft,
                         double *w, double gamma2, double w0, int plasm, double deltat, char *gas, double rho_c,
                         double rho_nt, double n0, double *puls_e, double *r, double *BP1, double *BP2, double *y1,
                         double *y2, double *bps1, double *bps2, double Ab_M, int change, int mpirank) {
    (void)pR1fft; (void)pR1ifft; (void)pR2fft; (void)pR2ifft; (void)pTifft; (void)mpirank;
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!pTfft) { set_err(MPC_EINVAL); return; }
    std::vector<double> rr = zeros_or(r, dims[1]);
    mpc_params p{};
    std::memcpy(p.dims, dims, sizeof p.dims);
    p.sst_on = sst_on; p.plasm = plasm; p.steps_per_pass = 1; p.npasses = 1;
    p.dzz = dzz; p.alpha = alpha; p.Cav = Cav; p.k0 = k0; p.gamma2 = gamma2; p.w0 = w0; p.deltat = deltat;
    p.rho_c = rho_c; p.rho_nt = rho_nt; p.n0 = n0; p.Ab_M = Ab_M; p.dist = Cav;
    p.betas = betas; p.wr = wr; p.w = w; p.r = rr.data();
    std::snprintf(p.gas, sizeof p.gas, "%s", gas ? gas : "Argon");
    mpc_ctx *c = nullptr;
    int rc = legacy_ctx(pTfft, p, &c);
    if (!rc) rc = mpc_upload(c, A);
    if (!rc) rc = legacy_tune(pTfft->sh);
    if (!rc) rc = mpc_step(c, dzz, z, signmem);
    if (!rc && change == 1) {
        /* propagation.c:334-341 (the absorption branch cannot trigger once
         * calcexpR has set signmem; kept for the same observable behaviour) */
        const int signum = ((int)(std::floor(z / Cav) + 1) % 2) * 2 - 1;
        if (*signmem == -signum) { *puls_e = *puls_e * Ab_M; *signmem *= -1; }
        mpc_diag d{};
        d.BP1 = BP1; d.BP2 = BP2;
        rc = mpc_diagnostics(c, &d);
        if (!rc) {
            calc1oes(BP1, dims, rr.data(), bps1, y1);
            calc1oes(BP2, dims, rr.data(), bps2, y2);
        }
    }
    if (!rc) rc = mpc_download(c, A);
    set_err(rc);
}

extern "C" void linearstep(cufftDoubleComplex *A, cuDoubleComplex *exp_D0, cuDoubleComplex *expR, fftw_plan pTfft,
                           fftw_plan pR2fft, fftw_plan pR1fft, int *dims) {
    (void)pR2fft; (void)pR1fft;
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!pTfft) { set_err(MPC_EINVAL); return; }
    auto *sh = pTfft->sh;
    int rc = 0;
    if (!sh->ctx) {
        /* a minimal parameter set: only the tables passed here are used */
        std::vector<double> zt(dims[2], 0.0), zr(dims[1], 0.0);
        mpc_params p{};
        std::memcpy(p.dims, dims, sizeof p.dims);
        p.k0 = 1.0; p.w0 = 1.0; p.Cav = 1.0; p.Ab_M = 1.0; p.steps_per_pass = 1; p.npasses = 1;
        p.betas = zt.data(); p.w = zt.data(); p.wr = zr.data(); p.r = zr.data();
        std::snprintf(p.gas, sizeof p.gas, "Argon");
        rc = mpc_create(&p, MPC_F64, &sh->ctx);
    }
    mpc_ctx *c = sh->ctx;
    if (!rc) rc = mpc_upload(c, A);
    if (!rc) {
        const cx<double> *P = user_table<double>(c, "t", exp_D0, dims[2], 1.0 / dims[2], false);
        const cx<double> *Dy = user_table<double>(c, "y", expR, dims[0], 1.0 / dims[1], true);
        const cx<double> *Dx = user_table<double>(c, "x", expR, dims[0], 1.0 / dims[0], true);
        if (!P || !Dy || !Dx) rc = MPC_ENOMEM;
        if (!rc) rc = sweep_t<double>(c, P, nullptr, false, false, true, 0.0);
        if (!rc) rc = sweep_axis<double>(c, 1, Dy, false);
        if (!rc) rc = sweep_axis<double>(c, 0, Dx, false);
        if (!rc) rc = finish(c, 0);
    }
    if (!rc) rc = mpc_download(c, A);
    set_err(rc);
}

extern "C" void setStep(int *dims, double *dzz, cufftDoubleComplex *C1, cufftDoubleComplex *C2, double def_err,
                        double dzmin, double *y1, double *y2, int *bad, double *err) {
    const long N = (long)dims[0] * dims[1] * dims[2];
    int rc = 0;
    void *d1 = nullptr, *d2 = nullptr, *part = nullptr;
    const int nb = 1024;
    if (hipMalloc(&d1, 16 * (size_t)N) != hipSuccess || hipMalloc(&d2, 16 * (size_t)N) != hipSuccess ||
        hipMalloc(&part, sizeof(double) * 2 * nb) != hipSuccess)
        rc = MPC_ENOMEM;
    std::vector<double> h(2 * nb);
    if (!rc && (hipMemcpy(d1, C1, 16 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d2, C2, 16 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess))
        rc = MPC_EHIP;
    if (!rc) rc = launch_err_partials<double>((const cx<double> *)d1, (const cx<double> *)d2, N, (double *)part, nb, nullptr);
    if (!rc && hipMemcpy(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost) != hipSuccess) rc = MPC_EHIP;
    if (d1) (void)hipFree(d1);
    if (d2) (void)hipFree(d2);
    if (part) (void)hipFree(part);
    if (rc) { set_err(rc); return; }
    double a = 0, b = 0;
    for (int i = 0; i < nb; i++) { a += h[2 * i]; b += h[2 * i + 1]; }
    *err = std::sqrt(a) / std::sqrt(b);
    printf("With step size: %.4f, the local error is: %.8g\nThe defined error is:%.8g\n ", *dzz, *err, def_err);
    setstep_decide(dzz, *err, def_err, dzmin, y1, y2, bad);
}

extern "C" void ioniz(cufftDoubleComplex *A, int *dims, double w0, double deltat, char *gas, double rho_c,
                      double *w, double rho_nt, double n0, double *tau_c, double *freeelectrons) {
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    std::vector<double> zt(w, w + dims[2]), zr(dims[1], 0.0);
    mpc_params p{};
    std::memcpy(p.dims, dims, sizeof p.dims);
    p.plasm = 1; p.w0 = w0; p.deltat = deltat; p.rho_c = rho_c; p.rho_nt = rho_nt; p.n0 = n0;
    p.k0 = w0 / C0; p.Cav = 1.0; p.Ab_M = 1.0; p.steps_per_pass = 1; p.npasses = 1;
    p.betas = zt.data(); p.w = zt.data(); p.wr = zr.data(); p.r = zr.data();
    std::snprintf(p.gas, sizeof p.gas, "%s", gas ? gas : "");
    mpc_ctx *c = nullptr;
    int rc = mpc_create(&p, MPC_F64, &c);
    void *out = nullptr;
    const long N = (long)dims[0] * dims[1] * dims[2];
    if (!rc) rc = mpc_upload(c, A);
    if (!rc && hipMalloc(&out, 16 * (size_t)N) != hipSuccess) rc = MPC_ENOMEM;
    if (!rc) {
        /* tau_c in the real part, F in the imaginary part, so the layout
         * transpose used for fields brings both back to A[t][y][x] order */
        PlasmaK pk = c->pk;
        pk.hbar_w0 *= c->ascale * c->ascale;
        pk.cs *= c->ascale * c->ascale;
        rc = launch_ioniz<double>((const cx<double> *)c->field, c->nt, c->nxy, pk, (cx<double> *)out, c->stream);
    }
    if (!rc) rc = launch_to_reference<double>((const cx<double> *)out, c->staging, c->nx, c->ny, c->nt, 1.0, c->stream);
    std::vector<double> h;
    if (!rc) {
        h.resize(2 * (size_t)N);
        if (hipMemcpyAsync(h.data(), c->staging, 16 * (size_t)N, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = MPC_EHIP;
        if (!rc) rc = finish(c, 0);
    }
    if (!rc)
        for (long i = 0; i < N; i++) { tau_c[i] = h[2 * i]; freeelectrons[i] = h[2 * i + 1]; }
    if (out) (void)hipFree(out);
    mpc_destroy(c);
    set_err(rc);
}

extern "C" void ADK(double *ionizationrates1, double *ionizationrates2, int *dims, double *C_nl, double *f_lm,
                    double *nq, double *mq, double *Ei, double *E) {
    /* plasma.c:45-51 with |C_nl| taken as a double */
    double kon[6];
    kon[0] = std::pow(std::fabs(C_nl[0]), 2.0) * std::sqrt(6.0 / M_PI) * f_lm[0] * Ei[0];
    kon[1] = std::pow(std::fabs(C_nl[1]), 2.0) * std::sqrt(6.0 / M_PI) * f_lm[1] * Ei[1];
    kon[2] = 2.0 * std::pow(2.0 * Ei[0], 3.0 / 2.0);
    kon[3] = 2.0 * std::pow(2.0 * Ei[1], 3.0 / 2.0);
    kon[4] = 2.0 * nq[0] - mq[0] - 3.0 / 2.0;
    kon[5] = 2.0 * nq[1] - mq[1] - 3.0 / 2.0;
    const long N = (long)dims[0] * dims[1] * dims[2];
    double *dE = nullptr, *d1 = nullptr, *d2 = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&dE, 8 * (size_t)N) != hipSuccess || hipMalloc((void **)&d1, 8 * (size_t)N) != hipSuccess ||
        hipMalloc((void **)&d2, 8 * (size_t)N) != hipSuccess)
        rc = MPC_ENOMEM;
    if (!rc && hipMemcpy(dE, E, 8 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess) rc = MPC_EHIP;
    if (!rc) rc = launch_adk(dE, N, kon, d1, d2, nullptr);
    if (!rc && (hipMemcpy(ionizationrates1, d1, 8 * (size_t)N, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(ionizationrates2, d2, 8 * (size_t)N, hipMemcpyDeviceToHost) != hipSuccess))
        rc = MPC_EHIP;
    if (dE) (void)hipFree(dE);
    if (d1) (void)hipFree(d1);
    if (d2) (void)hipFree(d2);
    set_err(rc);
}
