/*
 * kernels.hip -- the gfx950 kernels of the split-step propagator.
 *
 * Sweeps over the resident field (layout A[y][x][t], t fastest):
 *   k_tline  : one workgroup owns whole t-lines.  [t-FFT] x P(omega) ->
 *              t-IFFT of A and of -i omega A (spectral d/dt) -> Kerr +
 *              self-steepening + plasma (ADK + avalanche, reverse-t affine
 *              scan in FP64) -> A exp(dz A_nl) -> t-FFT [x Q -> t-IFFT].
 *              Replaces propagation.c:164-327 and plasma.c:118-246 (which
 *              the reference runs on the host) and the t part of linearstep
 *              (propagation.c:11-30, 67-76).
 *   k_axis   : FFT along x or y of a tile of G consecutive t values,
 *              x D(k), inverse FFT [x mirror].  Replaces the x/y parts of
 *              linearstep (propagation.c:31-57, 77-100); the diffraction
 *              operator is separable, so x and y each take one sweep.
 * plus layout transposes, diagnostics reductions, the adaptive-step error
 * and Richardson sweeps, and the plasma/ADK kernels behind the legacy
 * ioniz()/ADK() entry points.
 */
#include "internal.h"
#include <cmath>
#include <cstdio>

namespace mpc {

/* ---------------------------------------------------------------------- */
/* FFT driver: unrolls Plan<N>::R into stages with LDS exchange between    */
/* ---------------------------------------------------------------------- */
template <class T, int N, int E, int DIR, int PAD, int Ns, int R, int... Rest>
struct DevFft {
    static __device__ __forceinline__ void run(cx<T> (&v)[E], int tid, cx<T> *lds, int ls,
                                               const cx<T> *tw) {
        constexpr bool last = sizeof...(Rest) == 0;
        fft_stage<T, N, E, R, Ns, DIR, last, PAD>(v, tid, lds, ls, tw);
        if constexpr (!last) {
            __syncthreads();
            fft_gather<T, N, E, PAD>(v, tid, lds, ls);
            __syncthreads();
            DevFft<T, N, E, DIR, PAD, Ns * R, Rest...>::run(v, tid, lds, ls, tw);
        }
    }
};
template <class T, int N, int E, int DIR, int PAD, class Rad> struct DevFftP;
template <class T, int N, int E, int DIR, int PAD, int... Rs> struct DevFftP<T, N, E, DIR, PAD, Radices<Rs...>> {
    static __device__ __forceinline__ void run(cx<T> (&v)[E], int tid, cx<T> *lds, int ls, const cx<T> *tw) {
        DevFft<T, N, E, DIR, PAD, 1, Rs...>::run(v, tid, lds, ls, tw);
    }
};
/* line FFT with plan PL (PL::E elements per thread, radix list PL::R) */
template <class T, int N, class PL, int DIR, int PAD = 0>
__device__ __forceinline__ void line_fft(cx<T> (&v)[PL::E], int tid, cx<T> *lds, int ls, const cx<T> *tw) {
    DevFftP<T, N, PL::E, DIR, PAD, typename PL::R>::run(v, tid, lds, ls, tw);
}

/* t-line geometry: plan, LDS padding and lines per 256-thread block.  The
 * LDS pad slot every E elements matches the stride of the first Stockham
 * stage's writes. */
template <int NT_, class PL_ = TPlan<NT_>> struct TGeo {
    static constexpr int NT = NT_;
    using PL = PL_;
    static constexpr int E = PL::E, TPL = NT / E;
    static constexpr int LPB = (256 / TPL) > 0 ? (256 / TPL) : 1;
    static constexpr int THREADS = LPB * TPL;
    static constexpr int PAD = E;
    static constexpr int FFT_ELEMS = NT + NT / PAD;
};

/* ---------------------------------------------------------------------- */
/* plasma model pieces (FP64 in both precisions)                           */
/* ---------------------------------------------------------------------- */
/* ADK rate Kon_a (Kon_b/Ea)^Kon_c exp(-Kon_b/(3 Ea)) (plasma.c:55-56) in log
 * form with e = Kon_b/(3 Ea) = ke * (1/s) and log(Kon_b/Ea) = L + log(1/s);
 * exactly 0 where the exponential underflows in double precision. */
__device__ __forceinline__ double adk_rate(double ka, double ke, double L, double kc, double invs) {
    const double e = ke * invs;
    if (ka == 0.0 || e > 760.0) return 0.0;
    const double w = ka * exp(kc * (L + log(invs)) - e);
    return (w != w) ? 0.0 : w;
}

/* 1/s, s = sqrt(2 ZW0 I / NC) the field scale from |A|^2, with I floored at
 * 1e-30 (plasma.c:135-138, :147) */
__device__ __forceinline__ double plasma_invs(double a2, const PlasmaK &k, double *Iout) {
    double I = a2 * k.hbar_w0;
    if (I == 0.0) I = 1e-30;
    *Iout = I;
    return rsqrt(k.zf * I);
}
/* collision time and inverse-bremsstrahlung cross-section from 1/s
 * (plasma.c:207-210, quiver velocity at w0) */
__device__ __forceinline__ double plasma_sigma(double invs, const PlasmaK &k, double *tau) {
    const double tc = k.kt * invs;
    *tau = tc;
    return k.K4 * tc / (1.0 + k.K5 * (tc * tc));
}
/* -expm1(-R): a 5-term series below 1e-4 (truncation < 1e-20 relative) */
__device__ __forceinline__ double one_minus_exp_neg(double R) {
    if (R < 1e-4) return R * (1.0 + R * (-0.5 + R * (1.0 / 6.0 + R * (-1.0 / 24.0 + R * (1.0 / 120.0)))));
    return -expm1(-R);
}

/* plasma arrays: one pad slot per E-element segment, so the per-thread
 * consecutive segments of the scan start in different LDS banks */
template <class G> __device__ __forceinline__ int padp(int i) { return i + i / G::E; }

/*
 * Reverse-t affine scan over one line.  Element k carries the map
 * f_k(R) = m_k R + c_k, given by coef(k, m_k, c_k); the result
 * R_k = f_k(R_{k+1}), R_NT = 0, is handed to out(k, R_k).  Thread tid owns the
 * consecutive segment [tid*E, tid*E + E): segment maps are composed serially,
 * combined across lanes with a shuffle-based suffix scan and across wavefronts
 * through LDS (scr), then each thread re-walks its segment.  coef must give
 * the same values on both walks; out may overwrite what coef read at k.
 */
template <class G, class Coef, class Out>
__device__ __forceinline__ void affine_suffix_scan(int tid, double *scr, Coef coef, Out out) {
    constexpr int E = G::E, TPL = G::NT / E;
    constexpr int W = TPL < 64 ? TPL : 64;
    constexpr int NW = (TPL + 63) / 64;
    const int lo = tid * E;
    double gm = 1.0, gc = 0.0;
#pragma unroll
    for (int i = E - 1; i >= 0; --i) {
        double mk, ck;
        coef(lo + i, mk, ck);
        gc = mk * gc + ck;
        gm = mk * gm;
    }
    const int lane = tid % W;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        const double om = __shfl_down(gm, d, W);
        const double oc = __shfl_down(gc, d, W);
        if (lane + d < W) {
            gc = gm * oc + gc;
            gm = gm * om;
        }
    }
    if constexpr (NW > 1) {
        const int wv = tid / 64;
        if (lane == 0) { scr[2 * wv] = gm; scr[2 * wv + 1] = gc; }
        __syncthreads();
        for (int w2 = wv + 1; w2 < NW; ++w2) {
            const double om = scr[2 * w2], oc = scr[2 * w2 + 1];
            gc = gm * oc + gc;
            gm = gm * om;
        }
    }
    scr[2 * NW + tid] = gc;
    __syncthreads();
    double R = (tid + 1 < TPL) ? scr[2 * NW + tid + 1] : 0.0;
#pragma unroll
    for (int i = E - 1; i >= 0; --i) {
        double mk, ck;
        coef(lo + i, mk, ck);
        R = mk * R + ck;
        out(lo + i, R);
    }
    __syncthreads();
}

template <class G> struct PlasmaLds {
    static constexpr int NT = G::NT, E = G::E, TPL = NT / E, NW = (TPL + 63) / 64;
    static constexpr int ARR = NT + NT / E;            /* padded doubles per array */
    static constexpr int SCR = 2 * NW + TPL;
    static constexpr int BYTES = (2 * ARR + SCR) * 8;
};

/*
 * Density of one line (plasma.c:218-241, the consistent two-species form):
 *   R_k     = (R_{k+1} + b_k)(1 - a_k) + a_k,                 R_NT = 0
 *   u_k     = 1 - exp(-R_k)      (evaluated as -expm1(-R_k))
 *   ions2_k = (1 - c_k) ions2_{k+1} + c_k u_{k+1},            ions2_NT = 0
 *   ions1_k = u_k - ions2_k ;  F_k = (ions1_k + 2 ions2_k) sigma_k
 * Element k = tid + TPL*m of the line is held by thread tid in slot m.  Two
 * padded LDS arrays: U holds m_k = 1 - a_k, then u_k; X holds the first
 * scan's c_k, then c_k of the second species, then ions2_k.  The second scan
 * forms its coefficients from U and X as it walks, so no per-element state
 * beyond 1/s stays in registers across the scans.
 * sink(m, F, tau) receives the result per element.
 */
template <class T, class G, class Sink>
__device__ __forceinline__ void plasma_line(const cx<T> (&A)[G::E], int tid, double *lds,
                                            const PlasmaK &pk, double &rho_max, Sink sink) {
    using L = PlasmaLds<G>;
    constexpr int NT = G::NT;
    constexpr int E = L::E, TPL = L::TPL;
    double *U = lds, *X = lds + L::ARR, *scr = lds + 2 * L::ARR;
    int anyc = 0;
    double invs[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = tid + TPL * m;
        const double a2 = (double)A[m].re * (double)A[m].re + (double)A[m].im * (double)A[m].im;
        double I;
        invs[m] = plasma_invs(a2, pk, &I);
        double tc;
        const double sig = plasma_sigma(invs[m], pk, &tc);
        const double b = adk_rate(pk.kon1, pk.ke1, pk.L1, pk.kon5, invs[m]) * pk.Kon;
        const double a = sig * I * pk.dt_I1e;             /* avalanche, SI units */
        U[padp<G>(k)] = 1.0 - a;
        X[padp<G>(k)] = b * (1.0 - a) + a;
        anyc |= (pk.kon2 != 0.0 && pk.ke2 * invs[m] <= 760.0);
    }
    __syncthreads();
    /* U <- u = 1 - exp(-R) */
    affine_suffix_scan<G>(tid, scr,
        [&](int k, double &mk, double &ck) { mk = U[padp<G>(k)]; ck = X[padp<G>(k)]; },
        [&](int k, double R) { U[padp<G>(k)] = one_minus_exp_neg(R); });
    const int block_anyc = __syncthreads_or(anyc);
    if (block_anyc) {
#pragma unroll
        for (int m = 0; m < E; ++m)
            X[padp<G>(tid + TPL * m)] = adk_rate(pk.kon2, pk.ke2, pk.L2, pk.kon6, invs[m]) * pk.Kon;
        __syncthreads();
        /* X <- ions2: m_k = 1 - c_k, c'_k = c_k u_{k+1} */
        affine_suffix_scan<G>(tid, scr,
            [&](int k, double &mk, double &ck) {
                const double c = X[padp<G>(k)];
                const double un = (k + 1 < NT) ? U[padp<G>(k + 1)] : 0.0;
                mk = 1.0 - c;
                ck = c * un;
            },
            [&](int k, double R) { X[padp<G>(k)] = R; });
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = tid + TPL * m;
        const double i2 = block_anyc ? X[padp<G>(k)] : 0.0;
        const double i1 = U[padp<G>(k)] - i2;
        const double fr = i1 + 2.0 * i2;
        double tc;
        const double sig = plasma_sigma(invs[m], pk, &tc);
        rho_max = fmax(rho_max, fr);
        sink(m, fr * sig, tc);
    }
    __syncthreads();
}

/* ---------------------------------------------------------------------- */
/* t-line kernel                                                           */
/* ---------------------------------------------------------------------- */
template <class T, class G, bool PLASMA> struct TCfg {
    static constexpr int E = G::E, TPL = G::TPL, LPB = G::LPB, THREADS = G::THREADS;
    static constexpr int FFT_BYTES = G::FFT_ELEMS * (int)sizeof(cx<T>);
    static constexpr int PL_BYTES = PLASMA ? PlasmaLds<G>::BYTES : 0;
    static constexpr int LINE_BYTES = ((FFT_BYTES > PL_BYTES ? FFT_BYTES : PL_BYTES) + 15) / 16 * 16;
};

/*
 * sin/cos of the per-step nonlinear phase (small in practice).  Cody-Waite
 * reduction by pi/2 with a three-part constant, then Taylor polynomials on
 * [-pi/4, pi/4] (truncation below 1e-19 in FP64, 2e-9 in FP32) and the
 * quadrant swap.  |x| > 1e5 takes the library routine instead.
 */
__device__ __noinline__ void sincos_big(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __noinline__ void sincos_bigf(float x, float *s, float *c) { sincosf(x, s, c); }

template <class T> __device__ __forceinline__ void sincos_t(T x, T *s, T *c);
template <> __device__ __forceinline__ void sincos_t<double>(double x, double *s, double *c) {
    if (!(fabs(x) <= 1e5)) { sincos_big(x, s, c); return; }
    const double k = rint(x * 0.63661977236758134308);          /* 2/pi */
    double r = fma(-k, 1.57079632673412561417e+00, x);            /* pi/2 = h + m + l */
    r = fma(-k, 6.07710050630396597660e-11, r);
    r = fma(-k, 2.02226624879595063154e-21, r);
    const double r2 = r * r;
    /* sin r = r - r^3/3! + ... ; cos r = 1 - r^2/2! + ... (Horner in r^2) */
    double sp = -1.0 / 121645100408832000.0;                      /* -1/19! */
    sp = fma(sp, r2, 1.0 / 355687428096000.0);                    /* 1/17! */
    sp = fma(sp, r2, -1.0 / 1307674368000.0);                     /* -1/15! */
    sp = fma(sp, r2, 1.0 / 6227020800.0);                         /* 1/13! */
    sp = fma(sp, r2, -1.0 / 39916800.0);                          /* -1/11! */
    sp = fma(sp, r2, 1.0 / 362880.0);                             /* 1/9! */
    sp = fma(sp, r2, -1.0 / 5040.0);                              /* -1/7! */
    sp = fma(sp, r2, 1.0 / 120.0);                                /* 1/5! */
    sp = fma(sp, r2, -1.0 / 6.0);                                 /* -1/3! */
    const double sr = fma(sp * r2, r, r);
    double cp = 1.0 / 6402373705728000.0;                         /* 1/18! */
    cp = fma(cp, r2, -1.0 / 20922789888000.0);                    /* -1/16! */
    cp = fma(cp, r2, 1.0 / 87178291200.0);                        /* 1/14! */
    cp = fma(cp, r2, -1.0 / 479001600.0);                         /* -1/12! */
    cp = fma(cp, r2, 1.0 / 3628800.0);                            /* 1/10! */
    cp = fma(cp, r2, -1.0 / 40320.0);                             /* -1/8! */
    cp = fma(cp, r2, 1.0 / 720.0);                                /* 1/6! */
    cp = fma(cp, r2, -1.0 / 24.0);                                /* -1/4! */
    cp = fma(cp, r2, 0.5);
    const double cr = fma(-cp, r2, 1.0);
    const int q = ((int)k) & 3;
    const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
template <> __device__ __forceinline__ void sincos_t<float>(float x, float *s, float *c) {
    if (!(fabsf(x) <= 1e5f)) { sincos_bigf(x, s, c); return; }
    const float k = rintf(x * 0.636619772f);
    float r = fmaf(-k, 1.5703125f, x);                            /* pi/2 in three parts */
    r = fmaf(-k, 4.8351287841796875e-4f, r);
    r = fmaf(-k, 3.13916473e-7f, r);
    const float r2 = r * r;
    float sp = 1.0f / 362880.0f;                                  /* Taylor to r^9 */
    sp = fmaf(sp, r2, -1.0f / 5040.0f);
    sp = fmaf(sp, r2, 1.0f / 120.0f);
    sp = fmaf(sp, r2, -1.0f / 6.0f);
    const float sr = fmaf(sp * r2, r, r);
    float cp = -1.0f / 3628800.0f;                                /* Taylor to r^10 */
    cp = fmaf(cp, r2, 1.0f / 40320.0f);
    cp = fmaf(cp, r2, -1.0f / 720.0f);
    cp = fmaf(cp, r2, 1.0f / 24.0f);
    cp = fmaf(cp, r2, -0.5f);
    const float cr = fmaf(cp, r2, 1.0f);
    const int q = ((int)k) & 3;
    const float ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
template <class T> __device__ __forceinline__ T exp_t(T x);
template <> __device__ __forceinline__ double exp_t<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float exp_t<float>(float x) { return expf(x); }

template <class T, class G, bool PLASMA>
__global__ void __launch_bounds__((TCfg<T, G, PLASMA>::THREADS)) k_tline(TArgs<T> a) {
    using C = TCfg<T, G, PLASMA>;
    using PL = typename G::PL;
    constexpr int NT = G::NT;
    constexpr int E = C::E, TPL = C::TPL;
    __shared__ __align__(16) unsigned char smem[C::LPB * C::LINE_BYTES];
    const int lslot = threadIdx.x / TPL, tid = threadIdx.x % TPL;
    const long line = (long)blockIdx.x * C::LPB + lslot;
    unsigned char *lraw = smem + lslot * C::LINE_BYTES;
    cx<T> *L = reinterpret_cast<cx<T> *>(lraw);
    cx<T> *g = a.field + line * NT;

    cx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = g[tid + TPL * m];
    if (!a.in_freq) line_fft<T, NT, PL, -1, G::PAD>(v, tid, L, 1, a.tw);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(v[m], a.P[tid + TPL * m]);

    if (a.do_nl) {
        cx<T> d[E];
        if (a.sst_on) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const T wk = a.wfft[tid + TPL * m];
                d[m] = mk(wk * v[m].im, -wk * v[m].re);   /* -i omega A(omega) */
            }
            line_fft<T, NT, PL, +1, G::PAD>(d, tid, L, 1, a.tw);   /* dA */
        }
        line_fft<T, NT, PL, +1, G::PAD>(v, tid, L, 1, a.tw);       /* A  */
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T a2 = v[m].re * v[m].re + v[m].im * v[m].im;
            if (a.sst_on) {
                /* S = gamma2/w0 (2 dA conj(A) + A conj(dA))  (propagation.c:232) */
                const cx<T> t1 = cmulc(d[m], v[m]);
                const cx<T> t2 = cmulc(v[m], d[m]);
                const T sre = a.sst_coef * ((T)2 * t1.re + t2.re);
                const T sim = a.sst_coef * ((T)2 * t1.im + t2.im);
                d[m] = mk(-sre, a.gamma * a2 - sim);
            } else {
                d[m] = mk((T)0, a.gamma * a2);
            }
        }
        if constexpr (PLASMA) {
            double rmax = 0.0;
            const double half_rho = 0.5 * a.pk.rho_nt;
            plasma_line<T, G>(v, tid, reinterpret_cast<double *>(lraw), a.pk, rmax,
                               [&](int m, double F, double tau) {
                                   /* (1 + i w0 tau_c) rho_nt F / 2  (propagation.c:289-292) */
                                   const double z = half_rho * F;
                                   d[m].re -= (T)z;
                                   d[m].im -= (T)(a.pk.w0 * tau * z);
                               });
            if (a.rho_block) {
                /* block max of the ionisation fraction */
                __shared__ double red[32];
                double r = rmax;
                for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_down(r, o));
                if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
                __syncthreads();
                if (threadIdx.x == 0) {
                    double mx = 0.0;
                    for (int i = 0; i < (C::THREADS + 63) / 64; ++i) mx = fmax(mx, red[i]);
                    a.rho_block[blockIdx.x] = mx;
                }
            }
        }
        /* A <- A exp(dz A_nl)  (propagation.c:317-327, complex product) */
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T er = exp_t<T>(a.dz * d[m].re);
            T sn, cs;
            sincos_t<T>(a.dz * d[m].im, &sn, &cs);
            const cx<T> z = mk(v[m].re * cs - v[m].im * sn, v[m].re * sn + v[m].im * cs);
            v[m] = mk(er * z.re, er * z.im);
        }
        line_fft<T, NT, PL, -1, G::PAD>(v, tid, L, 1, a.tw);
    }
    if (a.Q) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], a.Q[tid + TPL * m]);
    }
    if (a.out_time) line_fft<T, NT, PL, +1, G::PAD>(v, tid, L, 1, a.tw);
#pragma unroll
    for (int m = 0; m < E; ++m) g[tid + TPL * m] = v[m];
}

bool tline_supported(int nt) { return nt == 512 || nt == 1024 || nt == 2048 || nt == 4096; }
int tline_variants(void) { return 2; }

template <class T, class G, bool PL>
static int tline_go(const TArgs<T> &a, hipStream_t s) {
    using C = TCfg<T, G, PL>;
    if (a.nlines % C::LPB) return MPC_EINVAL;
    const long nb = a.nlines / C::LPB;
    hipLaunchKernelGGL((k_tline<T, G, PL>), dim3((unsigned)nb), dim3(C::THREADS), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
/* variant 0: 8 elements per thread (TPlan); 1: 16 per thread (Plan) */
template <class T, int NT> static int tline_v(const TArgs<T> &a, bool plasma, int variant, hipStream_t s) {
    if (variant == 1)
        return plasma ? tline_go<T, TGeo<NT, Plan<NT>>, true>(a, s) : tline_go<T, TGeo<NT, Plan<NT>>, false>(a, s);
    return plasma ? tline_go<T, TGeo<NT>, true>(a, s) : tline_go<T, TGeo<NT>, false>(a, s);
}
template <class T> int launch_tline(const TArgs<T> &a, int nt, bool plasma, int variant, hipStream_t s) {
    switch (nt) {
    case 512:  return tline_v<T, 512>(a, plasma, variant, s);
    case 1024: return tline_v<T, 1024>(a, plasma, variant, s);
    case 2048: return tline_v<T, 2048>(a, plasma, variant, s);
    case 4096: return tline_v<T, 4096>(a, plasma, variant, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_tline<double>(const TArgs<double> &, int, bool, int, hipStream_t);
template int launch_tline<float>(const TArgs<float> &, int, bool, int, hipStream_t);

/* lines per block are the same for every t-line variant: (256 / TPL) with
 * TPL = NT / E, and every variant's rho_block is sized by the largest grid */
/* t-line kernel block geometry: the smallest lines-per-block over the
 * variants, so per-block buffers sized by it fit every variant's grid */
int tline_lines_per_block(int nt, bool plasma, int prec) {
    (void)plasma; (void)prec;
    auto mn = [](int a, int b) { return a < b ? a : b; };
    switch (nt) {
    case 512: return mn(TGeo<512>::LPB, TGeo<512, Plan<512>>::LPB);
    case 1024: return mn(TGeo<1024>::LPB, TGeo<1024, Plan<1024>>::LPB);
    case 2048: return mn(TGeo<2048>::LPB, TGeo<2048, Plan<2048>>::LPB);
    case 4096: return mn(TGeo<4096>::LPB, TGeo<4096, Plan<4096>>::LPB);
    default: return 1;
    }
}

/* ---------------------------------------------------------------------- */
/* x / y line kernel                                                       */
/* ---------------------------------------------------------------------- */
template <class T> struct ATile { static constexpr int G = 256 / (int)sizeof(cx<T>); };  /* 256 B rows */
int axis_tile(int precision) { return precision == MPC_F32 ? ATile<float>::G : ATile<double>::G; }

/* 8-elements-per-thread plans for the transverse lines */
template <int N> struct APlan8 { using P = Plan<N>; static constexpr int E = P::E; using R = typename P::R; };
template <> struct APlan8<512> { static constexpr int E = 8; using R = Radices<8, 8, 8>; };
template <> struct APlan8<256> { static constexpr int E = 8; using R = Radices<8, 8, 4>; };
template <> struct APlan8<128> { static constexpr int E = 8; using R = Radices<8, 8, 2>; };
template <> struct APlan8<64>  { static constexpr int E = 8; using R = Radices<8, 8>; };
template <> struct APlan8<16>  { static constexpr int E = 8; using R = Radices<8, 2>; };

template <class T, int N, class PL, int G_> struct ACfg {
    static constexpr int E = PL::E, TPL = N / E, G = G_;
    static constexpr int THREADS = G * TPL;
};

template <class T, int N, class PL, int G_>
__global__ void __launch_bounds__((ACfg<T, N, PL, G_>::THREADS)) k_axis(AArgs<T> a) {
    using C = ACfg<T, N, PL, G_>;
    constexpr int E = C::E, TPL = C::TPL, G = C::G;
    __shared__ __align__(16) cx<T> lds[N * G];
    const int gcol = threadIdx.x % G, tid = threadIdx.x / G;
    const long outer = blockIdx.y;
    const long t0 = (long)blockIdx.x * G;
    cx<T> *base = a.field + outer * a.s_outer + t0 + gcol;
    cx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = base[(long)(tid + TPL * m) * a.s_line];
    line_fft<T, N, PL, -1>(v, tid, lds + gcol, G, a.tw);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(v[m], a.D[tid + TPL * m]);
    line_fft<T, N, PL, +1>(v, tid, lds + gcol, G, a.tw);
    if (a.M) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], a.M[(long)(tid + TPL * m) * a.m_line + outer * a.m_outer]);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) base[(long)(tid + TPL * m) * a.s_line] = v[m];
}

bool axis_supported(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 192 || n == 256 || n == 512; }
int axis_variants(void) { return 4; }

template <class T, int N, class PL, int G> static int axis_go(const AArgs<T> &a, int outer, hipStream_t s) {
    using C = ACfg<T, N, PL, G>;
    if (a.nt % G) return MPC_EINVAL;
    hipLaunchKernelGGL((k_axis<T, N, PL, G>), dim3((unsigned)(a.nt / G), (unsigned)outer), dim3(C::THREADS), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
/* variants: 0: Plan, 256-B rows; 1: Plan, 128-B rows; 2: 8 per thread,
 * 128-B rows; 3: 8 per thread, 256-B rows */
template <class T, int N> static int axis_v(const AArgs<T> &a, int outer, int variant, hipStream_t s) {
    constexpr int G = ATile<T>::G;
    switch (variant) {
    case 1: return axis_go<T, N, Plan<N>, G / 2>(a, outer, s);
    case 2: return axis_go<T, N, APlan8<N>, G / 2>(a, outer, s);
    case 3: return axis_go<T, N, APlan8<N>, G>(a, outer, s);
    default: return axis_go<T, N, Plan<N>, G>(a, outer, s);
    }
}
template <class T> int launch_axis(const AArgs<T> &a, int n, int outer, int variant, hipStream_t s) {
    switch (n) {
    case 16:  return axis_v<T, 16>(a, outer, variant, s);
    case 32:  return axis_v<T, 32>(a, outer, variant, s);
    case 64:  return axis_v<T, 64>(a, outer, variant, s);
    case 128: return axis_v<T, 128>(a, outer, variant, s);
    case 192: return axis_v<T, 192>(a, outer, variant, s);
    case 256: return axis_v<T, 256>(a, outer, variant, s);
    case 512: return axis_v<T, 512>(a, outer, variant, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_axis<double>(const AArgs<double> &, int, int, int, hipStream_t);
template int launch_axis<float>(const AArgs<float> &, int, int, int, hipStream_t);

/* ---------------------------------------------------------------------- */
/* layout conversion: reference A[t][y][x] (FP64) <-> internal A[y][x][t]  */
/* ---------------------------------------------------------------------- */
template <class T>
__global__ void __launch_bounds__(256) k_to_internal(const double *__restrict__ src, cx<T> *__restrict__ dst,
                                                     int nx, int ny, int nt, double scale) {
    __shared__ double2 tile[32][33];
    const int x0 = blockIdx.x * 32, t0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   /* 32 x 8 */
    for (int i = ty; i < 32; i += 8) {
        if (x0 + tx < nx) {
            const long s = ((long)(t0 + i) * ny + y) * nx + x0 + tx;
            tile[i][tx] = reinterpret_cast<const double2 *>(src)[s];
        }
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        if (x0 + i < nx) {
            const double2 e = tile[tx][i];                       /* t = t0+tx, x = x0+i */
            const long d = ((long)y * nx + x0 + i) * nt + t0 + tx;
            dst[d] = mk((T)(e.x * scale), (T)(e.y * scale));
        }
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_to_reference(const cx<T> *__restrict__ src, double *__restrict__ dst,
                                                      int nx, int ny, int nt, double scale) {
    __shared__ double2 tile[32][33];
    const int x0 = blockIdx.x * 32, t0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        if (x0 + i < nx) {
            const cx<T> e = src[((long)y * nx + x0 + i) * nt + t0 + tx];   /* x = x0+i, t = t0+tx */
            tile[tx][i] = make_double2((double)e.re * scale, (double)e.im * scale);
        }
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        if (x0 + tx < nx) {
            const long d = ((long)(t0 + i) * ny + y) * nx + x0 + tx;
            reinterpret_cast<double2 *>(dst)[d] = tile[i][tx];
        }
    }
}
template <class T> int launch_to_internal(const double *src, cx<T> *dst, int nx, int ny, int nt, double scale, hipStream_t s) {
    if (nt % 32) return MPC_EINVAL;
    hipLaunchKernelGGL((k_to_internal<T>), dim3((nx + 31) / 32, nt / 32, ny), dim3(256), 0, s, src, dst, nx, ny, nt, scale);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T> int launch_to_reference(const cx<T> *src, double *dst, int nx, int ny, int nt, double scale, hipStream_t s) {
    if (nt % 32) return MPC_EINVAL;
    hipLaunchKernelGGL((k_to_reference<T>), dim3((nx + 31) / 32, nt / 32, ny), dim3(256), 0, s, src, dst, nx, ny, nt, scale);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_to_internal<double>(const double *, cx<double> *, int, int, int, double, hipStream_t);
template int launch_to_internal<float>(const double *, cx<float> *, int, int, int, double, hipStream_t);
template int launch_to_reference<double>(const cx<double> *, double *, int, int, int, double, hipStream_t);
template int launch_to_reference<float>(const cx<float> *, double *, int, int, int, double, hipStream_t);

/* per-block max of |re|, |im| over n doubles (host reduces the 1024 partials) */
__global__ void __launch_bounds__(256) k_maxabs(const double *v, long n, double *part) {
    __shared__ double red[256];
    double m = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) m = fmax(m, fabs(v[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
int launch_maxabs(const double *v, long n, double *part, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(256), 0, s, v, n, part);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

template <class T> __global__ void k_scale(cx<T> *f, long n, T s) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        f[i] = mk(f[i].re * s, f[i].im * s);
}
template <class T> int launch_scale(cx<T> *f, long n, T s, hipStream_t st) {
    hipLaunchKernelGGL((k_scale<T>), dim3(2048), dim3(256), 0, st, f, n, s);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_scale<double>(cx<double> *, long, double, hipStream_t);
template int launch_scale<float>(cx<float> *, long, float, hipStream_t);

/* ---------------------------------------------------------------------- */
/* diagnostics: per-line power, and per-block partial temp[t] / spec[t]    */
/* ---------------------------------------------------------------------- */
template <class T, int NT>
__global__ void __launch_bounds__(256) k_diag_lines(const cx<T> *f, const cx<T> *tw, long nlines,
                                                    double *line_power, double *temp_part, double *spec_part) {
    using G = TGeo<NT>;
    constexpr int E = G::E, TPL = G::TPL, LPB = G::LPB;
    __shared__ __align__(16) cx<T> lds[LPB * G::FFT_ELEMS];
    __shared__ double accT[NT], accS[NT];
    __shared__ double lp[LPB][TPL / 64 + 1];
    const int lslot = threadIdx.x / TPL, tid = threadIdx.x % TPL;
    for (int i = threadIdx.x; i < NT; i += blockDim.x) { accT[i] = 0.0; accS[i] = 0.0; }
    double at[E], as[E];
#pragma unroll
    for (int m = 0; m < E; ++m) { at[m] = 0.0; as[m] = 0.0; }
    /* a fixed set of line groups per block: deterministic order */
    const long ngroups = nlines / LPB;
    for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long line = grp * LPB + lslot;
        const cx<T> *g = f + line * NT;
        cx<T> v[E];
        double pw = 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            v[m] = g[tid + TPL * m];
            const double p = (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
            at[m] += p;
            pw += p;
        }
        /* line power: reduce over the TPL threads of the line */
        constexpr int W = TPL < 64 ? TPL : 64;
        for (int o = W / 2; o > 0; o >>= 1) pw += __shfl_down(pw, o, W);
        if (tid % 64 == 0) lp[lslot][tid / 64] = pw;
        line_fft<T, NT, TPlan<NT>, -1, G::PAD>(v, tid, lds + lslot * G::FFT_ELEMS, 1, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) as[m] += (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int i = 0; i < (TPL + 63) / 64; ++i) sum += lp[lslot][i];
            line_power[line] = sum;
        }
        __syncthreads();
    }
    /* combine the LPB line slots in slot order */
    for (int sl = 0; sl < LPB; ++sl) {
        if (lslot == sl) {
#pragma unroll
            for (int m = 0; m < E; ++m) { accT[tid + TPL * m] += at[m]; accS[tid + TPL * m] += as[m]; }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < NT; i += blockDim.x) {
        temp_part[(long)blockIdx.x * NT + i] = accT[i];
        spec_part[(long)blockIdx.x * NT + i] = accS[i];
    }
}
template <class T, int NT>
static int diag_go(const cx<T> *f, const cx<T> *tw, long nlines, double *lp, double *tp, double *sp, int nb, hipStream_t s) {
    using G = TGeo<NT>;
    if (nlines % G::LPB) return MPC_EINVAL;
    hipLaunchKernelGGL((k_diag_lines<T, NT>), dim3(nb), dim3(G::THREADS), 0, s, f, tw, nlines, lp, tp, sp);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T> int launch_diag_lines(const cx<T> *f, const cx<T> *tw, int nt, long nlines, double *lp,
                                         double *tp, double *sp, int nb, hipStream_t s) {
    switch (nt) {
    case 512: return diag_go<T, 512>(f, tw, nlines, lp, tp, sp, nb, s);
    case 1024: return diag_go<T, 1024>(f, tw, nlines, lp, tp, sp, nb, s);
    case 2048: return diag_go<T, 2048>(f, tw, nlines, lp, tp, sp, nb, s);
    case 4096: return diag_go<T, 4096>(f, tw, nlines, lp, tp, sp, nb, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_diag_lines<double>(const cx<double> *, const cx<double> *, int, long, double *, double *, double *, int, hipStream_t);
template int launch_diag_lines<float>(const cx<float> *, const cx<float> *, int, long, double *, double *, double *, int, hipStream_t);

/* out[c] = sum_r part[r][c], in row order (deterministic) */
__global__ void k_reduce_cols(const double *part, int nrows, int ncols, double *out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double s = 0.0;
    for (int r = 0; r < nrows; ++r) s += part[(long)r * ncols + c];
    out[c] = s;
}
int launch_reduce_cols(const double *part, int nrows, int ncols, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_cols, dim3((ncols + 255) / 256), dim3(256), 0, s, part, nrows, ncols, out);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

/* ---------------------------------------------------------------------- */
/* adaptive step: error partials and Richardson (propagation.c:625-631,    */
/* MPC.c:377-381)                                                          */
/* ---------------------------------------------------------------------- */
template <class T>
__global__ void __launch_bounds__(256) k_err_partials(const cx<T> *c1, const cx<T> *c2, long n, double *part) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double dr = (double)c2[i].re - (double)c1[i].re, di = (double)c2[i].im - (double)c1[i].im;
        a += dr * dr + di * di;
        b += (double)c2[i].re * c2[i].re + (double)c2[i].im * c2[i].im;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { s1[threadIdx.x] += s1[threadIdx.x + o]; s2[threadIdx.x] += s2[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1[0]; part[2 * blockIdx.x + 1] = s2[0]; }
}
template <class T> int launch_err_partials(const cx<T> *c1, const cx<T> *c2, long n, double *part, int nb, hipStream_t s) {
    hipLaunchKernelGGL((k_err_partials<T>), dim3(nb), dim3(256), 0, s, c1, c2, n, part);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_err_partials<double>(const cx<double> *, const cx<double> *, long, double *, int, hipStream_t);
template int launch_err_partials<float>(const cx<float> *, const cx<float> *, long, double *, int, hipStream_t);

template <class T> __global__ void k_richardson(cx<T> *a, const cx<T> *c1, const cx<T> *c2, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double r = (4.0 / 3.0) * (double)c2[i].re - (1.0 / 3.0) * (double)c1[i].re;
        const double m = (4.0 / 3.0) * (double)c2[i].im - (1.0 / 3.0) * (double)c1[i].im;
        a[i] = mk((T)r, (T)m);
    }
}
template <class T> int launch_richardson(cx<T> *a, const cx<T> *c1, const cx<T> *c2, long n, hipStream_t s) {
    hipLaunchKernelGGL((k_richardson<T>), dim3(2048), dim3(256), 0, s, a, c1, c2, n);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_richardson<double>(cx<double> *, const cx<double> *, const cx<double> *, long, hipStream_t);
template int launch_richardson<float>(cx<float> *, const cx<float> *, const cx<float> *, long, hipStream_t);

/* ---------------------------------------------------------------------- */
/* legacy ioniz(): tau_c and F per element (internal layout outputs)       */
/* ---------------------------------------------------------------------- */
template <class T, int NT>
__global__ void __launch_bounds__((TCfg<T, TGeo<NT>, true>::THREADS)) k_ioniz(const cx<T> *f, long nlines, PlasmaK pk,
                                                                     cx<double> *out) {
    using G = TGeo<NT>;
    using C = TCfg<T, G, true>;
    constexpr int E = C::E, TPL = C::TPL;
    __shared__ __align__(16) unsigned char smem[C::LPB * C::LINE_BYTES];
    const int lslot = threadIdx.x / TPL, tid = threadIdx.x % TPL;
    const long line = (long)blockIdx.x * C::LPB + lslot;
    cx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = f[line * NT + tid + TPL * m];
    double rmax = 0.0;
    plasma_line<T, G>(v, tid, reinterpret_cast<double *>(smem + lslot * C::LINE_BYTES), pk, rmax,
                       [&](int m, double Fv, double tv) {
                           out[line * NT + tid + TPL * m] = mk(tv, Fv);   /* (tau_c, F) */
                       });
}
template <class T, int NT> static int ioniz_go(const cx<T> *f, long nlines, const PlasmaK &pk, cx<double> *out, hipStream_t s) {
    using C = TCfg<T, TGeo<NT>, true>;
    if (nlines % C::LPB) return MPC_EINVAL;
    hipLaunchKernelGGL((k_ioniz<T, NT>), dim3((unsigned)(nlines / C::LPB)), dim3(C::THREADS), 0, s, f, nlines, pk, out);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T> int launch_ioniz(const cx<T> *f, int nt, long nlines, const PlasmaK &pk, cx<double> *out, hipStream_t s) {
    switch (nt) {
    case 512: return ioniz_go<T, 512>(f, nlines, pk, out, s);
    case 1024: return ioniz_go<T, 1024>(f, nlines, pk, out, s);
    case 2048: return ioniz_go<T, 2048>(f, nlines, pk, out, s);
    case 4096: return ioniz_go<T, 4096>(f, nlines, pk, out, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_ioniz<double>(const cx<double> *, int, long, const PlasmaK &, cx<double> *, hipStream_t);
template int launch_ioniz<float>(const cx<float> *, int, long, const PlasmaK &, cx<double> *, hipStream_t);

/* legacy ADK(): rates from field amplitudes in atomic units (plasma.c:43-58) */
__global__ void k_adk(const double *E, long n, double k0, double k1, double k2, double k3, double k4,
                      double k5, double *W1, double *W2) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        /* no NaN filtering here: the reference does that in ioniz (plasma.c:184-192) */
        W1[i] = k0 * pow(k2 / E[i], k4) * exp(-k2 / (3.0 * E[i]));
        W2[i] = k1 * pow(k3 / E[i], k5) * exp(-k3 / (3.0 * E[i]));
    }
}
int launch_adk(const double *E, long n, const double *kon, double *W1, double *W2, hipStream_t s) {
    hipLaunchKernelGGL(k_adk, dim3(1024), dim3(256), 0, s, E, n, kon[0], kon[1], kon[2], kon[3], kon[4], kon[5], W1, W2);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

} // namespace mpc
