/*
 * device.h -- device building blocks shared by the kernel translation units:
 * the line-FFT driver over fft_core.h's stages, t-line geometry, the FP64
 * plasma model (ADK rate, avalanche, reverse-t affine scan), and the
 * nonlinear phase's sin/cos and exp.
 */
#ifndef MPC_DEVICE_H
#define MPC_DEVICE_H

#include "internal.h"
#include <cmath>
#include <cstdio>
#include <type_traits>

namespace mpc {

/*
 * Loads and stores through global-address-space pointers.  A kernel that
 * reads its pointers from device memory (the batched sweeps take them from an
 * argument array) cannot tell global from LDS or private addresses and would
 * access through flat instructions.  Flat accesses also count against the LDS
 * wait counter, so every LDS wait of a transform would wait for the prefetch
 * of the next line as well.  All pointers given here address global memory.
 */
template <class T> __device__ __forceinline__ T gload(const T *p, long i) {
    return ((const __attribute__((address_space(1))) T *)p)[i];
}
template <class T> __device__ __forceinline__ cx<T> gload(const cx<T> *p, long i) {
    typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    const v2 r = ((const __attribute__((address_space(1))) v2 *)p)[i];
    return mk(r.x, r.y);
}
template <class T> __device__ __forceinline__ void gstore(cx<T> *p, long i, cx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    v2 r;
    r.x = v.re;
    r.y = v.im;
    ((__attribute__((address_space(1))) v2 *)p)[i] = r;
}

/* ---------------------------------------------------------------------- */
/* FFT driver: unrolls Plan<N>::R into stages with LDS exchange between    */
/* ---------------------------------------------------------------------- */
/*
 * BOFF = 0: one exchange buffer, two barriers per exchange.  BOFF != 0: the
 * stages alternate between lds and lds + BOFF*ls, so a stage's writes can
 * never meet a slower thread's reads of the previous exchange and one
 * barrier per exchange suffices; the caller then puts a barrier before the
 * transform (the buffer it writes first may still be read by the previous
 * user of the LDS).  The sign of BOFF flips at every stage.
 */
template <class T, int N, int E, int DIR, int PAD, int BOFF, int Ns, int R, int... Rest>
struct DevFft {
    static __device__ __forceinline__ void run(cx<T> (&v)[E], int tid, cx<T> *lds, int ls,
                                               const cx<T> *tw) {
        constexpr bool last = sizeof...(Rest) == 0;
        fft_stage<T, N, E, R, Ns, DIR, last, PAD>(v, tid, lds, ls, tw);
        if constexpr (!last) {
            __syncthreads();
            fft_gather<T, N, E, PAD>(v, tid, lds, ls);
            if constexpr (BOFF == 0) {
                __syncthreads();
                DevFft<T, N, E, DIR, PAD, 0, Ns * R, Rest...>::run(v, tid, lds, ls, tw);
            } else {
                DevFft<T, N, E, DIR, PAD, -BOFF, Ns * R, Rest...>::run(v, tid, lds + BOFF * ls, ls, tw);
            }
        }
    }
};
template <class T, int N, int E, int DIR, int PAD, int BOFF, class Rad> struct DevFftP;
template <class T, int N, int E, int DIR, int PAD, int BOFF, int... Rs>
struct DevFftP<T, N, E, DIR, PAD, BOFF, Radices<Rs...>> {
    static __device__ __forceinline__ void run(cx<T> (&v)[E], int tid, cx<T> *lds, int ls, const cx<T> *tw) {
        if constexpr (BOFF != 0) __syncthreads();
        DevFft<T, N, E, DIR, PAD, BOFF, 1, Rs...>::run(v, tid, lds, ls, tw);
    }
};
/* line FFT with plan PL (PL::E elements per thread, radix list PL::R) */
template <class T, int N, class PL, int DIR, int PAD = 0, int BOFF = 0>
__device__ __forceinline__ void line_fft(cx<T> (&v)[PL::E], int tid, cx<T> *lds, int ls, const cx<T> *tw) {
    DevFftP<T, N, PL::E, DIR, PAD, BOFF, typename PL::R>::run(v, tid, lds, ls, tw);
}

/* t-line geometry: plan, LDS layout (XOR swizzle keyed to E) and lines per
 * 256-thread block */
template <int NT_, class PL_ = TPlan<NT_>> struct TGeo {
    static constexpr int NT = NT_;
    using PL = PL_;
    static constexpr int E = PL::E, TPL = NT / E;
    static constexpr int LPB = (256 / TPL) > 0 ? (256 / TPL) : 1;
    static constexpr int THREADS = LPB * TPL;
    static constexpr int PAD = -ilog2(E);              /* XOR swizzle, see padi */
    static constexpr int FFT_ELEMS = padded_elems<PAD>(NT);
};

/* ---------------------------------------------------------------------- */
/* plasma model pieces (FP64 in both precisions)                           */
/* ---------------------------------------------------------------------- */
/*
 * ADK rate Kon_a (Kon_b/Ea)^Kon_c exp(-Kon_b/(3 Ea)) (plasma.c:55-56) in log
 * form with e = Kon_b/(3 Ea) = ke * (1/s) and log(Kon_b/Ea) = L + log(1/s).
 * Rates with e > ADK_E_MAX are taken as 0: there Kon_b/Ea = 3e > 300, and
 * for Kon_c < 4 the rate is below Kon_a * 300^4 * exp(-100) < 1e-33 Kon_a,
 * so per step it adds less than 1e-30 to an ionised fraction.  That skips
 * the exp/log in the field's wings, where it would only produce such values.
 */
constexpr double ADK_E_MAX = 100.0;
__device__ __forceinline__ double adk_rate(double ka, double ke, double L, double kc, double invs) {
    const double e = ke * invs;
    if (ka == 0.0 || e > ADK_E_MAX) return 0.0;
    const double w = ka * exp(kc * (L + log(invs)) - e);
    return (w != w) ? 0.0 : w;
}
/* 1/d for normal d > 0: hardware estimate and two Newton steps (within an
 * ulp or so of the correctly rounded quotient) */
__device__ __forceinline__ double fast_rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
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
    /* 1 + K5 tc^2 is in [1, 1e60] for any I >= the 1e-30 floor */
    return k.K4 * tc * fast_rcp(1.0 + k.K5 * (tc * tc));
}
/* 1 - exp(-R) = sum_{k>=1} (-1)^(k+1) R^k / k!: four terms when a whole
 * wave has R <= 2^-12, twelve up to R = 0.1 (truncation below 2e-22
 * relative), -expm1(-R) beyond */
__device__ __forceinline__ double one_minus_exp_neg(double R) {
    if (__all(R <= 0x1p-12)) {
        /* a whole wave of small densities: four terms (truncation below
         * R^4/120 < 3e-17 relative) */
        double p = -1.0 / 24.0;
        p = fma(p, R, 1.0 / 6.0);
        p = fma(p, R, -0.5);
        p = fma(p, R, 1.0);
        return p * R;
    }
    if (R <= 0.1) {
        double p = -1.0 / 479001600.0;                  /* -1/12! */
        p = fma(p, R, 1.0 / 39916800.0);
        p = fma(p, R, -1.0 / 3628800.0);
        p = fma(p, R, 1.0 / 362880.0);
        p = fma(p, R, -1.0 / 40320.0);
        p = fma(p, R, 1.0 / 5040.0);
        p = fma(p, R, -1.0 / 720.0);
        p = fma(p, R, 1.0 / 120.0);
        p = fma(p, R, -1.0 / 24.0);
        p = fma(p, R, 1.0 / 6.0);
        p = fma(p, R, -0.5);
        p = fma(p, R, 1.0);
        return p * R;
    }
    return -expm1(-R);
}

/* plasma arrays: the exchange buffers' swizzle, so both the per-element
 * accesses (unit lane stride) and the scan's per-thread segments (lane
 * stride E) are free of bank conflicts */
template <class G> __device__ __forceinline__ int padp(int i) { return padi<G::PAD>(i); }
/* the swizzled position of element tid + TPL*m (constant offsets where the
 * layout allows, see padi_strided) */
template <class G> __device__ __forceinline__ int padp_tm(int tid, int m) {
    return padi_strided<G::PAD, G::NT / G::E>(tid, m);
}

/*
 * Lane moves within a wavefront by DPP (no LDS round trip): CTRL is
 * row_shr:d (0x110 + d, lanes of a 16-lane row take lane - d of the same row),
 * row_bcast:15 (0x142, the next row takes lane 15 of a row) or row_bcast:31
 * (0x143, rows 2 and 3 take lane 31); ROWS masks the rows written.  Lanes with
 * no source, or in rows not written, return old.
 */
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_move(double x, double old) {
    const int2 v = __builtin_bit_cast(int2, x), o = __builtin_bit_cast(int2, old);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(o.x, v.x, CTRL, ROWS, 0xf, false);
    r.y = __builtin_amdgcn_update_dpp(o.y, v.y, CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, r);
}
/* sum of s over the W lanes of a group: W = 64, a whole wavefront by DPP
 * (inclusive prefix sums, the total in lane 63); W < 64, butterfly shuffles
 * (the total in every lane of the group) */
template <int W> __device__ __forceinline__ void lanes_sum(double &s) {
    if constexpr (W == 64) {
        s += dpp_move<0x111, 0xf>(s, 0.0);
        s += dpp_move<0x112, 0xf>(s, 0.0);
        s += dpp_move<0x114, 0xf>(s, 0.0);
        s += dpp_move<0x118, 0xf>(s, 0.0);
        s += dpp_move<0x142, 0xa>(s, 0.0);
        s += dpp_move<0x143, 0xc>(s, 0.0);
    } else {
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, W);
    }
}

/* Knuth's TwoSum: s + e == a + b exactly, s = fl(a + b) */
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
/*
 * Composes an affine segment map (gm, gc + gl) with the map of the segment
 * that follows it in the scan, (om, oc + ol): R -> gm (om R + oc + ol) + gc +
 * gl.  The constant term is kept as an unevaluated sum hi + lo (a
 * compensated carry): the product's rounding error comes from an FMA, the
 * sum's from TwoSum, and both go into lo.
 */
__device__ __forceinline__ void affine_combine(double &gm, double &gc, double &gl, double om, double oc, double ol) {
    const double p = gm * oc;
    const double ep = fma(gm, oc, -p);
    double sh, es;
    two_sum(p, gc, sh, es);
    gl = gl + fma(gm, ol, ep + es);
    gc = sh;
    gm = gm * om;
}

/*
 * Reverse-t affine scan over one line.  Element k carries the map
 * f_k(R) = m_k R + c_k, given by coef(k, m_k, c_k); the result
 * R_k = f_k(R_{k+1}), R_NT = 0, is handed to out(k, R_k).  Thread tid owns the
 * consecutive segment [tid*E, tid*E + E): segment maps are composed serially,
 * combined across lanes with a shuffle-based suffix scan and across wavefronts
 * through LDS (scr, 3 NW + TPL doubles), then each thread re-walks its
 * segment from the suffix value of the segments after it.  The carries of
 * the lane and wavefront combines are compensated (affine_combine); the
 * suffix value is rounded once, where a thread's walk starts.  coef must give
 * the same values on both walks; out may overwrite what coef read at k.
 */
template <class G, bool GATED = false, class Coef, class Out, class Gate = int>
__device__ __forceinline__ bool affine_suffix_scan(int tid, double *scr, Coef coef, Out out,
                                                   double gate_init = 0.0, Gate gate = 0) {
    constexpr int E = G::E, TPL = G::NT / E;
    constexpr int W = TPL < 64 ? TPL : 64;
    constexpr int NW = (TPL + 63) / 64;
    const int lo = tid * E;
    double gm = 1.0, gc = 0.0, gl = 0.0, cs = gate_init;
#pragma unroll
    for (int i = E - 1; i >= 0; --i) {
        double mk, ck;
        coef(lo + i, mk, ck);
        gc = mk * gc + ck;
        gm = mk * gm;
        if constexpr (GATED) cs += ck;
    }
    if constexpr (GATED) {
        /* the line's sum of c_k (plus gate_init of every thread) goes to
         * gate(); when it returns true on every line of the workgroup the
         * scan stops here and returns false (out is not called) */
        double *part = scr + 3 * NW + TPL;                 /* NW doubles */
        lanes_sum<W>(cs);
        if ((tid % W) == (W == 64 ? 63 : 0)) part[tid / W] = cs;
        __syncthreads();
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += part[w];
        bool stop = gate(sum);
        if constexpr (G::LPB > 1) {
            stop = !__syncthreads_or(!stop);                /* whole workgroup or none */
        } else {
            stop = __builtin_amdgcn_readfirstlane((int)stop);   /* the same on every lane */
        }
        if (stop) return false;
    }
    const int lane = tid % W;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        const double om = __shfl_down(gm, d, W);
        const double oc = __shfl_down(gc, d, W);
        const double ol = __shfl_down(gl, d, W);
        if (lane + d < W) affine_combine(gm, gc, gl, om, oc, ol);
    }
    if constexpr (NW > 1) {
        const int wv = tid / 64;
        if (lane == 0) { scr[3 * wv] = gm; scr[3 * wv + 1] = gc; scr[3 * wv + 2] = gl; }
        __syncthreads();
        for (int w2 = wv + 1; w2 < NW; ++w2) affine_combine(gm, gc, gl, scr[3 * w2], scr[3 * w2 + 1], scr[3 * w2 + 2]);
    }
    scr[3 * NW + tid] = gc + gl;
    __syncthreads();
    double R = (tid + 1 < TPL) ? scr[3 * NW + tid + 1] : 0.0;
#pragma unroll
    for (int i = E - 1; i >= 0; --i) {
        double mk, ck;
        coef(lo + i, mk, ck);
        R = mk * R + ck;
        out(lo + i, R);
    }
    __syncthreads();
    return true;
}

/* per-line plasma LDS: U and X (FP64) and the scan's scratch, plus S (sigma)
 * and V (1/s) in the field precision T while the workgroup's share stays
 * within 80 KB; otherwise 1/s stays in registers and sigma is formed again
 * for the result */
template <class G, class T> struct PlasmaLds {
    static constexpr int NT = G::NT, E = G::E, TPL = NT / E, NW = (TPL + 63) / 64;
    static constexpr int ARR = padded_elems<G::PAD>(NT);   /* elements per array */
    static constexpr int SCR = 3 * NW + TPL + NW;           /* doubles: the scan's scratch, then the
                                                               cold-line test's per-wave partials */
    static constexpr int SVB = 2 * ARR * (int)sizeof(T);
    static constexpr bool SV = G::LPB * ((2 * ARR + SCR) * 8 + SVB) <= 80 * 1024;
    static constexpr int BYTES = (2 * ARR + SCR) * 8 + (SV ? SVB : 0);
};

/*
 * Cold lines.  The first scan below is R_k = R_{k+1} (1 - a_k) + X_k with
 * X_k = b_k (1 - a_k) + a_k >= a_k >= 0, so with S = sum_k X_k <= 1 over a
 * line, R_{k+1} <= R_k <= R_0 <= S and R_0 >= S (1 - S).  Without the second
 * species (c = 0) the ionised fraction is u_k = 1 - exp(-R_k) <= S, and the
 * line's maximum u_0 >= S (1 - 3S/2).  The plasma part of the exponent,
 * dz rho_nt/2 u sigma (1 + i w0 tau_c), is then at most
 * |dz| rho_nt/2 S sigma_bound in modulus at every element, with
 * sigma_bound >= sigma (1 + w0 tau) for every tau.  A line where that is
 * below EPS (the field changes by less than that, relative: under the
 * rounding of T) and S <= S_MAX skips the rest of the scans: its plasma part
 * is taken as zero and its ionised-fraction maximum as S (relative error
 * below 1.5 S_MAX).
 */
template <class T> struct ColdLine;
template <> struct ColdLine<double> { static constexpr double S_MAX = 0x1p-42, EPS = 0x1p-53; };
template <> struct ColdLine<float> { static constexpr double S_MAX = 0x1p-24, EPS = 0x1p-26; };

/*
 * The per-element plasma quantities from |A|^2 (a2, the field scale folded
 * into pk):  1/s (inv_s), the collision time and cross-section (tau, sigma),
 * the avalanche coefficient a = sigma I dt / (I1 e) and the field-ionisation
 * coefficient b = W1(1/s) dt/T0 of the first species.
 *
 * FP64 path: everything in FP64.  FP32 path: 1/s, tau, sigma and a in FP32;
 * the ADK exponent and log stay FP64 (an exponent near 100 amplifies any
 * error of 1/s a hundredfold), formed from |A|^2 in FP64 and 1/s refined by
 * one FP64 Newton step from the FP32 estimate, and only on lanes whose FP32
 * exponent says the rate can be non-zero.
 */
template <class T> struct PlasmaElem;
template <> struct PlasmaElem<double> {
    double inv_s, I;
    __device__ __forceinline__ void set(double re, double im, const PlasmaK &k) {
        inv_s = plasma_invs(re * re + im * im, k, &I);
    }
    __device__ __forceinline__ double sigma(const PlasmaK &k, double *tau) const { return plasma_sigma(inv_s, k, tau); }
    __device__ __forceinline__ double avalanche(const PlasmaK &k, double sig) const { return sig * I * k.dt_I1e; }
    /* ADK rate of species 1 or 2 times Kon */
    __device__ __forceinline__ double adk_b(const PlasmaK &k, int species) const {
        return species == 1 ? adk_rate(k.kon1, k.ke1, k.L1, k.kon5, inv_s) * k.Kon
                            : adk_rate(k.kon2, k.ke2, k.L2, k.kon6, inv_s) * k.Kon;
    }
    __device__ __forceinline__ bool adk_live(const PlasmaK &k, int species) const {
        return species == 1 ? (k.kon1 != 0.0 && k.ke1 * inv_s <= ADK_E_MAX)
                            : (k.kon2 != 0.0 && k.ke2 * inv_s <= ADK_E_MAX);
    }
};
template <> struct PlasmaElem<float> {
    float inv_s, I;
    double a2;                         /* |A|^2 in FP64 (exact from the FP32 components) */
    __device__ __forceinline__ void set(float re, float im, const PlasmaK &k) {
        a2 = (double)re * (double)re + (double)im * (double)im;
        I = (float)a2 * (float)k.hbar_w0;
        if (I == 0.0f) I = 1e-30f;                           /* plasma.c:135-138 */
        inv_s = __builtin_amdgcn_rsqf((float)k.zf * I);      /* 1 ulp */
    }
    __device__ __forceinline__ float sigma(const PlasmaK &k, float *tau) const {
        const float tc = (float)k.kt * inv_s;
        *tau = tc;
        /* 1 + K5 tc^2 may overflow to inf at the 1e-30 floor: sigma -> 0 there */
        return (float)k.K4 * tc * __builtin_amdgcn_rcpf(1.0f + (float)k.K5 * (tc * tc));
    }
    __device__ __forceinline__ float avalanche(const PlasmaK &k, float sig) const {
        return sig * I * (float)k.dt_I1e;
    }
    /* FP64 1/s: one Newton step from the FP32 estimate (relative error
     * about 1.5 x (6e-8)^2) */
    __device__ __forceinline__ double inv_s64(const PlasmaK &k) const {
        double Id = a2 * k.hbar_w0;
        if (Id == 0.0) Id = 1e-30;
        const double x = k.zf * Id, y = (double)inv_s;
        const double r = fma(-x * y, y, 1.0);
        return fma(0.5 * y, r, y);
    }
    __device__ __forceinline__ bool adk_live(const PlasmaK &k, int species) const {
        /* the FP32 exponent with a margin for its rounding */
        return species == 1 ? (k.kon1 != 0.0 && (float)k.ke1 * inv_s <= (float)(ADK_E_MAX * 1.001))
                            : (k.kon2 != 0.0 && (float)k.ke2 * inv_s <= (float)(ADK_E_MAX * 1.001));
    }
    __device__ __forceinline__ double adk_b(const PlasmaK &k, int species) const {
        if (!adk_live(k, species)) return 0.0;
        const double iv = inv_s64(k);
        return species == 1 ? adk_rate(k.kon1, k.ke1, k.L1, k.kon5, iv) * k.Kon
                            : adk_rate(k.kon2, k.ke2, k.L2, k.kon6, iv) * k.Kon;
    }
};

/*
 * Density of one line (plasma.c:218-241, the consistent two-species form):
 *   R_k     = (R_{k+1} + b_k)(1 - a_k) + a_k,                 R_NT = 0
 *   u_k     = 1 - exp(-R_k)      (evaluated as -expm1(-R_k))
 *   ions2_k = (1 - c_k) ions2_{k+1} + c_k u_{k+1},            ions2_NT = 0
 *   ions1_k = u_k - ions2_k ;  F_k = (ions1_k + 2 ions2_k) sigma_k
 * Element k = tid + TPL*m of the line is held by thread tid in slot m.  Two
 * padded FP64 LDS arrays: U holds m_k = 1 - a_k, then u_k; X holds the first
 * scan's c_k, then c_k of the second species, then ions2_k.  The second scan
 * forms its coefficients from U and X as it walks, so no per-element state
 * beyond 1/s stays in registers across the scans.  The scans and u run in
 * FP64 in both precisions.
 * sink(m, F, tau) receives the result per element, in precision T.
 */
template <class T, class G, bool COLD = false, class Sink>
__device__ __forceinline__ void plasma_line(const cx<T> (&A)[G::E], int tid, double *lds,
                                            const PlasmaK &pk, double &rho_max, Sink sink,
                                            double dz_abs = 0.0) {
    using L = PlasmaLds<G, T>;
    constexpr int NT = G::NT;
    constexpr int E = L::E, TPL = L::TPL;
    double *U = lds, *X = lds + L::ARR, *scr = lds + 2 * L::ARR;
    T *S = reinterpret_cast<T *>(lds + 2 * L::ARR + L::SCR), *V = S + L::ARR;   /* used when L::SV */
    int anyc = 0;
    T invs[L::SV ? 1 : E];
    __syncthreads();              /* the LDS may still be read by the caller's last exchange */
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = padp_tm<G>(tid, m);                  /* element tid + TPL*m, swizzled */
        PlasmaElem<T> pe;
        pe.set(A[m].re, A[m].im, pk);
        T tc;
        const T sig = pe.sigma(pk, &tc);
        const double b = pe.adk_b(pk, 1);
        const double a = (double)pe.avalanche(pk, sig);    /* avalanche, SI units */
        U[k] = 1.0 - a;
        X[k] = b * (1.0 - a) + a;
        if constexpr (L::SV) {
            S[k] = sig;
            V[k] = pe.inv_s;
        } else {
            invs[m] = pe.inv_s;
        }
        anyc |= pe.adk_live(pk, 2);
    }
    auto elem = [&](int m) {
        PlasmaElem<T> pe;
        if constexpr (std::is_same_v<T, float>) {
            pe.a2 = (double)A[m].re * (double)A[m].re + (double)A[m].im * (double)A[m].im;
        } else {
            pe.I = 0.0;
        }
        pe.inv_s = L::SV ? V[padp_tm<G>(tid, m)] : invs[L::SV ? 0 : m];
        return pe;
    };
    __syncthreads();
    /* U <- u = 1 - exp(-R); a cold line (COLD) stops after the sum of X */
    double s_line = 0.0;
    double s_lim = 0.0;
    if constexpr (COLD) {
        /* the bound |dz| rho_nt/2 S sigma_bound taken twice over, for the
         * rounding of sigma and tau in T; fmin drops the NaN of 0 * inf
         * (no plasma part then) */
        s_lim = fmin(ColdLine<T>::S_MAX, ColdLine<T>::EPS / (pk.rho_nt * dz_abs * pk.sig_bound));
    }
    const bool scanned = affine_suffix_scan<G, COLD>(tid, scr,
        [&](int k, double &mk, double &ck) { mk = U[padp<G>(k)]; ck = X[padp<G>(k)]; },
        [&](int k, double R) { U[padp<G>(k)] = one_minus_exp_neg(R); },
        anyc ? __builtin_inf() : 0.0,                      /* a second species: never cold */
        [&](double sum) { s_line = sum; return sum <= s_lim; });
    if (!scanned) {
        rho_max = fmax(rho_max, s_line);
        __syncthreads();          /* the caller's next exchange overwrites this LDS */
        return;
    }
    const int block_anyc = __syncthreads_or(anyc);
    if (block_anyc) {
#pragma unroll
        for (int m = 0; m < E; ++m) X[padp_tm<G>(tid, m)] = elem(m).adk_b(pk, 2);
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
        const int k = padp_tm<G>(tid, m);
        const double i2 = block_anyc ? X[k] : 0.0;
        const double i1 = U[k] - i2;
        const double fr = i1 + 2.0 * i2;
        T tc, sig;
        if constexpr (L::SV) {
            sig = S[k];
            tc = (T)pk.kt * V[k];
        } else {
            sig = elem(m).sigma(pk, &tc);
        }
        rho_max = fmax(rho_max, fr);
        sink(m, (T)fr * sig, tc);
    }
    __syncthreads();
}

/* ---------------------------------------------------------------------- */
/* t-line kernel                                                           */
/* ---------------------------------------------------------------------- */
/* per-line LDS: the FFT exchange buffer (doubled, BOFF != 0, while a
 * workgroup's LDS stays within 80 KB, i.e. two workgroups per CU) unioned
 * with the plasma arrays */
template <class T, class G, bool PLASMA> struct TCfg {
    static constexpr int E = G::E, TPL = G::TPL, LPB = G::LPB, THREADS = G::THREADS;
    static constexpr int FFT1 = G::FFT_ELEMS * (int)sizeof(cx<T>);
    static constexpr int PL_BYTES = PLASMA ? PlasmaLds<G, T>::BYTES : 0;
    static constexpr bool DB = LPB * 2 * FFT1 <= 80 * 1024;
    static constexpr int BOFF = DB ? G::FFT_ELEMS : 0;
    static constexpr int FFT_BYTES = (DB ? 2 : 1) * FFT1;
    static constexpr int LINE_BYTES = ((FFT_BYTES > PL_BYTES ? FFT_BYTES : PL_BYTES) + 15) / 16 * 16;
    static_assert(LPB * LINE_BYTES <= 160 * 1024, "t-line workgroup exceeds the LDS");
};

/*
 * sin/cos of the per-step nonlinear phase (small in practice).  Cody-Waite
 * reduction by pi/2 with a three-part constant, then Taylor polynomials on
 * [-pi/4, pi/4] (truncation below 1e-19 in FP64, 2e-9 in FP32) and the
 * quadrant swap.  |x| > 1e5 takes the library routine instead.
 */
static __device__ __noinline__ void sincos_big(double x, double *s, double *c) { sincos(x, s, c); }
static __device__ __noinline__ void sincos_bigf(float x, float *s, float *c) { sincosf(x, s, c); }

template <class T> __device__ __forceinline__ void sincos_t(T x, T *s, T *c);
/* the polynomials on the reduced argument r, |r| <= pi/4 */
__device__ __forceinline__ void sincos_poly(double r, double *sr_out, double *cr_out) {
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
    *sr_out = fma(sp * r2, r, r);
    double cp = 1.0 / 6402373705728000.0;                         /* 1/18! */
    cp = fma(cp, r2, -1.0 / 20922789888000.0);                    /* -1/16! */
    cp = fma(cp, r2, 1.0 / 87178291200.0);                        /* 1/14! */
    cp = fma(cp, r2, -1.0 / 479001600.0);                         /* -1/12! */
    cp = fma(cp, r2, 1.0 / 3628800.0);                            /* 1/10! */
    cp = fma(cp, r2, -1.0 / 40320.0);                             /* -1/8! */
    cp = fma(cp, r2, 1.0 / 720.0);                                /* 1/6! */
    cp = fma(cp, r2, -1.0 / 24.0);                                /* -1/4! */
    cp = fma(cp, r2, 0.5);
    *cr_out = fma(-cp, r2, 1.0);
}
template <> __device__ __forceinline__ void sincos_t<double>(double x, double *s, double *c) {
    /* a whole wave inside (-pi/4, pi/4) (the usual nonlinear phase of a step)
     * skips the reduction, which is the identity there (k = 0, r = x, no
     * quadrant swap): same results, fewer instructions */
    if (__all(fabs(x) < 0.785)) { sincos_poly(x, s, c); return; }
    if (!(fabs(x) <= 1e5)) { sincos_big(x, s, c); return; }
    const double k = rint(x * 0.63661977236758134308);          /* 2/pi */
    double r = fma(-k, 1.57079632673412561417e+00, x);            /* pi/2 = h + m + l */
    r = fma(-k, 6.07710050630396597660e-11, r);
    r = fma(-k, 2.02226624879595063154e-21, r);
    double sr, cr;
    sincos_poly(r, &sr, &cr);
    const int q = ((int)k) & 3;
    const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
__device__ __forceinline__ void sincos_poly(float r, float *sr_out, float *cr_out) {
    const float r2 = r * r;
    float sp = 1.0f / 362880.0f;                                  /* Taylor to r^9 */
    sp = fmaf(sp, r2, -1.0f / 5040.0f);
    sp = fmaf(sp, r2, 1.0f / 120.0f);
    sp = fmaf(sp, r2, -1.0f / 6.0f);
    *sr_out = fmaf(sp * r2, r, r);
    float cp = -1.0f / 3628800.0f;                                /* Taylor to r^10 */
    cp = fmaf(cp, r2, 1.0f / 40320.0f);
    cp = fmaf(cp, r2, -1.0f / 720.0f);
    cp = fmaf(cp, r2, 1.0f / 24.0f);
    cp = fmaf(cp, r2, -0.5f);
    *cr_out = fmaf(cp, r2, 1.0f);
}
template <> __device__ __forceinline__ void sincos_t<float>(float x, float *s, float *c) {
    if (__all(fabsf(x) < 0.785f)) { sincos_poly(x, s, c); return; }   /* as in the FP64 form */
    if (!(fabsf(x) <= 1e5f)) { sincos_bigf(x, s, c); return; }
    const float k = rintf(x * 0.636619772f);
    float r = fmaf(-k, 1.5703125f, x);                            /* pi/2 in three parts */
    r = fmaf(-k, 4.8351287841796875e-4f, r);
    r = fmaf(-k, 3.13916473e-7f, r);
    float sr, cr;
    sincos_poly(r, &sr, &cr);
    const int q = ((int)k) & 3;
    const float ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
template <class T> __device__ __forceinline__ T exp_t(T x);
template <> __device__ __forceinline__ double exp_t<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float exp_t<float>(float x) { return expf(x); }

/* line FFT with twiddle powers cached per thread (TwChain of the plan) */
template <class T, int N, int E, int DIR, int PAD, int BOFF, int Ns, int R, int... Rest>
struct DevFftC {
    static __device__ __forceinline__ void run(cx<T> (&v)[E], int tid, cx<T> *lds, int ls,
                                               const TwChain<T, N, E, Ns, R, Rest...> &ch) {
        constexpr bool last = sizeof...(Rest) == 0;
        fft_stage_t<T, N, E, R, Ns, DIR, last, PAD>(v, tid, lds, ls, TwFromPowers<T, R, E / R, DIR>{ch.p});
        if constexpr (!last) {
            __syncthreads();
            fft_gather<T, N, E, PAD>(v, tid, lds, ls);
            if constexpr (BOFF == 0) {
                __syncthreads();
                DevFftC<T, N, E, DIR, PAD, 0, Ns * R, Rest...>::run(v, tid, lds, ls, ch.next);
            } else {
                DevFftC<T, N, E, DIR, PAD, -BOFF, Ns * R, Rest...>::run(v, tid, lds + BOFF * ls, ls, ch.next);
            }
        }
    }
};
template <class T, int N, int E, class Rad> struct ChainOf;
template <class T, int N, int E, int... Rs> struct ChainOf<T, N, E, Radices<Rs...>> {
    using type = TwChain<T, N, E, 1, Rs...>;
    /* line FFT with the thread's cached twiddle powers; BOFF as for DevFft */
    template <int DIR, int PAD, int BOFF = 0>
    static __device__ __forceinline__ void fft(cx<T> (&v)[E], int tid, cx<T> *lds, int ls, const type &ch) {
        if constexpr (BOFF != 0) __syncthreads();
        DevFftC<T, N, E, DIR, PAD, BOFF, 1, Rs...>::run(v, tid, lds, ls, ch);
    }
};
struct NoChain {};
/* mark every cached twiddle power as redefined here, so products of them are
 * formed where a stage uses them instead of being hoisted out of the line
 * loop and held in registers */
template <class T, int N, int E, int Ns> __device__ __forceinline__ void opaque_chain(TwChain<T, N, E, Ns> &) {}
template <class T, int N, int E, int Ns, int R, int... Rest>
__device__ __forceinline__ void opaque_chain(TwChain<T, N, E, Ns, R, Rest...> &ch) {
    if constexpr (Ns > 1) {
#pragma unroll
        for (int b = 0; b < E / R; ++b)
#pragma unroll
            for (int i = 0; i < tw_npow<R, T>(); ++i) asm volatile("" : "+v"(ch.p[b][i].re), "+v"(ch.p[b][i].im));
    }
    opaque_chain(ch.next);
}

/* workgroups resident on the device at once for kernel k (0: unknown) */
static long resident_blocks(const void *k, int threads, int lds_bytes) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, lds_bytes) != hipSuccess) return 0;
    return (long)ncu * per;
}


} // namespace mpc

#endif
