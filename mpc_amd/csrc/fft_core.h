/*
 * fft_core.h -- register-resident Stockham FFT building blocks for CDNA4.
 *
 * This replaces the reference's cuFFT/FFTW plans (propagation.c:508-613) for
 * the three line geometries (x, y, t).  Conventions are FFTW's: forward
 * (DIR = -1) is exp(-2 pi i jk/n), inverse (DIR = +1) exp(+...), both
 * unnormalised.
 *
 * Work split: a line of N points is owned by TPL = N/E threads; thread tid
 * holds x[tid + TPL*m] for m in [0, E) in registers.  A stage of radix R
 * does E/R butterflies per thread in registers, then writes its outputs to
 * the line's LDS buffer in Stockham order; after a barrier every thread
 * reads back x[tid + TPL*m].  The last stage's outputs land in the same
 * register slots they were read from, so a transform of S stages costs
 * S-1 LDS round trips, and both the first load and the last store use the
 * coalesced pattern x[tid + TPL*m].
 *
 * Everything here is host+device so the index arithmetic can be exercised
 * on the CPU by a loop-over-threads emulation (tests/).
 */
#ifndef MPC_FFT_CORE_H
#define MPC_FFT_CORE_H

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPC_HD __host__ __device__ __forceinline__
#else
#define MPC_HD inline
#endif

namespace mpc {

template <class T> struct cx { T re, im; };

template <class T> MPC_HD cx<T> mk(T a, T b) { cx<T> c; c.re = a; c.im = b; return c; }
template <class T> MPC_HD cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <class T> MPC_HD cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <class T> MPC_HD cx<T> cmul(cx<T> a, cx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T> MPC_HD cx<T> cmulc(cx<T> a, cx<T> b) { /* a * conj(b) */
    return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
template <class T> MPC_HD cx<T> cscale(cx<T> a, T s) { return mk(a.re * s, a.im * s); }
/* a * exp(DIR * i pi/2): DIR = -1 -> a * (-i), DIR = +1 -> a * (+i) */
template <int DIR, class T> MPC_HD cx<T> rot90(cx<T> a) {
    return DIR < 0 ? mk(a.im, -a.re) : mk(-a.im, a.re);
}

/* ---- compile-time twiddles: exp(DIR * 2 pi i m / N) ---------------------- */
constexpr double ce_pi = 3.14159265358979323846264338327950288;
constexpr double ce_sin_poly(double x) {
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; ++i) { term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0)); sum += term; }
    return sum;
}
constexpr double ce_cos_poly(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) { term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i)); sum += term; }
    return sum;
}
/* cos / sin of 2 pi m / N with octant reduction in integer arithmetic */
constexpr double ce_cos2pi(int m, int N) {
    m %= N; if (m < 0) m += N;
    /* reduce to [0, N/2] using cos(2pi - a) = cos a */
    if (2 * m > N) m = N - m;
    /* [0, N/4]: direct; (N/4, N/2]: cos a = -cos(pi - a) */
    if (4 * m > N) return -ce_cos_poly(ce_pi - 2.0 * ce_pi * m / N);
    return ce_cos_poly(2.0 * ce_pi * m / N);
}
/* not recursive, so that device code folds it to a constant instead of
 * calling it (a call gives the kernel a stack in scratch memory) */
constexpr double ce_sin2pi(int m, int N) {
    m %= N; if (m < 0) m += N;
    /* sin(2 pi - a) = -sin a */
    const double sgn = 2 * m > N ? -1.0 : 1.0;
    if (2 * m > N) m = N - m;
    if (4 * m > N) return sgn * ce_sin_poly(ce_pi - 2.0 * ce_pi * m / N);
    return sgn * ce_sin_poly(2.0 * ce_pi * m / N);
}
template <int DIR, class T> MPC_HD cx<T> ctw(int m, int N) {
    return mk((T)ce_cos2pi(m, N), (T)(DIR * ce_sin2pi(m, N)));
}
/* a * exp(DIR * 2 pi i e / N) for an exponent fixed at compile time (after
 * unrolling): quarter turns are exact sign swaps, the rest a complex product */
template <int DIR, class T> MPC_HD cx<T> ctwmul(cx<T> a, int e, int N) {
    if ((4 * e) % N == 0) {
        switch (((4 * e) / N) & 3) {
        case 0: return a;
        case 1: return rot90<DIR>(a);
        case 2: return mk(-a.re, -a.im);
        default: return rot90<-DIR>(a);
        }
    }
    return cmul(a, ctw<DIR, T>(e, N));
}

/* ---- small DFT codelets, in place, natural order ------------------------- */
template <int R, int DIR, class T> struct Dft;

template <int DIR, class T> struct Dft<2, DIR, T> {
    static MPC_HD void run(cx<T> *x) {
        const cx<T> a = x[0], b = x[1];
        x[0] = a + b;
        x[1] = a - b;
    }
};

template <int DIR, class T> struct Dft<3, DIR, T> {
    static MPC_HD void run(cx<T> *x) {
        const T h = (T)0.5, s3 = (T)0.86602540378443864676372317075294;
        const cx<T> t1 = x[1] + x[2];
        const cx<T> t2 = x[1] - x[2];
        const cx<T> m = mk(x[0].re - h * t1.re, x[0].im - h * t1.im);
        /* DIR * i * sqrt(3)/2 * t2 */
        const cx<T> j = rot90<DIR>(cscale(t2, s3));
        x[0] = x[0] + t1;
        x[1] = m + j;
        x[2] = m - j;
    }
};

template <int DIR, class T> struct Dft<4, DIR, T> {
    static MPC_HD void run(cx<T> *x) {
        const cx<T> a0 = x[0] + x[2], a1 = x[0] - x[2];
        const cx<T> a2 = x[1] + x[3], a3 = rot90<DIR>(x[1] - x[3]);
        x[0] = a0 + a2;
        x[2] = a0 - a2;
        x[1] = a1 + a3;
        x[3] = a1 - a3;
    }
};

/* Cooley-Tukey composite R1*R2: n = R2*n1 + n2, k = k1 + R1*k2 */
template <int R1, int R2, int DIR, class T> struct DftCT {
    static MPC_HD void run(cx<T> *x) {
        constexpr int N = R1 * R2;
        cx<T> z[N];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            cx<T> a[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) a[n1] = x[R2 * n1 + n2];
            Dft<R1, DIR, T>::run(a);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1)
                z[n2 * R1 + k1] = ctwmul<DIR, T>(a[k1], n2 * k1, N);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            cx<T> b[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) b[n2] = z[n2 * R1 + k1];
            Dft<R2, DIR, T>::run(b);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) x[k1 + R1 * k2] = b[k2];
        }
    }
};
template <int DIR, class T> struct Dft<8, DIR, T> {
    static MPC_HD void run(cx<T> *x) { DftCT<2, 4, DIR, T>::run(x); }
};
template <int DIR, class T> struct Dft<12, DIR, T> {
    static MPC_HD void run(cx<T> *x) { DftCT<4, 3, DIR, T>::run(x); }
};
template <int DIR, class T> struct Dft<16, DIR, T> {
    static MPC_HD void run(cx<T> *x) { DftCT<4, 4, DIR, T>::run(x); }
};

/*
 * LDS slot of element i of an exchange buffer.  PAD > 0: one padding slot
 * every PAD elements.  PAD < 0: no padding; the low five bits of i are XORed
 * with bits [s, s+5), s = -PAD.  With s = log2(elements per thread) the
 * strided writes of the early Stockham stages (lane stride E) and the
 * unit-stride gathers both hit distinct banks for every 16 or 32 lanes of
 * 8-byte accesses, which padding cannot give for both at once.  The XOR form
 * is a bijection on every multiple of 32 elements and stays inside it.
 * PAD = 0: identity.
 */
template <int PAD> MPC_HD int padi(int i) {
    if constexpr (PAD > 0) return i + i / PAD;
    else if constexpr (PAD < 0) return i ^ ((i >> (-PAD)) & 31);
    else return i;
}
/* padi<PAD>(a + STRIDE * m), written so that the m-dependent part becomes a
 * constant offset (an immediate in the LDS instruction) where the layout
 * allows it.  XOR swizzle (PAD = -s): the mask reads index bits s..s+4 and
 * flips bits 0..4, and a multiple of 2^(s+5) touches neither, so
 * padi(a + c) = padi(a) + c for such c.  Padding (PAD > 0): a multiple c of
 * PAD gives padi(a + c) = padi(a) + c + c / PAD. */
template <int PAD, int STRIDE> MPC_HD int padi_strided(int a, int m) {
    if constexpr (PAD < 0 && STRIDE % (1 << (-PAD + 5)) == 0) return padi<PAD>(a) + STRIDE * m;
    else if constexpr (PAD > 0 && STRIDE % PAD == 0) return padi<PAD>(a) + (STRIDE + STRIDE / PAD) * m;
    else return padi<PAD>(a + STRIDE * m);
}
/* elements an exchange buffer of n elements occupies under padi<PAD> */
template <int PAD> constexpr int padded_elems(int n) { return PAD > 0 ? n + n / PAD : n; }
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

/*
 * Stage twiddles w[r] = W^(r*idx) (conjugated for the inverse), r < R, from
 * the table W^m = tw[m] = exp(-2 pi i m / N), each entry rounded once from
 * its exact value.  r*idx < N, so no wrap.  The split (load_powers, then
 * expand_powers) lets a kernel load them once and use them for every
 * transform it runs.
 * FP32 loads every twiddle.  A product of rounded entries carries their
 * rounding errors and its own, the same in every transform, and over hundreds
 * of steps of FP32 transform pairs that drift grew linearly, at 2-4x the
 * drift with every twiddle loaded.
 * FP64 loads only the power-of-two powers and forms the rest as products, at
 * most three multiplications deep: a radix-8 stage loads 3 entries instead of
 * 7, a radix-16 stage 4 instead of 15, and FP64 products stay far below what
 * an FP64 run resolves.
 */
template <class T> constexpr bool tw_products(int R) {
    return sizeof(T) == 8 && (R == 2 || R == 3 || R == 4 || R == 8 || R == 16);
}
template <int R, class T> constexpr int tw_npow() {
    if constexpr (tw_products<T>(R)) return R == 2 ? 1 : (R == 3 || R == 4) ? 2 : R == 8 ? 3 : 4;
    else return R - 1;
}
template <int R, class T>
MPC_HD void load_powers(cx<T> (&p)[(tw_npow<R, T>())], const cx<T> *tw, int idx) {
    if constexpr (tw_products<T>(R)) {
        p[0] = tw[idx];
        if constexpr (R > 2) p[1] = tw[2 * idx];
        if constexpr (R >= 8) p[2] = tw[4 * idx];
        if constexpr (R == 16) p[3] = tw[8 * idx];
    } else {
#pragma unroll
        for (int r = 1; r < R; ++r) p[r - 1] = tw[r * idx];
    }
}
template <int R, int DIR, class T>
MPC_HD void expand_powers(cx<T> (&w)[R], const cx<T> (&p)[(tw_npow<R, T>())]) {
    auto cj = [](cx<T> v) {
        if (DIR > 0) v.im = -v.im;
        return v;
    };
    w[0] = mk((T)1, (T)0);
    if constexpr (tw_products<T>(R)) {
        w[1] = cj(p[0]);
        if constexpr (R > 2) {
            w[2] = cj(p[1]);
            if constexpr (R > 3) {
                w[3] = cmul(w[1], w[2]);
                if constexpr (R > 4) {
                    w[4] = cj(p[2]);
                    w[5] = cmul(w[1], w[4]);
                    w[6] = cmul(w[2], w[4]);
                    w[7] = cmul(w[3], w[4]);
                    if constexpr (R > 8) {
                        w[8] = cj(p[3]);
#pragma unroll
                        for (int r = 9; r < R; ++r) w[r] = cmul(w[r - 8], w[8]);
                    }
                }
            }
        }
    } else {
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cj(p[r - 1]);
    }
}
template <int R, int DIR, class T>
MPC_HD void stage_twiddles(cx<T> (&w)[R], const cx<T> *tw, int idx) {
    cx<T> p[(tw_npow<R, T>())];
    load_powers<R, T>(p, tw, idx);
    expand_powers<R, DIR>(w, p);
}

/* twiddle sources for fft_stage: tw(b, k, w) fills w[R] for butterfly b
 * (of the thread's E/R) whose in-group index is k */
template <class T, int N, int R, int Ns, int DIR> struct TwFromTable {
    const cx<T> *tw;
    MPC_HD void operator()(int, int k, cx<T> (&w)[R]) const { stage_twiddles<R, DIR>(w, tw, k * (N / (Ns * R))); }
};
template <class T, int R, int B, int DIR> struct TwFromPowers {
    const cx<T> (*p)[(tw_npow<R, T>())];   /* p[b] */
    MPC_HD void operator()(int b, int, cx<T> (&w)[R]) const { expand_powers<R, DIR>(w, p[b]); }
};

/*
 * One Stockham stage.  Ns = product of the radices of earlier stages.
 * v: the thread's E registers (x[tid + TPL*m]); lds: the line's exchange
 * buffer, element idx at lds[padi<PAD>(idx) * ls]; twf: twiddle source.
 */
template <class T, int N, int E, int R, int Ns, int DIR, bool LAST, int PAD, class TwF>
MPC_HD void fft_stage_t(cx<T> (&v)[E], int tid, cx<T> *lds, int ls, const TwF &twf) {
    constexpr int TPL = N / E;
    constexpr int B = E / R;
    static_assert(E % R == 0, "radix must divide elements per thread");
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + TPL * b;
        const int k = j % Ns;
        cx<T> u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[b + r * B];
        if (Ns > 1) {
            cx<T> w[R];
            twf(b, k, w);
#pragma unroll
            for (int r = 1; r < R; ++r) u[r] = cmul(u[r], w[r]);
        }
        Dft<R, DIR, T>::run(u);
        if (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = u[r];
        } else {
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[padi<PAD>(base + r * Ns) * ls] = u[r];
        }
    }
}
/* a stage with twiddles from the N-entry table exp(-2 pi i k/N) */
template <class T, int N, int E, int R, int Ns, int DIR, bool LAST, int PAD = 0>
MPC_HD void fft_stage(cx<T> (&v)[E], int tid, cx<T> *lds, int ls, const cx<T> *tw) {
    fft_stage_t<T, N, E, R, Ns, DIR, LAST, PAD>(v, tid, lds, ls, TwFromTable<T, N, R, Ns, DIR>{tw});
}

/*
 * Per-thread twiddle powers of every stage of a plan (radices Rs, first
 * stage at group size Ns), loaded once with load(tw, tid) and reused by
 * every transform of that length the thread runs, in either direction.
 */
template <class T, int N, int E, int Ns, int... Rs> struct TwChain {
    MPC_HD void load(const cx<T> *, int) {}
};
template <class T, int N, int E, int Ns, int R, int... Rest> struct TwChain<T, N, E, Ns, R, Rest...> {
    static constexpr int B = E / R, NP = tw_npow<R, T>(), TPL = N / E;
    cx<T> p[B][NP];
    TwChain<T, N, E, Ns * R, Rest...> next;
    MPC_HD void load(const cx<T> *tw, int tid) {
        if constexpr (Ns > 1) {
#pragma unroll
            for (int b = 0; b < B; ++b) load_powers<R, T>(p[b], tw, ((tid + TPL * b) % Ns) * (N / (Ns * R)));
        }
        next.load(tw, tid);
    }
};

template <class T, int N, int E, int PAD = 0>
MPC_HD void fft_gather(cx<T> (&v)[E], int tid, const cx<T> *lds, int ls) {
    constexpr int TPL = N / E;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[padi_strided<PAD, TPL>(tid, m) * ls];
}

/* radix plans: a compile-time list of radices whose product is N */
template <int... Rs> struct Radices {};

/* t lines: 8 elements per thread keeps the fused nonlinear kernel's register
 * footprint low (two line copies live at once) */
template <int N> struct TPlan;
template <> struct TPlan<4096> { static constexpr int E = 8; using R = Radices<8, 8, 8, 8>; };
template <> struct TPlan<2048> { static constexpr int E = 8; using R = Radices<8, 8, 8, 4>; };
template <> struct TPlan<1024> { static constexpr int E = 8; using R = Radices<8, 8, 8, 2>; };
template <> struct TPlan<512>  { static constexpr int E = 8; using R = Radices<8, 8, 8>; };
template <> struct TPlan<256>  { static constexpr int E = 8; using R = Radices<8, 8, 4>; };

template <int N> struct Plan;
/* t lines */
template <> struct Plan<4096> { static constexpr int E = 16; using R = Radices<16, 16, 16>; };
template <> struct Plan<2048> { static constexpr int E = 16; using R = Radices<16, 16, 8>; };
template <> struct Plan<1024> { static constexpr int E = 16; using R = Radices<16, 16, 4>; };
template <> struct Plan<512>  { static constexpr int E = 16; using R = Radices<16, 16, 2>; };
/* x / y lines */
template <> struct Plan<256>  { static constexpr int E = 16; using R = Radices<16, 16>; };
template <> struct Plan<192>  { static constexpr int E = 24; using R = Radices<8, 8, 3>; };
template <> struct Plan<128>  { static constexpr int E = 16; using R = Radices<16, 8>; };
template <> struct Plan<64>   { static constexpr int E = 16; using R = Radices<16, 4>; };
template <> struct Plan<32>   { static constexpr int E = 8;  using R = Radices<8, 4>; };
template <> struct Plan<16>   { static constexpr int E = 16; using R = Radices<16>; };

} // namespace mpc

#endif
