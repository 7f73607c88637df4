/*
 * axis.hip -- the x- and y-line sweeps of the split-step propagator.
 *
 * k_axis: FFT along x or y of a tile of G consecutive t values (layout
 * A[y][x][t], t fastest, so a tile row is G*sizeof(complex) contiguous
 * bytes), x D(k), inverse FFT [x mirror] [x per-t factor], in place or from
 * a separate source field (src).  Replaces the x/y
 * parts of linearstep (propagation.c:31-57, 77-100); the diffraction operator
 * is separable, so x and y each take one sweep.
 *
 * Built with machine-level loop-invariant code motion off, like tline.hip:
 * the persistent form's tile loop would otherwise keep every product of the
 * cached twiddle powers in registers.
 *
 * Multiply-adds are contracted per source expression only (the pragma below
 * precedes the headers, so it covers the FFT code they define): the
 * per-field and the batched kernels then round identically, whatever the
 * code around the shared arithmetic.  These sweeps are bound by HBM, not by
 * their FMAs.
 */
#pragma clang fp contract(on)
#include "device.h"

namespace mpc {

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

/*
 * Tile lin -> (t-chunk, outer): runs of tg consecutive t-chunks, then the
 * outer index, then the next run.  tg = all chunks walks t fastest; a small
 * tg spreads the concurrently running tiles over the outer index.
 */
template <class T> __device__ __forceinline__ void axis_tile_of(const AArgs<T> &a, long lin, long &outer, long &t0, int G) {
    const long rest = lin / a.tg;
    outer = rest % a.n_outer;
    t0 = ((rest / a.n_outer) * a.tg + lin % a.tg) * G;
}

/* the arithmetic of one tile (this thread's E elements of one line in v):
 * FFT, x D(k), inverse FFT [x mirror] [x per-t factor p].  Shared by the
 * per-field and the batched sweeps. */
template <class T, int N, class PL, int G_, bool PERSIST, class Fft>
__device__ __forceinline__ void axis_apply(const AArgs<T> &a, cx<T> (&v)[ACfg<T, N, PL, G_>::E], cx<T> p, long outer,
                                           int tid, const cx<T> (&Dc)[ACfg<T, N, PL, G_>::E], Fft &fft) {
    using C = ACfg<T, N, PL, G_>;
    constexpr int E = C::E, TPL = C::TPL;
    fft(v, std::integral_constant<int, -1>{});
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(v[m], PERSIST ? Dc[m] : gload(a.D, tid + TPL * m));
    fft(v, std::integral_constant<int, +1>{});
    if (a.M) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], gload(a.M, (long)(tid + TPL * m) * a.m_line + outer * a.m_outer));
    }
    if (a.Pt) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], p);
    }
}

/*
 * With PERSIST the grid is what fits on the device at once; each workgroup
 * walks tiles blockIdx.x, + gridDim.x, ..., keeps its twiddle powers and its
 * D(k) entries in registers, and loads the next tile (and its per-t factor)
 * while it transforms the current one.  Without PERSIST there is one
 * workgroup per tile and the tables are read per tile.
 */
template <class T, int N, class PL, int G_, bool PERSIST>
__device__ __forceinline__ void axis_work(const AArgs<T> &a, long lin, long lstride) {
    using C = ACfg<T, N, PL, G_>;
    using CO = ChainOf<T, N, C::E, typename PL::R>;
    constexpr int E = C::E, TPL = C::TPL, G = C::G;
    __shared__ __align__(16) cx<T> lds[N * G];
    const int gcol0 = threadIdx.x % G, tid0 = threadIdx.x / G;
    const long ntiles = (long)(a.nt / G) * a.n_outer;
    if (lin >= ntiles) return;                        /* whole workgroup */

    std::conditional_t<PERSIST, typename CO::type, NoChain> twc;
    cx<T> Dc[E];
    if constexpr (PERSIST) {
        twc.load(a.tw, tid0);
#pragma unroll
        for (int m = 0; m < E; ++m) Dc[m] = gload(a.D, tid0 + TPL * m);
    }
    int tid = tid0, gcol = gcol0;
    auto fft = [&](cx<T>(&x)[E], auto dir) {
        constexpr int DIR = decltype(dir)::value;
        if constexpr (PERSIST) CO::template fft<DIR, 0>(x, tid, lds + gcol, G, twc);
        else line_fft<T, N, PL, DIR>(x, tid, lds + gcol, G, a.tw);
    };

    const cx<T> *in = a.src ? a.src : a.field;
    cx<T> nxt[E], pn = mk((T)1, (T)0);
    {
        long outer, t0;
        axis_tile_of(a, lin, outer, t0, G);
        const cx<T> *b0 = in + outer * a.s_outer + t0 + gcol0;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = gload(b0, (long)(tid0 + TPL * m) * a.s_line);
        if (a.Pt) pn = gload(a.Pt, t0 + gcol0);
    }
    for (;;) {
        /* tid and gcol re-enter as opaque values each iteration, so index
         * arithmetic is recomputed where it is used instead of being held */
        tid = tid0;
        gcol = gcol0;
        asm volatile("" : "+v"(tid), "+v"(gcol));
        if constexpr (PERSIST) opaque_chain(twc);
        long outer, t0;
        axis_tile_of(a, lin, outer, t0, G);
        cx<T> *base = a.field + outer * a.s_outer + t0 + gcol;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        const cx<T> p = pn;
        const long lnext = lin + lstride;
        if constexpr (PERSIST) {
            if (lnext < ntiles) {
                long on, tn;
                axis_tile_of(a, lnext, on, tn, G);
                const cx<T> *bn = in + on * a.s_outer + tn + gcol;
#pragma unroll
                for (int m = 0; m < E; ++m) nxt[m] = gload(bn, (long)(tid + TPL * m) * a.s_line);
                if (a.Pt) pn = gload(a.Pt, tn + gcol);
            }
        }
        axis_apply<T, N, PL, G_, PERSIST>(a, v, p, outer, tid, Dc, fft);
#pragma unroll
        for (int m = 0; m < E; ++m) gstore(base, (long)(tid + TPL * m) * a.s_line, v[m]);
        if (!PERSIST || lnext >= ntiles) break;
        lin = lnext;
    }
}
template <class T, int N, class PL, int G_, bool PERSIST>
__global__ void __launch_bounds__((ACfg<T, N, PL, G_>::THREADS)) k_axis(AArgs<T> a) {
    axis_work<T, N, PL, G_, PERSIST>(a, blockIdx.x, gridDim.x);
}
/*
 * Batched form: one workgroup per tile of every field in d[0..n), which
 * share the line length; workgroup b takes field p, tile b - tstart[p], for
 * the largest p with tstart[p] <= b.  Same bits per field as k_axis<...,
 * false> with that field's tile order.
 */
template <class T, int N, class PL, int G_>
__global__ void __launch_bounds__((ACfg<T, N, PL, G_>::THREADS))
k_axis_batch(const AArgs<T> *__restrict__ d, const long *__restrict__ tstart, int n) {
    const long b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const AArgs<T> a = d[lo];
    const long t = b - tstart[lo];
    axis_work<T, N, PL, G_, false>(a, t, (long)(a.nt / G_) * a.n_outer);
}

/*
 * Persistent batched form: the grid is what fits on the device at once and
 * workgroup w walks the concatenated tiles of every field in d[0..n) (field
 * p owns [tstart[p], tstart[p + 1]), tstart[n] = total), reloading the
 * field's arguments, twiddle powers and D(k) entries at each field boundary
 * and prefetching the next tile across boundaries.  Same arithmetic per tile
 * as the persistent per-field kernel (axis_apply), with each field's tile
 * order.
 */
template <class T, int N, class PL, int G_>
__global__ void __launch_bounds__((ACfg<T, N, PL, G_>::THREADS))
k_axis_batch_p(const AArgs<T> *__restrict__ d, const long *__restrict__ tstart, int n) {
    using C = ACfg<T, N, PL, G_>;
    using CO = ChainOf<T, N, C::E, typename PL::R>;
    constexpr int E = C::E, TPL = C::TPL, G = C::G;
    __shared__ __align__(16) cx<T> lds[N * G];
    const int gcol0 = threadIdx.x % G, tid0 = threadIdx.x / G;
    const long total = tstart[n];
    long l = blockIdx.x;
    if (l >= total) return;                           /* whole workgroup */
    int p = 0;
    {
        int hi = n - 1;
        while (p < hi) {
            const int mid = (p + hi + 1) / 2;
            if (tstart[mid] <= l) p = mid;
            else hi = mid - 1;
        }
    }
    AArgs<T> a = d[p];
    typename CO::type twc;
    cx<T> Dc[E];
    auto load_field = [&] {
        twc.load(a.tw, tid0);
#pragma unroll
        for (int m = 0; m < E; ++m) Dc[m] = gload(a.D, tid0 + TPL * m);
    };
    load_field();
    int tid = tid0, gcol = gcol0;
    auto fft = [&](cx<T>(&x)[E], auto dir) {
        constexpr int DIR = decltype(dir)::value;
        CO::template fft<DIR, 0>(x, tid, lds + gcol, G, twc);
    };

    cx<T> nxt[E], pn = mk((T)1, (T)0);
    {
        long outer, t0;
        axis_tile_of(a, l - tstart[p], outer, t0, G);
        const cx<T> *b0 = (a.src ? a.src : a.field) + outer * a.s_outer + t0 + gcol0;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = gload(b0, (long)(tid0 + TPL * m) * a.s_line);
        if (a.Pt) pn = gload(a.Pt, t0 + gcol0);
    }
    for (;;) {
        tid = tid0;
        gcol = gcol0;
        asm volatile("" : "+v"(tid), "+v"(gcol));
        opaque_chain(twc);
        long outer, t0;
        axis_tile_of(a, l - tstart[p], outer, t0, G);
        cx<T> *base = a.field + outer * a.s_outer + t0 + gcol;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        const cx<T> pf = pn;
        const long ln = l + gridDim.x;
        int q = p;
        while (q + 1 < n && tstart[q + 1] <= ln) ++q;
        if (ln < total) {
            const AArgs<T> &an = d[q];
            long on, tn;
            axis_tile_of(an, ln - tstart[q], on, tn, G);
            const cx<T> *bn = (an.src ? an.src : an.field) + on * an.s_outer + tn + gcol;
#pragma unroll
            for (int m = 0; m < E; ++m) nxt[m] = gload(bn, (long)(tid + TPL * m) * an.s_line);
            if (an.Pt) pn = gload(an.Pt, tn + gcol);
        }
        axis_apply<T, N, PL, G_, true>(a, v, pf, outer, tid, Dc, fft);
#pragma unroll
        for (int m = 0; m < E; ++m) gstore(base, (long)(tid + TPL * m) * a.s_line, v[m]);
        if (ln >= total) break;
        if (q != p) {
            /* the next field's tables; the LDS of the last tile is read by
             * now (the inverse FFT's last exchange precedes its last stage) */
            p = q;
            a = d[p];
            load_field();
        }
        l = ln;
    }
}

/*
 * The last y sweep of an adaptive attempt's fine branch C2 (two dz steps),
 * fused with what the step-size control needs (propagation.c:343-362,
 * :616-631; MPC.c:377-381).  Per tile (x, t-chunk) it finishes C2 as the y
 * sweep does, then, with C1 (one 2dz step) read at the same points:
 *   err_part[2 tile + 0/1] = sum |C2 - C1|^2, sum |C2|^2      (FP64)
 *   lp_part[chunk][y][x]   = sum over the chunk's t of |C2|^2 (FP64, t order)
 * and stores R = 4/3 C2 - 1/3 C1 (formed in FP64) in place of C2, so an
 * accepted step only has to adopt that buffer.  Tiles with the chunks of
 * each x consecutive; the plan is Plan<N> (radix 16 where it
 * applies) where its workgroup fits in 1024 threads, else the 8-per-thread
 * one.
 */
template <class T, int N> struct RichCfg {
    static constexpr int G = ATile<T>::G;
    using PL = std::conditional_t<(G * (N / Plan<N>::E) <= 1024), Plan<N>, APlan8<N>>;
    using C = ACfg<T, N, PL, G>;
};
/* PERSIST: the grid is what fits on the device at once and each workgroup
 * walks tiles blockIdx.x, + gridDim.x, ..., with its twiddle powers and D(k)
 * entries in registers (as k_axis's persistent form); the per-tile work and
 * outputs are the same either way */
template <class T, int N, bool PERSIST>
__global__ void __launch_bounds__((RichCfg<T, N>::C::THREADS))
k_axis_richardson(AArgs<T> a, const cx<T> *__restrict__ c1, double *__restrict__ err_part,
                  double *__restrict__ lp_part, long nxy, int nx) {
    using RC = RichCfg<T, N>;
    using PL = typename RC::PL;
    using C = typename RC::C;
    using CO = ChainOf<T, N, C::E, typename PL::R>;
    constexpr int E = C::E, TPL = C::TPL, G = C::G, THREADS = C::THREADS;
    __shared__ __align__(16) cx<T> lds[N * G + G];   /* G more: the padded line-power rows below */
    constexpr int W = THREADS < 64 ? THREADS : 64, NWV = THREADS / W;   /* lanes per wave, waves */
    __shared__ double wred[2 * NWV];
    const int gcol0 = threadIdx.x % G, tid0 = threadIdx.x / G;
    const long ntiles = (long)(a.nt / G) * a.n_outer;
    std::conditional_t<PERSIST, typename CO::type, NoChain> twc;
    cx<T> Dc[E];
    if constexpr (PERSIST) {
        twc.load(a.tw, tid0);
#pragma unroll
        for (int m = 0; m < E; ++m) Dc[m] = gload(a.D, tid0 + TPL * m);
    }
    int tid = tid0, gcol = gcol0;
    auto fft = [&](cx<T>(&x)[E], auto dir) {
        constexpr int DIR = decltype(dir)::value;
        if constexpr (PERSIST) CO::template fft<DIR, 0>(x, tid, lds + gcol, G, twc);
        else line_fft<T, N, PL, DIR>(x, tid, lds + gcol, G, a.tw);
    };
    const cx<T> *in = a.src ? a.src : a.field;
    for (long lin = blockIdx.x; lin < ntiles; lin += PERSIST ? (long)gridDim.x : ntiles) {
        tid = tid0;
        gcol = gcol0;
        if constexpr (PERSIST) {
            /* index arithmetic is recomputed where it is used instead of
             * being held across the tile loop */
            asm volatile("" : "+v"(tid), "+v"(gcol));
            opaque_chain(twc);
        }
        long outer, t0;
        axis_tile_of(a, lin, outer, t0, G);
        const long off0 = outer * a.s_outer + t0 + gcol;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = gload(in + off0, (long)(tid + TPL * m) * a.s_line);
        const cx<T> p = a.Pt ? gload(a.Pt, t0 + gcol) : mk((T)1, (T)0);
        axis_apply<T, N, PL, G, PERSIST>(a, v, p, outer, tid, Dc, fft);
        __syncthreads();                              /* the transform's last LDS reads are done */
        /* |C2|^2 per (t in chunk, line along y), rows padded to N + 1 doubles
         * so that both the writes (consecutive lanes: consecutive t) and the
         * reads (consecutive lanes: consecutive lines) are free of bank
         * conflicts */
        constexpr int PW = N + 1;
        static_assert(G * PW * (int)sizeof(double) <= (N * G + G) * (int)sizeof(cx<T>),
                      "line-power scratch must fit in the transform's LDS");
        double *pw_t = reinterpret_cast<double *>(lds);
        double e1 = 0.0, e2 = 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const long off = (long)(tid + TPL * m) * a.s_line;
            const cx<T> u = gload(c1 + off0, off);
            const double cr = (double)v[m].re, ci = (double)v[m].im, ur = (double)u.re, ui = (double)u.im;
            const double dr = cr - ur, di = ci - ui, pw = cr * cr + ci * ci;
            e1 += dr * dr + di * di;
            e2 += pw;
            pw_t[gcol * PW + tid + TPL * m] = pw;
            gstore(a.field + off0, off,
                   mk((T)((4.0 / 3.0) * cr - (1.0 / 3.0) * ur), (T)((4.0 / 3.0) * ci - (1.0 / 3.0) * ui)));
        }
        /* block sums of the two error norms: waves, then the wave totals in order */
        for (int o = W / 2; o > 0; o >>= 1) {
            e1 += __shfl_down(e1, o, W);
            e2 += __shfl_down(e2, o, W);
        }
        if (threadIdx.x % W == 0) {
            wred[2 * (threadIdx.x / W)] = e1;
            wred[2 * (threadIdx.x / W) + 1] = e2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s1 = 0.0, s2 = 0.0;
            for (int w = 0; w < NWV; ++w) { s1 += wred[2 * w]; s2 += wred[2 * w + 1]; }
            err_part[2 * lin] = s1;
            err_part[2 * lin + 1] = s2;
        }
        /* per-line power over the chunk's t values, in t order: line y at x = outer */
        const long chunk = t0 / G;
        for (int j = threadIdx.x; j < N; j += THREADS) {
            double sum = 0.0;
            for (int g = 0; g < G; ++g) sum += pw_t[g * PW + j];
            lp_part[chunk * nxy + (long)j * nx + outer] = sum;
        }
        /* wred and pw_t are read above; the next tile's transform writes lds */
        if constexpr (PERSIST) __syncthreads();
    }
}
template <class T, int N>
static int rich_go(AArgs<T> a, int nx, const cx<T> *c1, double *err_part, double *lp_part, hipStream_t s) {
    using C = typename RichCfg<T, N>::C;
    if (a.nt % C::G) return MPC_EINVAL;
    const int nch = a.nt / C::G;
    a.n_outer = nx;
    a.tg = nch;                                        /* the chunks of one x consecutive */
    const long nb = (long)nch * nx;
    if (nb <= 0 || nb > 0x7fffffffL) return MPC_EINVAL;
    /* FP32 walks the tiles with resident workgroups (0.87 against 0.97 ms at
     * 256^2 x 2048); in FP64 the cached tables cost the form its second wave
     * per SIMD (1.71 against 1.52 ms), so it keeps one workgroup per tile */
    if constexpr (sizeof(T) == 4) {
        static long resident = -1;                    /* per instantiation: the same on every device of one model */
        if (resident < 0) resident = resident_blocks((const void *)k_axis_richardson<T, N, true>, C::THREADS, 0);
        if (resident > 0) {
            hipLaunchKernelGGL((k_axis_richardson<T, N, true>), dim3((unsigned)std::min(nb, resident)),
                               dim3(C::THREADS), 0, s, a, c1, err_part, lp_part, (long)nx * N, nx);
            return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
        }
    }
    hipLaunchKernelGGL((k_axis_richardson<T, N, false>), dim3((unsigned)nb), dim3(C::THREADS), 0, s, a, c1, err_part,
                       lp_part, (long)nx * N, nx);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T>
int launch_axis_richardson(const AArgs<T> &a, int n, const cx<T> *c1, double *err_part, double *lp_part, hipStream_t s) {
    switch (n) {
    case 16:  return rich_go<T, 16>(a, n, c1, err_part, lp_part, s);
    case 32:  return rich_go<T, 32>(a, n, c1, err_part, lp_part, s);
    case 64:  return rich_go<T, 64>(a, n, c1, err_part, lp_part, s);
    case 128: return rich_go<T, 128>(a, n, c1, err_part, lp_part, s);
    case 192: return rich_go<T, 192>(a, n, c1, err_part, lp_part, s);
    case 256: return rich_go<T, 256>(a, n, c1, err_part, lp_part, s);
    case 512: return rich_go<T, 512>(a, n, c1, err_part, lp_part, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_axis_richardson<double>(const AArgs<double> &, int, const cx<double> *, double *, double *,
                                            hipStream_t);
template int launch_axis_richardson<float>(const AArgs<float> &, int, const cx<float> *, double *, double *,
                                           hipStream_t);
/* tiles of the fused sweep (its partial-sum slots): t-chunks x Nx */
int axis_richardson_tiles(int precision, int nt, int nx) {
    const int G = precision == MPC_F32 ? ATile<float>::G : ATile<double>::G;
    return nt / G * nx;
}

bool axis_supported(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 192 || n == 256 || n == 512; }
int axis_variants(void) { return 32; }

/* tile orders: runs of all / 1 / 4 / 16 t-chunks before the outer index moves */
static int axis_tg(int order, int nchunks) {
    static const int tg[4] = {0, 1, 4, 16};
    const int g = tg[order & 3];
    return (g == 0 || nchunks % g) ? nchunks : g;
}
template <class T, int N, class PL, int G, bool PERSIST>
static int axis_go(AArgs<T> a, int outer, int order, hipStream_t s) {
    using C = ACfg<T, N, PL, G>;
    if (a.nt % G || outer <= 0) return MPC_EINVAL;
    const int nch = a.nt / G;
    a.n_outer = outer;
    a.tg = axis_tg(order, nch);
    long nb = (long)nch * outer;
    if (PERSIST) {
        static long resident = -1;              /* per instantiation: the same on every device of one model */
        if (resident < 0) resident = resident_blocks((const void *)k_axis<T, N, PL, G, PERSIST>, C::THREADS, 0);
        if (resident > 0 && resident < nb) nb = resident;
    }
    if (nb <= 0 || nb > 0x7fffffffL) return MPC_EINVAL;
    hipLaunchKernelGGL((k_axis<T, N, PL, G, PERSIST>), dim3((unsigned)nb), dim3(C::THREADS), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
/* variant = plan + 4 * order + 16 * mode.  plan 0: 16-or-more per thread
 * (Plan), 256-B rows; 1: Plan, 128-B rows; 2: 8 per thread, 128-B rows; 3: 8
 * per thread, 256-B rows.  order: see axis_tg.  mode 0: one workgroup per
 * tile; 1: persistent workgroups with cached tables and prefetch. */
template <class T, int N, int PERSIST>
static int axis_p(const AArgs<T> &a, int outer, int plan, int order, hipStream_t s) {
    constexpr int G = ATile<T>::G;
    constexpr bool P = PERSIST != 0;
    switch (plan) {
    case 1: return axis_go<T, N, Plan<N>, G / 2, P>(a, outer, order, s);
    case 2: return axis_go<T, N, APlan8<N>, G / 2, P>(a, outer, order, s);
    case 3: return axis_go<T, N, APlan8<N>, G, P>(a, outer, order, s);
    default: return axis_go<T, N, Plan<N>, G, P>(a, outer, order, s);
    }
}
template <class T, int N> static int axis_v(const AArgs<T> &a, int outer, int variant, hipStream_t s) {
    const int plan = variant % 4, order = (variant / 4) % 4;
    return (variant / 16) ? axis_p<T, N, 1>(a, outer, plan, order, s) : axis_p<T, N, 0>(a, outer, plan, order, s);
}
/* the tile geometry a variant gives: t values per tile, and tile order */
int axis_tile_g(int precision, int variant) {
    const int G = precision == MPC_F32 ? ATile<float>::G : ATile<double>::G;
    return (variant % 4 == 1 || variant % 4 == 2) ? G / 2 : G;
}
int axis_tile_tg(int variant, int nchunks) { return axis_tg((variant / 4) % 4, nchunks); }

template <class T, int N, class PL, int G>
static int axis_batch_go(const AArgs<T> *d, const long *tstart, int n, long total, bool persist, hipStream_t s) {
    using C = ACfg<T, N, PL, G>;
    if (n <= 0 || total <= 0 || total > 0x7fffffffL) return MPC_EINVAL;
    if (persist) {
        static long resident = -1;              /* per instantiation */
        if (resident < 0) resident = resident_blocks((const void *)k_axis_batch_p<T, N, PL, G>, C::THREADS, 0);
        const long nb = (resident > 0 && resident < total) ? resident : total;
        hipLaunchKernelGGL((k_axis_batch_p<T, N, PL, G>), dim3((unsigned)nb), dim3(C::THREADS), 0, s, d, tstart, n);
    } else {
        hipLaunchKernelGGL((k_axis_batch<T, N, PL, G>), dim3((unsigned)total), dim3(C::THREADS), 0, s, d, tstart, n);
    }
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T, int N>
static int axis_batch_v(const AArgs<T> *d, const long *tstart, int n, long total, int variant, hipStream_t s) {
    constexpr int G = ATile<T>::G;
    const bool persist = variant / 16 != 0;
    switch (variant % 4) {
    case 1: return axis_batch_go<T, N, Plan<N>, G / 2>(d, tstart, n, total, persist, s);
    case 2: return axis_batch_go<T, N, APlan8<N>, G / 2>(d, tstart, n, total, persist, s);
    case 3: return axis_batch_go<T, N, APlan8<N>, G>(d, tstart, n, total, persist, s);
    default: return axis_batch_go<T, N, Plan<N>, G>(d, tstart, n, total, persist, s);
    }
}
/* batched x/y sweep: d[p] carries each field's strides, tables, n_outer and
 * tg (as the variant's tile order gives them); the plan and tile width come
 * from variant % 4, persistence from variant / 16 (tstart holds n + 1
 * entries, the last the total tile count) */
template <class T>
int launch_axis_batch(const AArgs<T> *d, const long *tstart, int n, long nblocks, int len, int variant, hipStream_t s) {
    if (variant < 0 || variant >= axis_variants()) return MPC_EINVAL;
    switch (len) {
    case 16:  return axis_batch_v<T, 16>(d, tstart, n, nblocks, variant, s);
    case 32:  return axis_batch_v<T, 32>(d, tstart, n, nblocks, variant, s);
    case 64:  return axis_batch_v<T, 64>(d, tstart, n, nblocks, variant, s);
    case 128: return axis_batch_v<T, 128>(d, tstart, n, nblocks, variant, s);
    case 192: return axis_batch_v<T, 192>(d, tstart, n, nblocks, variant, s);
    case 256: return axis_batch_v<T, 256>(d, tstart, n, nblocks, variant, s);
    case 512: return axis_batch_v<T, 512>(d, tstart, n, nblocks, variant, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_axis_batch<double>(const AArgs<double> *, const long *, int, long, int, int, hipStream_t);
template int launch_axis_batch<float>(const AArgs<float> *, const long *, int, long, int, int, hipStream_t);

template <class T> int launch_axis(const AArgs<T> &a, int n, int outer, int variant, hipStream_t s) {
    if (variant < 0 || variant >= axis_variants()) return MPC_EINVAL;
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

} // namespace mpc
