/*
 * tline.hip -- the fused t-line sweep of the split-step propagator.
 *
 * k_tline: a workgroup owns whole t-lines (layout A[y][x][t], t fastest).
 * Per line: [t-FFT] x P(omega) -> t-IFFT of A and of -i omega A (spectral
 * d/dt) -> Kerr + self-steepening + plasma (ADK + avalanche, reverse-t affine
 * scan in FP64) -> A exp(dz A_nl) -> t-FFT [x Q -> t-IFFT].  Replaces
 * propagation.c:164-327 and plasma.c:118-246 (which the reference runs on
 * the host) and the t part of linearstep (propagation.c:11-30, 67-76).
 *
 * Built with machine-level loop-invariant code motion off: in the persistent
 * form the line loop would otherwise hoist every FP64 polynomial constant of
 * exp/log into a register pair for the whole kernel, which costs a wave per
 * SIMD of occupancy.
 */
#include "device.h"

namespace mpc {

/*
 * k_tline<T, G, PLASMA, PERSIST>: a workgroup owns G::LPB t-lines at a time.
 * With PERSIST the grid is what fits on the device at once and each
 * workgroup walks line groups blockIdx.x, + gridDim.x, ...; it loads the
 * twiddle powers and the Omega grid once, and loads the next group's field
 * while it works on the current one, so the HBM latency of a line is hidden
 * behind the previous line's FFTs.  Without PERSIST the grid has one
 * workgroup per line group and every table is read per line.
 */
/*
 * The arithmetic of one line (this thread's E elements in v, the line's LDS
 * at lraw): [t-FFT] [x P] -> nonlinear step -> t-FFT [x Q] [-> t-IFFT].
 * prefetch() runs once the input has been consumed, so a persistent caller
 * can start loading its next line there.  Shared by the per-field and the
 * batched sweeps, so both do the same arithmetic per line.
 */
template <class T, class G, bool PLASMA, bool PERSIST, class Fft, class Pre>
__device__ __forceinline__ void tline_line(const TArgs<T> &a, cx<T> (&v)[G::E], int tid, unsigned char *lraw,
                                            const T (&wk)[G::E], double &rmax, Fft &fft, Pre &&prefetch) {
    constexpr int E = G::E, TPL = G::TPL;
    using Fwd = std::integral_constant<int, -1>;
    using Inv = std::integral_constant<int, +1>;
    if (!a.in_freq) fft(v, Fwd{});
    if (a.P) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], gload(a.P, tid + TPL * m));
    }
    prefetch();

    if (a.do_nl) {
        cx<T> d[E];
        if (a.sst_on) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const T w = PERSIST ? wk[m] : gload(a.wfft, tid + TPL * m);
                d[m] = mk(w * v[m].im, -w * v[m].re);   /* -i omega A(omega) */
            }
            fft(d, Inv{});                             /* dA */
        }
        fft(v, Inv{});                                 /* A  */
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T a2 = v[m].re * v[m].re + v[m].im * v[m].im;
            if (a.sst_on) {
                /* S = gamma2/w0 (2 dA conj(A) + A conj(dA))  (propagation.c:232);
                 * A conj(dA) = conj(dA conj(A)), so S = gamma2/w0 (3 Re t + i Im t)
                 * with t = dA conj(A) */
                const cx<T> t = cmulc(d[m], v[m]);
                const T sre = a.sst_coef * ((T)3 * t.re);
                const T sim = a.sst_coef * t.im;
                d[m] = mk(-sre, a.gamma * a2 - sim);
            } else {
                d[m] = mk((T)0, a.gamma * a2);
            }
        }
        if constexpr (PLASMA) {
            const T half_rho = (T)(0.5 * a.pk.rho_nt), w0 = (T)a.pk.w0;
            plasma_line<T, G, true>(v, tid, reinterpret_cast<double *>(lraw), a.pk, rmax,
                                     [&](int m, T F, T tau) {
                                         /* (1 + i w0 tau_c) rho_nt F / 2  (propagation.c:289-292) */
                                         const T z = half_rho * F;
                                         d[m].re -= z;
                                         d[m].im -= w0 * tau * z;
                                     },
                                     fabs((double)a.dz));
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
        fft(v, Fwd{});
    }
    if (a.Q) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], gload(a.Q, tid + TPL * m));
    }
    if (a.out_time) fft(v, Inv{});
}

/* workgroup max of the ionisation fraction over its lines, into *slot (all
 * threads call it; the LDS scratch is free again on return) */
template <class G> __device__ __forceinline__ void tline_flush_rho(double rmax, double *slot) {
    __shared__ double red[32];
    double r = rmax;
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_down(r, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int i = 0; i < (G::THREADS + 63) / 64; ++i) mx = fmax(mx, red[i]);
        *slot = mx;
    }
    __syncthreads();
}

/* the work of one workgroup: line groups grp, grp + gstride, ... of the
 * field a describes (one group unless PERSIST); its plasma maximum goes to
 * a.rho_block[rho_slot] */
template <class T, class G, bool PLASMA, bool PERSIST>
__device__ __forceinline__ void tline_work(const TArgs<T> &a, long grp, long gstride, long rho_slot) {
    using C = TCfg<T, G, PLASMA>;
    using PL = typename G::PL;
    using CO = ChainOf<T, G::NT, C::E, typename PL::R>;
    constexpr int NT = G::NT;
    constexpr int E = C::E, TPL = C::TPL, LPB = C::LPB;
    __shared__ __align__(16) unsigned char smem[LPB * C::LINE_BYTES];
    const int lslot0 = threadIdx.x / TPL, tid0 = threadIdx.x % TPL;
    const long ngroups = a.nlines / LPB;
    if (grp >= ngroups) return;                       /* whole workgroup */
    int tid = tid0, lslot = lslot0;

    std::conditional_t<PERSIST, typename CO::type, NoChain> twc;
    T wk[E];
    if constexpr (PERSIST) {
        twc.load(a.tw, tid);
        if (a.do_nl && a.sst_on) {
#pragma unroll
            for (int m = 0; m < E; ++m) wk[m] = gload(a.wfft, tid + TPL * m);
        }
    }
    unsigned char *lraw;
    cx<T> *L;
    auto fft = [&](cx<T>(&x)[E], auto dir) {
        constexpr int DIR = decltype(dir)::value;
        if constexpr (PERSIST) CO::template fft<DIR, G::PAD, C::BOFF>(x, tid, L, 1, twc);
        else line_fft<T, NT, PL, DIR, G::PAD, C::BOFF>(x, tid, L, 1, a.tw);
    };

    cx<T> nxt[E];
    {
        const cx<T> *g0 = a.field + (grp * LPB + lslot) * NT;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = gload(g0, tid + TPL * m);
    }
    double rmax = 0.0;
    for (;;) {
        /* tid and lslot re-enter as opaque values each iteration, so the
         * per-thread index arithmetic is recomputed where it is used rather
         * than hoisted out of the loop and held in registers throughout */
        tid = tid0;
        lslot = lslot0;
        asm volatile("" : "+v"(tid), "+v"(lslot));
        if constexpr (PERSIST) opaque_chain(twc);
        lraw = smem + lslot * C::LINE_BYTES;
        L = reinterpret_cast<cx<T> *>(lraw);
        cx<T> *g = a.field + (grp * LPB + lslot) * NT;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        const long gnext = grp + gstride;
        tline_line<T, G, PLASMA, PERSIST>(a, v, tid, lraw, wk, rmax, fft, [&] {
            if constexpr (PERSIST) {
                if (gnext < ngroups) {
                    const cx<T> *gn = a.field + (gnext * LPB + lslot) * NT;
#pragma unroll
                    for (int m = 0; m < E; ++m) nxt[m] = gload(gn, tid + TPL * m);
                }
            }
        });
#pragma unroll
        for (int m = 0; m < E; ++m) gstore(g, tid + TPL * m, v[m]);
        if (!PERSIST || gnext >= ngroups) break;
        grp = gnext;
    }
    if constexpr (PLASMA) {
        if (a.rho_block && a.do_nl) tline_flush_rho<G>(rmax, a.rho_block + rho_slot);
    }
}

/* FP32: the persistent form is held to two waves per SIMD (256 VGPRs) and
 * the per-group form to three (168 VGPRs, which its plasma LDS allows too).
 * The persistent form fits; the per-group form spills (tens of VGPRs at 1024
 * and 2048 points, since it loads every stage twiddle), and the tuner keeps
 * whichever is faster */
template <class T, class G, bool PLASMA, bool PERSIST>
__global__ void __launch_bounds__((TCfg<T, G, PLASMA>::THREADS), sizeof(T) == 4 ? (PERSIST ? 2 : 3) : 1)
k_tline(TArgs<T> a) {
    tline_work<T, G, PLASMA, PERSIST>(a, blockIdx.x, gridDim.x, blockIdx.x);
}

/*
 * Batched form: one workgroup per line group of every field in d[0..n),
 * which share the t length and the plasma switch.  Workgroup b works on
 * field p, group b - gstart[p], for the largest p with gstart[p] <= b; the
 * per-line arithmetic is the per-field kernel's, so each field's result is
 * the same bits as a launch of k_tline<..., false> on it alone.
 */
template <class T, class G, bool PLASMA>
__global__ void __launch_bounds__((TCfg<T, G, PLASMA>::THREADS), sizeof(T) == 4 ? 3 : 1)
k_tline_batch(const TArgs<T> *__restrict__ d, const long *__restrict__ gstart, int n) {
    const long b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (gstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const TArgs<T> a = d[lo];
    const long g = b - gstart[lo];
    tline_work<T, G, PLASMA, false>(a, g, a.nlines, g);
}

/*
 * Persistent batched form: the grid is what fits on the device at once, and
 * workgroup w walks the concatenated line groups of every field in d[0..n)
 * (global group w, w + gridDim.x, ...; field p owns [gstart[p],
 * gstart[p + 1]), gstart[n] = total).  At each field boundary it reloads the
 * field's arguments, twiddle powers and Omega grid, and flushes its plasma
 * maximum for the field it leaves to slot j = its first group in that field,
 * so the slots [0, min(grid, groups)) of each field are written exactly
 * once.  The next group is prefetched across field boundaries too.  The
 * per-line arithmetic is the persistent per-field kernel's (tline_line).
 */
template <class T, class G, bool PLASMA>
__device__ __forceinline__ void tline_work_batch(const TArgs<T> *__restrict__ d, const long *__restrict__ gstart,
                                                 int n) {
    using C = TCfg<T, G, PLASMA>;
    using PL = typename G::PL;
    using CO = ChainOf<T, G::NT, C::E, typename PL::R>;
    constexpr int NT = G::NT;
    constexpr int E = C::E, TPL = C::TPL, LPB = C::LPB;
    __shared__ __align__(16) unsigned char smem[LPB * C::LINE_BYTES];
    const int lslot0 = threadIdx.x / TPL, tid0 = threadIdx.x % TPL;
    const long total = gstart[n];
    long g = blockIdx.x;
    if (g >= total) return;                           /* whole workgroup */
    int p = 0;
    {
        int hi = n - 1;
        while (p < hi) {
            const int mid = (p + hi + 1) / 2;
            if (gstart[mid] <= g) p = mid;
            else hi = mid - 1;
        }
    }
    TArgs<T> a = d[p];
    long first = g - gstart[p];
    int tid = tid0, lslot = lslot0;

    typename CO::type twc;
    T wk[E];
    auto load_field = [&] {
        twc.load(a.tw, tid);
        if (a.do_nl && a.sst_on) {
#pragma unroll
            for (int m = 0; m < E; ++m) wk[m] = gload(a.wfft, tid + TPL * m);
        }
    };
    load_field();
    unsigned char *lraw;
    cx<T> *L;
    auto fft = [&](cx<T>(&x)[E], auto dir) {
        constexpr int DIR = decltype(dir)::value;
        CO::template fft<DIR, G::PAD, C::BOFF>(x, tid, L, 1, twc);
    };

    cx<T> nxt[E];
    {
        const cx<T> *g0 = a.field + ((g - gstart[p]) * LPB + lslot) * NT;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = gload(g0, tid + TPL * m);
    }
    double rmax = 0.0;
    for (;;) {
        tid = tid0;
        lslot = lslot0;
        asm volatile("" : "+v"(tid), "+v"(lslot));
        opaque_chain(twc);
        lraw = smem + lslot * C::LINE_BYTES;
        L = reinterpret_cast<cx<T> *>(lraw);
        cx<T> *gp = a.field + ((g - gstart[p]) * LPB + lslot) * NT;
        cx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        const long gn = g + gridDim.x;
        int pn = p;
        while (pn + 1 < n && gstart[pn + 1] <= gn) ++pn;
        tline_line<T, G, PLASMA, true>(a, v, tid, lraw, wk, rmax, fft, [&] {
            if (gn < total) {
                const cx<T> *fn = pn == p ? a.field : d[pn].field;
                const cx<T> *src = fn + ((gn - gstart[pn]) * LPB + lslot) * NT;
#pragma unroll
                for (int m = 0; m < E; ++m) nxt[m] = gload(src, tid + TPL * m);
            }
        });
#pragma unroll
        for (int m = 0; m < E; ++m) gstore(gp, tid + TPL * m, v[m]);
        if (gn >= total || pn != p) {
            /* leaving field p */
            if constexpr (PLASMA) {
                if (a.rho_block && a.do_nl) tline_flush_rho<G>(rmax, a.rho_block + first);
            }
            rmax = 0.0;
            if (gn >= total) break;
            p = pn;
            a = d[p];
            first = gn - gstart[p];
            load_field();
        }
        g = gn;
    }
}

template <class T, class G, bool PLASMA>
__global__ void __launch_bounds__((TCfg<T, G, PLASMA>::THREADS), sizeof(T) == 4 ? 2 : 1)
k_tline_batch_p(const TArgs<T> *__restrict__ d, const long *__restrict__ gstart, int n) {
    tline_work_batch<T, G, PLASMA>(d, gstart, n);
}

bool tline_supported(int nt) { return nt == 256 || nt == 512 || nt == 1024 || nt == 2048 || nt == 4096; }
int tline_variants(void) { return 2; }

/* workgroups a per-field launch over ngroups line groups runs: all of them,
 * or for the persistent form what fits on the device at once */
template <class T, class G, bool PL, bool PERSIST> static long tline_grid_of(long ngroups) {
    using C = TCfg<T, G, PL>;
    long nb = ngroups;
    if (PERSIST) {
        static long resident = -1;              /* per instantiation: the same on every device of one model */
        if (resident < 0) resident = resident_blocks((const void *)k_tline<T, G, PL, PERSIST>, C::THREADS, 0);
        if (resident > 0 && resident < nb) nb = resident;
    }
    return nb;
}
template <class T, class G, bool PL, bool PERSIST>
static int tline_go(const TArgs<T> &a, hipStream_t s, int *nblocks) {
    using C = TCfg<T, G, PL>;
    if (a.nlines % C::LPB) return MPC_EINVAL;
    const long nb = tline_grid_of<T, G, PL, PERSIST>(a.nlines / C::LPB);
    if (nb <= 0 || nb > 0x7fffffffL) return MPC_EINVAL;
    hipLaunchKernelGGL((k_tline<T, G, PL, PERSIST>), dim3((unsigned)nb), dim3(C::THREADS), 0, s, a);
    if (nblocks) *nblocks = (int)nb;
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
/* variant 0: persistent workgroups with cached twiddles and prefetch;
 * 1: one workgroup per line group.  Both use 8 elements per thread (TPlan). */
template <class T, int NT> static int tline_v(const TArgs<T> &a, bool plasma, int variant, hipStream_t s, int *nb) {
    using GEO = TGeo<NT>;
    if (variant == 0)
        return plasma ? tline_go<T, GEO, true, true>(a, s, nb) : tline_go<T, GEO, false, true>(a, s, nb);
    return plasma ? tline_go<T, GEO, true, false>(a, s, nb) : tline_go<T, GEO, false, false>(a, s, nb);
}
/* grid of a batched launch over `total` line groups: one workgroup per group
 * (variant 1) or what fits on the device at once (variant 0, persistent) */
template <class T, int NT> static long tline_batch_grid_t(long total, bool plasma, int variant) {
    using GEO = TGeo<NT>;
    if (variant != 0) return total;
    static long res[2] = {-1, -1};              /* per instantiation and plasma switch */
    long &r = res[plasma ? 1 : 0];
    if (r < 0)
        r = plasma ? resident_blocks((const void *)k_tline_batch_p<T, GEO, true>, TCfg<T, GEO, true>::THREADS, 0)
                   : resident_blocks((const void *)k_tline_batch_p<T, GEO, false>, TCfg<T, GEO, false>::THREADS, 0);
    return (r > 0 && r < total) ? r : total;
}
template <class T, int NT>
static int tline_batch_go(const TArgs<T> *d, const long *gstart, int n, long total, bool plasma, int variant,
                          hipStream_t s) {
    using GEO = TGeo<NT>;
    if (n <= 0 || total <= 0 || total > 0x7fffffffL) return MPC_EINVAL;
    const long nb = tline_batch_grid_t<T, NT>(total, plasma, variant);
    if (variant == 0) {
        if (plasma)
            hipLaunchKernelGGL((k_tline_batch_p<T, GEO, true>), dim3((unsigned)nb), dim3(TCfg<T, GEO, true>::THREADS),
                               0, s, d, gstart, n);
        else
            hipLaunchKernelGGL((k_tline_batch_p<T, GEO, false>), dim3((unsigned)nb),
                               dim3(TCfg<T, GEO, false>::THREADS), 0, s, d, gstart, n);
    } else if (plasma) {
        hipLaunchKernelGGL((k_tline_batch<T, GEO, true>), dim3((unsigned)nb), dim3(TCfg<T, GEO, true>::THREADS),
                           0, s, d, gstart, n);
    } else {
        hipLaunchKernelGGL((k_tline_batch<T, GEO, false>), dim3((unsigned)nb), dim3(TCfg<T, GEO, false>::THREADS),
                           0, s, d, gstart, n);
    }
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T>
int launch_tline_batch(const TArgs<T> *d, const long *gstart, int n, long total, int nt, bool plasma, int variant,
                       hipStream_t s) {
    switch (nt) {
    case 256:  return tline_batch_go<T, 256>(d, gstart, n, total, plasma, variant, s);
    case 512:  return tline_batch_go<T, 512>(d, gstart, n, total, plasma, variant, s);
    case 1024: return tline_batch_go<T, 1024>(d, gstart, n, total, plasma, variant, s);
    case 2048: return tline_batch_go<T, 2048>(d, gstart, n, total, plasma, variant, s);
    case 4096: return tline_batch_go<T, 4096>(d, gstart, n, total, plasma, variant, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_tline_batch<double>(const TArgs<double> *, const long *, int, long, int, bool, int, hipStream_t);
template int launch_tline_batch<float>(const TArgs<float> *, const long *, int, long, int, bool, int, hipStream_t);

template <class T> long tline_batch_grid(long total, int nt, bool plasma, int variant) {
    switch (nt) {
    case 256:  return tline_batch_grid_t<T, 256>(total, plasma, variant);
    case 512:  return tline_batch_grid_t<T, 512>(total, plasma, variant);
    case 1024: return tline_batch_grid_t<T, 1024>(total, plasma, variant);
    case 2048: return tline_batch_grid_t<T, 2048>(total, plasma, variant);
    case 4096: return tline_batch_grid_t<T, 4096>(total, plasma, variant);
    default: return total;
    }
}
template long tline_batch_grid<double>(long, int, bool, int);
template long tline_batch_grid<float>(long, int, bool, int);

template <class T> int launch_tline(const TArgs<T> &a, int nt, bool plasma, int variant, hipStream_t s, int *nblocks) {
    if (variant < 0 || variant >= tline_variants()) return MPC_EINVAL;
    switch (nt) {
    case 256:  return tline_v<T, 256>(a, plasma, variant, s, nblocks);
    case 512:  return tline_v<T, 512>(a, plasma, variant, s, nblocks);
    case 1024: return tline_v<T, 1024>(a, plasma, variant, s, nblocks);
    case 2048: return tline_v<T, 2048>(a, plasma, variant, s, nblocks);
    case 4096: return tline_v<T, 4096>(a, plasma, variant, s, nblocks);
    default: return MPC_EUNSUPP;
    }
}
template int launch_tline<double>(const TArgs<double> &, int, bool, int, hipStream_t, int *);
template int launch_tline<float>(const TArgs<float> &, int, bool, int, hipStream_t, int *);

template <class T, int NT> static long tline_grid_t(long nlines, bool plasma, int variant) {
    using GEO = TGeo<NT>;
    const long ng = nlines / GEO::LPB;
    if (variant == 0) return plasma ? tline_grid_of<T, GEO, true, true>(ng) : tline_grid_of<T, GEO, false, true>(ng);
    return ng;
}
template <class T> long tline_grid(long nlines, int nt, bool plasma, int variant) {
    switch (nt) {
    case 256:  return tline_grid_t<T, 256>(nlines, plasma, variant);
    case 512:  return tline_grid_t<T, 512>(nlines, plasma, variant);
    case 1024: return tline_grid_t<T, 1024>(nlines, plasma, variant);
    case 2048: return tline_grid_t<T, 2048>(nlines, plasma, variant);
    case 4096: return tline_grid_t<T, 4096>(nlines, plasma, variant);
    default: return 0;
    }
}
template long tline_grid<double>(long, int, bool, int);
template long tline_grid<float>(long, int, bool, int);

/* t-line kernel block geometry (lines per workgroup; every variant's) */
int tline_lines_per_block(int nt, bool plasma, int prec) {
    (void)plasma; (void)prec;
    switch (nt) {
    case 256: return TGeo<256>::LPB;
    case 512: return TGeo<512>::LPB;
    case 1024: return TGeo<1024>::LPB;
    case 2048: return TGeo<2048>::LPB;
    case 4096: return TGeo<4096>::LPB;
    default: return 1;
    }
}

} // namespace mpc
