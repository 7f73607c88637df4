/*
 * kernels.hip -- auxiliary kernels of the split-step propagator (the fused
 * t-line sweep is in tline.hip, the x/y sweeps in axis.hip): layout
 * transposes, diagnostics reductions, setStep's error norms (the legacy
 * entry point; the adaptive driver fuses them into its last sweep), fixed-order
 * sums, the guard's energy, and the plasma/ADK kernels behind the legacy
 * ioniz()/ADK() entry points.
 */
#include "device.h"

namespace mpc {

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

/* resident field <-> two float planes [re | im] of n points (same point order) */
template <class T>
__global__ void k_planar(cx<T> *__restrict__ f, float *__restrict__ planes, long n, double s, int to_planes) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (to_planes) {
            const cx<T> v = f[i];
            planes[i] = (float)((double)v.re * s);
            planes[n + i] = (float)((double)v.im * s);
        } else {
            f[i] = mk((T)((double)planes[i] * s), (T)((double)planes[n + i] * s));
        }
    }
}
template <class T> int launch_planar(cx<T> *f, float *planes, long n, double s, int to_planes, hipStream_t st) {
    hipLaunchKernelGGL(k_planar<T>, dim3(1024), dim3(256), 0, st, f, planes, n, s, to_planes);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_planar<double>(cx<double> *, float *, long, double, int, hipStream_t);
template int launch_planar<float>(cx<float> *, float *, long, double, int, hipStream_t);
template int launch_scale<float>(cx<float> *, long, float, hipStream_t);

/* ---------------------------------------------------------------------- */
/* diagnostics: per-line power, and per-block partial temp[t] / spec[t]    */
/* ---------------------------------------------------------------------- */
/* SPEC: the field holds the unnormalised t-FFT of each line (the t-spectral
 * state); spec is then read off directly and temp and the line power come
 * from the inverse transform, scaled by 1/nt^2 */
template <class T, int NT, bool SPEC>
__global__ void __launch_bounds__(256) k_diag_lines(const cx<T> *f, const cx<T> *tw, long nlines,
                                                    double *line_power, double *temp_part, double *spec_part) {
    using G = TGeo<NT>;
    constexpr int E = G::E, TPL = G::TPL, LPB = G::LPB;
    /* with one line per block each thread owns its t values and writes its
     * partial sums itself; with several, the slots are combined in LDS */
    constexpr bool SLOTS = LPB > 1;
    __shared__ __align__(16) cx<T> lds[LPB * G::FFT_ELEMS];
    __shared__ double accT[SLOTS ? NT : 1], accS[SLOTS ? NT : 1];
    __shared__ double lp[LPB][TPL / 64 + 1];
    const int lslot = threadIdx.x / TPL, tid = threadIdx.x % TPL;
    if constexpr (SLOTS)
        for (int i = threadIdx.x; i < NT; i += blockDim.x) { accT[i] = 0.0; accS[i] = 0.0; }
    /* the zeroing threads are not the ones that add each slot below, and a
     * block with no line group reaches no barrier inside the loop */
    __syncthreads();
    double at[E], as[E];
#pragma unroll
    for (int m = 0; m < E; ++m) { at[m] = 0.0; as[m] = 0.0; }
    /* a fixed set of line groups per block: deterministic order; the next
     * group's line is loaded while this one is transformed */
    const long ngroups = nlines / LPB;
    cx<T> nxt[E];
    if (blockIdx.x < ngroups) {
        const cx<T> *g0 = f + ((long)blockIdx.x * LPB + lslot) * NT;
#pragma unroll
        for (int m = 0; m < E; ++m) nxt[m] = g0[tid + TPL * m];
    }
    for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long line = grp * LPB + lslot;
        /* the table pointer re-enters as an opaque value each iteration, so
         * the stage twiddles load where they are used instead of being
         * hoisted out of the loop and held in registers throughout */
        const cx<T> *twi = tw;
        asm volatile("" : "+s"(twi));
        cx<T> v[E];
        double pw = 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        if (grp + gridDim.x < ngroups) {
            const cx<T> *gn = f + ((grp + gridDim.x) * LPB + lslot) * NT;
#pragma unroll
            for (int m = 0; m < E; ++m) nxt[m] = gn[tid + TPL * m];
        }
        if constexpr (SPEC) {
#pragma unroll
            for (int m = 0; m < E; ++m) as[m] += (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
            line_fft<T, NT, TPlan<NT>, +1, G::PAD>(v, tid, lds + lslot * G::FFT_ELEMS, 1, twi);
            constexpr double inv2 = 1.0 / ((double)NT * (double)NT);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const double p = ((double)v[m].re * v[m].re + (double)v[m].im * v[m].im) * inv2;
                at[m] += p;
                pw += p;
            }
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const double p = (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
                at[m] += p;
                pw += p;
            }
            line_fft<T, NT, TPlan<NT>, -1, G::PAD>(v, tid, lds + lslot * G::FFT_ELEMS, 1, twi);
#pragma unroll
            for (int m = 0; m < E; ++m) as[m] += (double)v[m].re * v[m].re + (double)v[m].im * v[m].im;
        }
        /* line power: reduce over the TPL threads of the line */
        constexpr int W = TPL < 64 ? TPL : 64;
        for (int o = W / 2; o > 0; o >>= 1) pw += __shfl_down(pw, o, W);
        if (tid % 64 == 0) lp[lslot][tid / 64] = pw;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int i = 0; i < (TPL + 63) / 64; ++i) sum += lp[lslot][i];
            line_power[line] = sum;
        }
        __syncthreads();
    }
    if constexpr (!SLOTS) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            temp_part[(long)blockIdx.x * NT + tid + TPL * m] = at[m];
            spec_part[(long)blockIdx.x * NT + tid + TPL * m] = as[m];
        }
        return;
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
static int diag_go(const cx<T> *f, const cx<T> *tw, long nlines, double *lp, double *tp, double *sp, int nb,
                   bool spectral, hipStream_t s) {
    using G = TGeo<NT>;
    if (nlines % G::LPB) return MPC_EINVAL;
    if (spectral)
        hipLaunchKernelGGL((k_diag_lines<T, NT, true>), dim3(nb), dim3(G::THREADS), 0, s, f, tw, nlines, lp, tp, sp);
    else
        hipLaunchKernelGGL((k_diag_lines<T, NT, false>), dim3(nb), dim3(G::THREADS), 0, s, f, tw, nlines, lp, tp, sp);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template <class T> int launch_diag_lines(const cx<T> *f, const cx<T> *tw, int nt, long nlines, double *lp,
                                         double *tp, double *sp, int nb, bool spectral, hipStream_t s) {
    switch (nt) {
    case 256: return diag_go<T, 256>(f, tw, nlines, lp, tp, sp, nb, spectral, s);
    case 512: return diag_go<T, 512>(f, tw, nlines, lp, tp, sp, nb, spectral, s);
    case 1024: return diag_go<T, 1024>(f, tw, nlines, lp, tp, sp, nb, spectral, s);
    case 2048: return diag_go<T, 2048>(f, tw, nlines, lp, tp, sp, nb, spectral, s);
    case 4096: return diag_go<T, 4096>(f, tw, nlines, lp, tp, sp, nb, spectral, s);
    default: return MPC_EUNSUPP;
    }
}
template int launch_diag_lines<double>(const cx<double> *, const cx<double> *, int, long, double *, double *, double *,
                                       int, bool, hipStream_t);
template int launch_diag_lines<float>(const cx<float> *, const cx<float> *, int, long, double *, double *, double *,
                                      int, bool, hipStream_t);

/* out[c] = sum_r part[r][c], in a fixed order (deterministic): RC_PARTS
 * threads per column sum rows q, q + RC_PARTS, ... in turn, then the parts
 * are added in q order; consecutive threads read consecutive columns */
constexpr int RC_COLS = 64, RC_PARTS = 16;
__global__ void __launch_bounds__(RC_COLS * RC_PARTS) k_reduce_cols(const double *part, int nrows, int ncols,
                                                                    double *out) {
    __shared__ double ps[RC_PARTS][RC_COLS];
    const int cl = threadIdx.x % RC_COLS, q = threadIdx.x / RC_COLS;
    const long c = (long)blockIdx.x * RC_COLS + cl;
    double s = 0.0;
    if (c < ncols)
        for (int r = q; r < nrows; r += RC_PARTS) s += part[(long)r * ncols + c];
    ps[q][cl] = s;
    __syncthreads();
    if (q == 0 && c < ncols) {
        double t = 0.0;
        for (int k = 0; k < RC_PARTS; ++k) t += ps[k][cl];
        out[c] = t;
    }
}
int launch_reduce_cols(const double *part, int nrows, int ncols, double *out, hipStream_t s) {
    if (nrows <= 0 || ncols <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_reduce_cols, dim3((ncols + RC_COLS - 1) / RC_COLS), dim3(RC_COLS * RC_PARTS), 0, s, part,
                       nrows, ncols, out);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

/* ---------------------------------------------------------------------- */
/* setStep's error norms (propagation.c:625-631), fixed-order sums, energy  */
/*                                                                         */
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

/* one block: sum of n (a, b) pairs in a fixed order */
__global__ void __launch_bounds__(256) k_sum_pairs(const double *part, long n, double *out) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (long i = threadIdx.x; i < n; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}
int launch_sum_pairs(const double *part, long n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, s, part, n, out);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

/* per-block sum |f|^2 in a fixed order (the solver guard's energy) */
template <class T>
__global__ void __launch_bounds__(256) k_energy_partials(const cx<T> *f, long n, double *part) {
    __shared__ double s1[256];
    double a = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        a += (double)f[i].re * f[i].re + (double)f[i].im * f[i].im;
    s1[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s1[threadIdx.x] += s1[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s1[0];
}
template <class T> int launch_energy_partials(const cx<T> *f, long n, double *part, int nb, hipStream_t s) {
    hipLaunchKernelGGL((k_energy_partials<T>), dim3(nb), dim3(256), 0, s, f, n, part);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
template int launch_energy_partials<double>(const cx<double> *, long, double *, int, hipStream_t);
template int launch_energy_partials<float>(const cx<float> *, long, double *, int, hipStream_t);


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
                       [&](int m, T Fv, T tv) {
                           out[line * NT + tid + TPL * m] = mk((double)tv, (double)Fv);   /* (tau_c, F) */
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
    case 256: return ioniz_go<T, 256>(f, nlines, pk, out, s);
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
