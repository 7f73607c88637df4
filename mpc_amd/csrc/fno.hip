/*
 * fno.hip -- kernels of the Fourier-neural-operator pass surrogate.
 *
 * The reference has no learned part.  The surrogate maps the field at one
 * mirror plane to the field after a pass.  On the matrix cores: the
 * truncated DFTs along t, y and x, their inverses and transposes, and the
 * weight gradients of the 1x1 channel maps, as batched GEMMs (k_gemm_bf16):
 * operands are rounded to BF16 as they are staged into LDS, products
 * accumulate in FP32 (v_mfma_f32_16x16x16_bf16).  The pointwise 1x1 channel
 * maps with bias, residual and GELU fused: FP32 on the matrix cores
 * (k_pointwise_mf, v_mfma_f32_32x32x2f32) at 16 to 64 input channels, on the
 * vector units otherwise (k_pointwise).  On the vector units, in FP32: the
 * per-mode complex channel mixing and its backward (k_mix, k_mix_bwd),
 * complex recombination after the real-block DFT products, bias reductions,
 * the loss and the Adam update.
 */
#include "internal.h"
#include "../../include/mpc_fno.h"
#include <cmath>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (finite inputs) */
__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

constexpr int TM = 64, TN = 64, TK = 32;     /* workgroup tile */
constexpr int LDK = TK + 4;                  /* LDS row length in bf16 (8-byte aligned rows) */

/*
 * C = alpha * A B + beta * C over a two-level batch, with arbitrary element
 * strides (A[i][k] = A[i*rsA + k*csA], etc.).  With ksplit > 1 the workgroup
 * z index also selects a K range and raw partial sums go to W[split][batch]
 * [M][N] for k_gemm_reduce.  256 threads: four waves, each a 32 x 32 block of
 * 2 x 2 MFMA 16 x 16 tiles.
 */
__global__ void __launch_bounds__(256) k_gemm_bf16(mpc_gemm g) {
    __shared__ __align__(16) unsigned short As[TM][LDK];
    __shared__ __align__(16) unsigned short Bs[TN][LDK];   /* B^T: [n][k] */
    const int nbat = g.nb1 * g.nb2;
    const int z = blockIdx.z;
    const int bat = z % nbat, split = z / nbat;
    const int b1 = bat / g.nb2, b2 = bat % g.nb2;
    const float *A = g.A + b1 * g.sA1 + b2 * g.sA2;
    const float *B = g.B + b1 * g.sB1 + b2 * g.sB2;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + TK - 1) / TK * TK;
    const int kbeg = split * kchunk;
    const int kend = min(g.K, kbeg + kchunk);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    /* coalesced staging: threads run along whichever operand index has unit stride */
    const bool a_krow = g.csA == 1 || g.rsA != 1;
    const bool b_ncol = g.csB == 1 || g.rsB != 1;
    f32x4 acc[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[p][q] = f32x4{0.f, 0.f, 0.f, 0.f};

    /* software pipeline: the next k tile's global loads are in flight while
     * the matrix cores work on the current one from LDS */
    float ra[8], rb[8];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int ii, kk;
            if (a_krow) { kk = t % TK; ii = t / TK + 8 * r; }
            else        { ii = t % TM; kk = t / TM + 4 * r; }
            const int gi = i0 + ii, gk = k0 + kk;
            ra[r] = (gi < g.M && gk < kend) ? A[(long)gi * g.rsA + (long)gk * g.csA] : 0.f;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int jj, kk;
            if (b_ncol) { jj = t % TN; kk = t / TN + 4 * r; }
            else        { kk = t % TK; jj = t / TK + 8 * r; }
            const int gj = j0 + jj, gk = k0 + kk;
            rb[r] = (gj < g.N && gk < kend) ? B[(long)gk * g.rsB + (long)gj * g.csB] : 0.f;
        }
    };
    if (kbeg < kend) load_tile(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int ii, kk;
            if (a_krow) { kk = t % TK; ii = t / TK + 8 * r; }
            else        { ii = t % TM; kk = t / TM + 4 * r; }
            As[ii][kk] = to_bf16(ra[r]);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int jj, kk;
            if (b_ncol) { jj = t % TN; kk = t / TN + 4 * r; }
            else        { kk = t % TK; jj = t / TK + 8 * r; }
            Bs[jj][kk] = to_bf16(rb[r]);
        }
        __syncthreads();
        if (k0 + TK < kend) load_tile(k0 + TK);
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            const int kq = ks + 4 * (lane >> 4);
            bf16x4 a[2], b[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) a[p] = *reinterpret_cast<const bf16x4 *>(&As[wi + 16 * p + (lane & 15)][kq]);
#pragma unroll
            for (int q = 0; q < 2; ++q) b[q] = *reinterpret_cast<const bf16x4 *>(&Bs[wj + 16 * q + (lane & 15)][kq]);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[p], b[q], acc[p][q], 0, 0, 0);
        }
        __syncthreads();
    }
    /* D[4*(lane/16) + e][lane%16] of each 16 x 16 tile */
    float *C = g.C + b1 * g.sC1 + b2 * g.sC2;
    float *Wp = g.ksplit > 1 ? g.W + ((long)split * nbat + bat) * (long)g.M * g.N : nullptr;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int gi = i0 + wi + 16 * p + 4 * (lane >> 4) + e;
                const int gj = j0 + wj + 16 * q + (lane & 15);
                if (gi >= g.M || gj >= g.N) continue;
                const float v = acc[p][q][e];
                if (Wp) {
                    Wp[(long)gi * g.N + gj] = v;
                } else {
                    float *c = C + (long)gi * g.rsC + (long)gj * g.csC;
                    *c = g.beta == 0.f ? g.alpha * v : g.alpha * v + g.beta * *c;
                }
            }
}

/* C = alpha * sum over splits of W + beta * C */
/* split-K reduction: workgroup of 64 outputs x 16 partial-sum lanes; lane r
 * adds splits r, r + 16, ... and the 16 lane sums are added in lane order */
constexpr int RD_OUT = 64, RD_SUB = 16;
__global__ void __launch_bounds__(RD_OUT * RD_SUB) k_gemm_reduce(mpc_gemm g) {
    __shared__ float part[RD_SUB][RD_OUT];
    const int nbat = g.nb1 * g.nb2;
    const long mn = (long)g.M * g.N;
    const long total = mn * nbat;
    const int o = threadIdx.x % RD_OUT, r = threadIdx.x / RD_OUT;
    const long x = blockIdx.x * (long)RD_OUT + o;
    float s = 0.f;
    if (x < total) {
        const int bat = (int)(x / mn);
        const long ij = x % mn;
        for (int sp = r; sp < g.ksplit; sp += RD_SUB) s += g.W[((long)sp * nbat + bat) * mn + ij];
    }
    part[r][o] = s;
    __syncthreads();
    if (r == 0 && x < total) {
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < RD_SUB; ++q) t += part[q][o];
        const int bat = (int)(x / mn);
        const long ij = x % mn;
        const int i = (int)(ij / g.N), j = (int)(ij % g.N);
        const int b1 = bat / g.nb2, b2 = bat % g.nb2;
        float *c = g.C + b1 * g.sC1 + b2 * g.sC2 + (long)i * g.rsC + (long)j * g.csC;
        *c = g.beta == 0.f ? g.alpha * t : g.alpha * t + g.beta * *c;
    }
}

/*
 * Complex recombination of a real-block DFT product.  P holds, per slice,
 * 2m rows: rows [0, m) are cos-weighted sums, rows [m, 2m) sin-weighted, and
 * each row nb column blocks of [re(0..h) | im(0..h)].  With sign sg (-1:
 * forward e^{-i theta}, +1: inverse) the output V (m rows) is
 *   Vr = Pc_r - sg Ps_i,   Vi = Pc_i + sg Ps_r.
 * bwd: the transpose map, dV -> dP.
 */
__global__ void k_combine(const float *__restrict__ in, float *__restrict__ out, long nslice, int m, int nb, int h,
                          float sg, int bwd) {
    const long per = (long)m * nb * h;
    const long total = per * nslice;
    const int w = 2 * h;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x) {
        const long s = x / per;
        long r = x % per;
        const int k = (int)(r % h); r /= h;
        const int q = (int)(r % nb);
        const int row = (int)(r / nb);
        const long rowlen = (long)nb * w;
        if (!bwd) {
            const float *P = in + s * 2 * m * rowlen;
            const float *pc = P + row * rowlen + q * w, *ps = P + (long)(m + row) * rowlen + q * w;
            float *V = out + s * m * rowlen + row * rowlen + q * w;
            V[k] = pc[k] - sg * ps[h + k];
            V[h + k] = pc[h + k] + sg * ps[k];
        } else {
            const float *dV = in + s * m * rowlen + row * rowlen + q * w;
            float *dP = out + s * 2 * m * rowlen;
            float *pc = dP + row * rowlen + q * w, *ps = dP + (long)(m + row) * rowlen + q * w;
            pc[k] = dV[k];
            pc[h + k] = dV[h + k];
            ps[h + k] = -sg * dV[k];
            ps[k] = sg * dV[h + k];
        }
    }
}

__device__ __forceinline__ float gelu(float x) {
    return 0.5f * x * (1.f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

/* z[b][c][p] += bias[c] (bias may be null); act: out = GELU(z) (or z) */
__global__ void k_bias_act(float *__restrict__ z, const float *__restrict__ bias, float *__restrict__ out, long nb,
                           int nc, long np, int act) {
    const long total = nb * nc * np;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x) {
        const int c = (int)((x / np) % nc);
        const float v = z[x] + (bias ? bias[c] : 0.f);
        z[x] = v;
        out[x] = act ? gelu(v) : v;
    }
}
/* dz = dout * GELU'(z) (act) or dout */
__global__ void k_act_bwd(const float *__restrict__ z, const float *__restrict__ dout, float *__restrict__ dz, long n,
                          int act) {
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x)
        dz[x] = act ? dout[x] * gelu_d(z[x]) : dout[x];
}
/*
 * Pointwise channel map (a 1x1 convolution) with its epilogue fused:
 *   z[b][o][p] = sum_i w[o*ldo + i*ldi] x[b][i][p] + beta zin[b][o][p] + bias[o]
 *   out[b][o][p] = GELU(z) (act) or z;  zs, if given, receives z.
 * FP32 on the vector ALUs, for the widths the matrix-core form below does
 * not take (cin < 16 or cout not a multiple of 32).  One thread per point
 * (coalesced along p); the weights are wave-uniform (scalar loads).
 * out may alias zin or zs (each element is read before it is written, by the
 * same thread).
 */
/* a float in the constant address space: loads through such a pointer are
 * known not to be clobbered by the kernel's stores, so uniform ones become
 * scalar loads */
typedef __attribute__((address_space(4))) const float cfloat;

template <int CIN, int OC, bool UNIT>
__global__ void __launch_bounds__(256) k_pointwise(mpc_pw a, const float *__restrict__ w,
                                                   const float *__restrict__ bias) {
    /* w and bias are read-only and wave-uniform: they load through the
     * scalar unit into SGPRs, the FMAs' scalar operand (no LDS traffic) */
    /* no early exit: the tail lanes of the last workgroup load a valid
     * point and skip their stores, so the control flow stays uniform and the
     * weight loads stay scalar */
    const long p0 = blockIdx.x * 256L + threadIdx.x;
    const bool live = p0 < a.np;
    const long p = live ? p0 : a.np - 1;
    const int b = blockIdx.y;
    float xr[CIN];
    const float *xb = a.x + b * a.x_bs + p;
#pragma unroll
    for (int i = 0; i < CIN; ++i) xr[i] = xb[(long)i * a.np];
    const long ob = (long)b * a.cout * a.np + p;
    const float *zb = a.beta != 0.f ? a.zin + b * a.zin_bs + p : nullptr;
    /* outputs in chunks of OC (cout is a multiple of OC): the next chunk's
     * residual loads are issued before this chunk's stores */
    float zc[OC], zn[OC];
#pragma unroll
    for (int j = 0; j < OC; ++j) zc[j] = zb ? zb[(long)j * a.np] : 0.f;
    for (int o0 = 0; o0 < a.cout; o0 += OC) {
        float acc[OC];
#pragma unroll
        for (int j = 0; j < OC; ++j) {
            const cfloat *wo = (const cfloat *)w + (o0 + j) * a.ldo;
            float s = 0.f;
#pragma unroll
            for (int i = 0; i < CIN; ++i) s = fmaf(wo[UNIT ? i : i * a.ldi], xr[i], s);
            acc[j] = s;
        }
        const bool more = zb && o0 + OC < a.cout;
#pragma unroll
        for (int j = 0; j < OC; ++j) zn[j] = more ? zb[(long)(o0 + OC + j) * a.np] : 0.f;
#pragma unroll
        for (int j = 0; j < OC; ++j) {
            const int o = o0 + j;
            float z = acc[j] + a.beta * zc[j];
            if (bias) z += bias[o];
            const float y = a.act ? gelu(z) : z;
            if (live) {
                if (a.zs) a.zs[ob + (long)o * a.np] = z;
                a.out[ob + (long)o * a.np] = y;
            }
            zc[j] = zn[j];
        }
    }
}

/* outputs per chunk: the chunk's CIN x OC weights sit in SGPRs at once, so
 * CIN x OC stays within 64 for unit-stride rows (wide scalar loads) and 16
 * for strided ones (a scalar load each); beyond that the compiler parks
 * SGPRs in VGPR lanes, two extra instructions per weight */
template <int CIN, bool UNIT> static void launch_pointwise_u(const mpc_pw *a, dim3 grid, hipStream_t s) {
    constexpr int LIM = UNIT ? 64 : 16;
    constexpr bool C8 = CIN * 8 <= LIM, C2 = CIN * 2 <= LIM;
    if (C8 && a->cout % 8 == 0)
        hipLaunchKernelGGL((k_pointwise<CIN, C8 ? 8 : 1, UNIT>), grid, dim3(256), 0, s, *a, a->w, a->bias);
    else if (C2 && a->cout % 2 == 0)
        hipLaunchKernelGGL((k_pointwise<CIN, C2 ? 2 : 1, UNIT>), grid, dim3(256), 0, s, *a, a->w, a->bias);
    else hipLaunchKernelGGL((k_pointwise<CIN, 1, UNIT>), grid, dim3(256), 0, s, *a, a->w, a->bias);
}
/* unit input stride (w[o][i] rows): the weight loads merge into wide scalar loads */
template <int CIN> static void launch_pointwise(const mpc_pw *a, dim3 grid, hipStream_t s) {
    if (a->ldi == 1) launch_pointwise_u<CIN, true>(a, grid, s);
    else launch_pointwise_u<CIN, false>(a, grid, s);
}

/*
 * The same map on the FP32 matrix cores, for cin in {16, 32, 64} and cout a
 * multiple of 32.  At these widths the vector ALUs spend cin FMAs per output
 * and run behind HBM; here a wave forms a tile of 32 outputs x 32 points as
 * cin/2 products v_mfma_f32_32x32x2f32 (K = 2 each), and the vector ALUs
 * are left the epilogue (residual, bias, GELU).
 * Operand layout of v_mfma_f32_32x32x2f32 (lane l, 64 lanes):
 *   A (32 x 2): lane l supplies A[l % 32][l / 32]  -> w[o = l % 32][i = l / 32]
 *   B (2 x 32): lane l supplies B[l / 32][l % 32]  -> x[i = l / 32][p = l % 32]
 *   D (32 x 32): register r of lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
 * so each x load reads two channel rows of 32 consecutive points, and each
 * epilogue register stores two output rows of 32 consecutive points.
 * A wave keeps its OT x 32 outputs' weights in registers and walks point tiles
 * grid-stride; blockIdx.z selects the group of OT x 32 outputs, blockIdx.y
 * the sample.  The tile loop is uniform per wave (MFMA runs on all 64 lanes):
 * the lanes past the last point load the last point and skip their stores.
 * out may alias zin or zs: each element is read and written by one lane.
 */
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int PW_MF_WAVES = 4;

template <int CIN, int OT>
__global__ void __launch_bounds__(64 * PW_MF_WAVES) k_pointwise_mf(mpc_pw a, const float *__restrict__ w,
                                                                   const float *__restrict__ bias) {
    constexpr int KS = CIN / 2;
    const int lane = threadIdx.x & 63, col = lane & 31, half = lane >> 5;
    const int b = blockIdx.y, o_base = blockIdx.z * (OT * 32);
    float av[OT][KS], bv[OT][16];
#pragma unroll
    for (int t = 0; t < OT; ++t) {
#pragma unroll
        for (int s = 0; s < KS; ++s)
            av[t][s] = w[(long)(o_base + t * 32 + col) * a.ldo + (long)(2 * s + half) * a.ldi];
#pragma unroll
        for (int r = 0; r < 16; ++r) bv[t][r] = bias ? bias[o_base + t * 32 + 8 * (r / 4) + 4 * half + r % 4] : 0.f;
    }
    /* addresses: a wave-uniform row pointer plus a 32-bit lane offset (the
     * half-wave's row and the point), so each load and store takes one VGPR
     * for its address */
    const float *xb = a.x + b * a.x_bs;
    const float *zb = a.beta != 0.f ? a.zin + b * a.zin_bs : nullptr;
    float *ob = a.out + (long)b * a.cout * a.np;
    float *sb = a.zs ? a.zs + (long)b * a.cout * a.np : nullptr;
    const long ntile = (a.np + 31) / 32;
    for (long tile = blockIdx.x * (long)PW_MF_WAVES + (threadIdx.x >> 6); tile < ntile;
         tile += (long)gridDim.x * PW_MF_WAVES) {
        const long p0 = tile * 32 + col;
        const bool live = p0 < a.np;
        const unsigned p = (unsigned)(live ? p0 : a.np - 1);
        const unsigned xo = (unsigned)(half * a.np) + p, zo = (unsigned)(4 * half * a.np) + p;
        float xv[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) xv[s] = (xb + 2 * s * a.np)[xo];
        /* the tile's residual loads go out with its x loads, ahead of every
         * store (out may alias zin, so the compiler keeps loads behind
         * earlier stores as written) */
        float zv[OT][16];
#pragma unroll
        for (int t = 0; t < OT; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                zv[t][r] = zb ? (zb + (long)(o_base + t * 32 + 8 * (r / 4) + r % 4) * a.np)[zo] : 0.f;
#pragma unroll
        for (int t = 0; t < OT; ++t) {
            f32x16 acc = {};
#pragma unroll
            for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t][s], xv[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long row = (long)(o_base + t * 32 + 8 * (r / 4) + r % 4) * a.np;
                float z = acc[r] + a.beta * zv[t][r];
                if (bias) z += bv[t][r];
                if (live) {
                    if (sb) (sb + row)[zo] = z;
                    (ob + row)[zo] = a.act ? gelu(z) : z;
                }
            }
        }
    }
}

static bool pointwise_mf_fits(const mpc_pw *a) {
    /* the lane offsets (below 5 np) are 32-bit */
    return (a->cin == 16 || a->cin == 32 || a->cin == 64) && a->cout % 32 == 0 && a->cout / 32 <= 65535 &&
           a->np <= (1L << 32) / 5;
}
/* workgroups of k resident on the current device at once (0: unknown) */
static long pw_resident(const void *k) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 64 * PW_MF_WAVES, 0) != hipSuccess)
        return 0;
    return (long)ncu * per;
}
template <int CIN, int OT> static void launch_pointwise_mf_t(const mpc_pw *a, hipStream_t s) {
    static const long resident = pw_resident((const void *)k_pointwise_mf<CIN, OT>);
    const long ntile = (a->np + 31) / 32, gz = a->cout / (32 * OT);
    /* one resident wave set walks the tiles: no second, partly filled round
     * of workgroups, and the weights load once per wave */
    long gx = (ntile + PW_MF_WAVES - 1) / PW_MF_WAVES;
    if (resident > 0) gx = std::min(gx, std::max(1L, resident / (a->nb * gz)));
    hipLaunchKernelGGL((k_pointwise_mf<CIN, OT>), dim3((unsigned)gx, a->nb, (unsigned)gz), dim3(64 * PW_MF_WAVES), 0,
                       s, *a, a->w, a->bias);
}
template <int CIN> static void launch_pointwise_mf(const mpc_pw *a, hipStream_t s) {
    if (a->cout % 64 == 0) launch_pointwise_mf_t<CIN, 2>(a, s);
    else launch_pointwise_mf_t<CIN, 1>(a, s);
}

/*
 * Per-mode complex channel mixing of the spectral convolution, FP32.
 * V, U: [b][channel][kxy][2 mt] (re at kt, im at mt + kt); R = Rr + i Ri:
 * [i][o][kxy][mt].  Forward U[b][o] = sum_i V[b][i] R[i][o] (per mode).
 * One thread per (mode, o); the weights stream once (coalesced along the
 * mode), V is re-read per o from cache.
 */
constexpr int MIX_BMAX = 8;
__global__ void __launch_bounds__(256) k_mix(const float *__restrict__ V, const float *__restrict__ Rr,
                                             const float *__restrict__ Ri, float *__restrict__ U, int nb, int w,
                                             int nxy, int mt) {
    const int nm = nxy * mt;
    const int m = blockIdx.x * 256 + threadIdx.x;
    const int o = blockIdx.y;
    if (m >= nm) return;
    const int kxy = m / mt, kt = m % mt;
    const long vm = (long)kxy * 2 * mt + kt;          /* re offset of the mode in a channel's block */
    const long cs = (long)nxy * 2 * mt;               /* channel stride in V, U */
    float ur[MIX_BMAX], ui[MIX_BMAX];
#pragma unroll
    for (int b = 0; b < MIX_BMAX; ++b) ur[b] = ui[b] = 0.f;
    for (int i = 0; i < w; ++i) {
        const long ri = ((long)i * w + o) * nm + m;
        const float rr = Rr[ri], rim = Ri[ri];
#pragma unroll
        for (int b = 0; b < MIX_BMAX; ++b) {
            if (b < nb) {
                const float *v = V + ((long)b * w + i) * cs + vm;
                const float vr = v[0], vi = v[mt];
                ur[b] = fmaf(vr, rr, fmaf(-vi, rim, ur[b]));
                ui[b] = fmaf(vr, rim, fmaf(vi, rr, ui[b]));
            }
        }
    }
#pragma unroll
    for (int b = 0; b < MIX_BMAX; ++b) {
        if (b < nb) {
            float *u = U + ((long)b * w + o) * cs + vm;
            u[0] = ur[b];
            u[mt] = ui[b];
        }
    }
}
/*
 * Its backward, one thread per (mode, i):  dV[b][i] = sum_o dU[b][o] conj(R[i][o])
 * (assigned) and dR[i][o] += sum_b conj(V[b][i]) dU[b][o] (accumulated).
 */
__global__ void __launch_bounds__(256) k_mix_bwd(const float *__restrict__ V, const float *__restrict__ dU,
                                                 const float *__restrict__ Rr, const float *__restrict__ Ri,
                                                 float *__restrict__ dV, float *__restrict__ dRr,
                                                 float *__restrict__ dRi, int nb, int w, int nxy, int mt) {
    const int nm = nxy * mt;
    const int m = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (m >= nm) return;
    const int kxy = m / mt, kt = m % mt;
    const long vm = (long)kxy * 2 * mt + kt;
    const long cs = (long)nxy * 2 * mt;
    float vr[MIX_BMAX], vi[MIX_BMAX], gr[MIX_BMAX], gi[MIX_BMAX];
#pragma unroll
    for (int b = 0; b < MIX_BMAX; ++b) {
        vr[b] = vi[b] = gr[b] = gi[b] = 0.f;
        if (b < nb) {
            const float *v = V + ((long)b * w + i) * cs + vm;
            vr[b] = v[0];
            vi[b] = v[mt];
        }
    }
    for (int o = 0; o < w; ++o) {
        const long ri = ((long)i * w + o) * nm + m;
        const float rr = Rr[ri], rim = Ri[ri];
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int b = 0; b < MIX_BMAX; ++b) {
            if (b < nb) {
                const float *d = dU + ((long)b * w + o) * cs + vm;
                const float dr = d[0], di = d[mt];
                gr[b] = fmaf(dr, rr, fmaf(di, rim, gr[b]));        /* dU conj(R) */
                gi[b] = fmaf(di, rr, fmaf(-dr, rim, gi[b]));
                sr = fmaf(vr[b], dr, fmaf(vi[b], di, sr));         /* conj(V) dU */
                si = fmaf(vr[b], di, fmaf(-vi[b], dr, si));
            }
        }
        dRr[ri] += sr;
        dRi[ri] += si;
    }
#pragma unroll
    for (int b = 0; b < MIX_BMAX; ++b) {
        if (b < nb) {
            float *d = dV + ((long)b * w + i) * cs + vm;
            d[0] = gr[b];
            d[mt] = gi[b];
        }
    }
}

/* db[c] (+)= sum over b, p of d[b][c][p], in two stages: workgroup (c, j)
 * sums a fixed chunk of channel c into part[c][j]; then one thread per
 * channel adds its chunks in order (deterministic) */
constexpr int BG_CHUNKS = 64;
__global__ void __launch_bounds__(256) k_bias_part(const float *__restrict__ d, float *__restrict__ part, long nb,
                                                   int nc, long np) {
    __shared__ float red[256];
    const int c = blockIdx.x, j = blockIdx.y;
    const long tot = nb * np;
    const long per = (tot + BG_CHUNKS - 1) / BG_CHUNKS;
    const long e0 = j * per, e1 = min(tot, e0 + per);
    float s = 0.f;
    for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const long b = e / np, p = e % np;
        s += d[(b * nc + c) * np + p];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[c * BG_CHUNKS + j] = red[0];
}
__global__ void k_bias_sum(const float *__restrict__ part, float *__restrict__ db, int nc, int accumulate) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    float s = 0.f;
    for (int j = 0; j < BG_CHUNKS; ++j) s += part[c * BG_CHUNKS + j];
    db[c] = accumulate ? db[c] + s : s;
}
/* Adam with bias correction: m, v moments; lr already corrected by the caller */
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float wd) {
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
        const float gi = g[x] + wd * w[x];
        const float mi = b1 * m[x] + (1.f - b1) * gi;
        const float vi = b2 * v[x] + (1.f - b2) * gi * gi;
        m[x] = mi;
        v[x] = vi;
        w[x] -= lr * mi / (sqrtf(vi) + eps);
    }
}

/*
 * relative-L2 loss: per sample b, s[b] = (|pred-tgt|^2, |tgt|^2) summed in
 * two stages (workgroup (b, j) sums chunk j into part, then in chunk order),
 * then loss[0] = mean_b s0/s1 and dpred = (2/nb) (pred - tgt) / s1[b].
 * loss holds 1 + 2 nb floats.
 */
constexpr int RL_CHUNKS = 64;
__global__ void __launch_bounds__(256) k_rel_l2_part(const float *__restrict__ pred, const float *__restrict__ tgt,
                                                     float *__restrict__ part, long nper) {
    __shared__ float r0[256], r1[256];
    const long b = blockIdx.x;
    const int j = blockIdx.y;
    const long per = (nper + RL_CHUNKS - 1) / RL_CHUNKS;
    const long e0 = j * per, e1 = min(nper, e0 + per);
    float a = 0.f, c = 0.f;
    for (long i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
        const float p = pred[b * nper + i], t = tgt[b * nper + i];
        a += (p - t) * (p - t);
        c += t * t;
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) { r0[threadIdx.x] += r0[threadIdx.x + o]; r1[threadIdx.x] += r1[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * (b * RL_CHUNKS + j)] = r0[0]; part[2 * (b * RL_CHUNKS + j) + 1] = r1[0]; }
}
__global__ void k_rel_l2_sums(const float *__restrict__ part, float *__restrict__ loss, int nb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float a = 0.f, c = 0.f;
    for (int j = 0; j < RL_CHUNKS; ++j) { a += part[2 * (b * RL_CHUNKS + j)]; c += part[2 * (b * RL_CHUNKS + j) + 1]; }
    loss[1 + 2 * b] = a;
    loss[2 + 2 * b] = c;
}
__global__ void k_rel_l2_grad(const float *__restrict__ pred, const float *__restrict__ tgt, float *__restrict__ dpred,
                              float *__restrict__ loss, int nb, long nper) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float l = 0.f;
        for (int b = 0; b < nb; ++b) l += loss[1 + 2 * b] / loss[2 + 2 * b];
        loss[0] = l / nb;
    }
    const long total = nper * nb;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total; x += (long)gridDim.x * blockDim.x) {
        const long b = x / nper;
        dpred[x] = 2.f / nb * (pred[x] - tgt[x]) / loss[2 + 2 * b];
    }
}

/* out[j][i] = in[i][j] for a rows x cols matrix (small: weight matrices) */
__global__ void k_transpose(const float *__restrict__ in, float *__restrict__ out, int rows, int cols) {
    const long n = (long)rows * cols;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
        const int i = (int)(x / cols), j = (int)(x % cols);
        out[(long)j * rows + i] = in[x];
    }
}

__global__ void k_fill(float *__restrict__ p, long n, float v) {
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) p[x] = v;
}

int grid_for(long n) {
    long b = (n + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

} // namespace

extern "C" int mpc_sizeof_gemm(void) { return (int)sizeof(mpc_gemm); }

extern "C" int mpc_gemm_bf16(const mpc_gemm *g, void *stream) {
    if (!g || g->M <= 0 || g->N <= 0 || g->K <= 0 || g->nb1 <= 0 || g->nb2 <= 0 || g->ksplit <= 0) return MPC_EINVAL;
    if (!g->A || !g->B || !g->C || (g->ksplit > 1 && !g->W)) return MPC_EINVAL;
    const long nbat = (long)g->nb1 * g->nb2;
    const long gz = nbat * g->ksplit;
    if (gz > 65535 || (g->M + TM - 1) / TM > 65535) return MPC_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((unsigned)((g->N + TN - 1) / TN), (unsigned)((g->M + TM - 1) / TM), (unsigned)gz);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, *g);
    if (hipGetLastError() != hipSuccess) return MPC_EHIP;
    if (g->ksplit > 1) {
        const long outs = (long)g->M * g->N * nbat;
        if ((outs + RD_OUT - 1) / RD_OUT > 0x7fffffffL) return MPC_EINVAL;
        hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)((outs + RD_OUT - 1) / RD_OUT)), dim3(RD_OUT * RD_SUB), 0, s,
                           *g);
        if (hipGetLastError() != hipSuccess) return MPC_EHIP;
    }
    return 0;
}

extern "C" int mpc_sizeof_pw(void) { return (int)sizeof(mpc_pw); }

extern "C" int mpc_fno_pointwise(const mpc_pw *a, void *stream) {
    if (!a || !a->x || !a->w || !a->out || a->nb <= 0 || a->nb > 65535 || a->cin <= 0 || a->cout <= 0 ||
        a->np <= 0 || (a->beta != 0.f && !a->zin))
        return MPC_EINVAL;
    const dim3 grid((unsigned)((a->np + 255) / 256), (unsigned)a->nb);
    hipStream_t s = (hipStream_t)stream;
    if (pointwise_mf_fits(a)) {
        switch (a->cin) {
        case 16: launch_pointwise_mf<16>(a, s); break;
        case 32: launch_pointwise_mf<32>(a, s); break;
        default: launch_pointwise_mf<64>(a, s); break;
        }
        return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
    }
    switch (a->cin) {
    case 1: launch_pointwise<1>(a, grid, s); break;
    case 2: launch_pointwise<2>(a, grid, s); break;
    case 4: launch_pointwise<4>(a, grid, s); break;
    case 8: launch_pointwise<8>(a, grid, s); break;
    case 16: launch_pointwise<16>(a, grid, s); break;
    case 32: launch_pointwise<32>(a, grid, s); break;
    case 64: launch_pointwise<64>(a, grid, s); break;
    default: return MPC_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_transpose(const float *in, float *out, int rows, int cols, void *stream) {
    if (!in || !out || in == out || rows <= 0 || cols <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_transpose, dim3(grid_for((long)rows * cols)), dim3(256), 0, (hipStream_t)stream, in, out,
                       rows, cols);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_mix(const float *V, const float *Rr, const float *Ri, float *U, int nb, int w, int nxy,
                           int mt, void *stream) {
    if (!V || !Rr || !Ri || !U || nb <= 0 || nb > MIX_BMAX || w <= 0 || w > 65535 || nxy <= 0 || mt <= 0)
        return MPC_EINVAL;
    const dim3 grid((unsigned)((nxy * mt + 255) / 256), (unsigned)w);
    hipLaunchKernelGGL(k_mix, grid, dim3(256), 0, (hipStream_t)stream, V, Rr, Ri, U, nb, w, nxy, mt);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_mix_bwd(const float *V, const float *dU, const float *Rr, const float *Ri, float *dV,
                               float *dRr, float *dRi, int nb, int w, int nxy, int mt, void *stream) {
    if (!V || !dU || !Rr || !Ri || !dV || !dRr || !dRi || nb <= 0 || nb > MIX_BMAX || w <= 0 || w > 65535 ||
        nxy <= 0 || mt <= 0)
        return MPC_EINVAL;
    const dim3 grid((unsigned)((nxy * mt + 255) / 256), (unsigned)w);
    hipLaunchKernelGGL(k_mix_bwd, grid, dim3(256), 0, (hipStream_t)stream, V, dU, Rr, Ri, dV, dRr, dRi, nb, w, nxy,
                       mt);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_combine(const float *in, float *out, long nslice, int m, int nb, int h, int sign, int bwd,
                               void *stream) {
    if (!in || !out || nslice <= 0 || m <= 0 || nb <= 0 || h <= 0 || (sign != 1 && sign != -1)) return MPC_EINVAL;
    hipLaunchKernelGGL(k_combine, dim3(grid_for(nslice * m * nb * h)), dim3(256), 0, (hipStream_t)stream, in, out,
                       nslice, m, nb, h, (float)sign, bwd);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_bias_act(float *z, const float *bias, float *out, long nb, int nc, long np, int act,
                                void *stream) {
    if (!z || !out || nb <= 0 || nc <= 0 || np <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(nb * nc * np)), dim3(256), 0, (hipStream_t)stream, z, bias, out, nb,
                       nc, np, act);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_act_bwd(const float *z, const float *dout, float *dz, long n, int act, void *stream) {
    if (!z || !dout || !dz || n <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, z, dout, dz, n, act);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_bias_grad(const float *d, float *db, long nb, int nc, long np, int accumulate, void *stream) {
    if (!d || !db || nb <= 0 || nc <= 0 || np <= 0) return MPC_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    float *part = nullptr;
    if (hipMallocAsync((void **)&part, sizeof(float) * nc * BG_CHUNKS, s) != hipSuccess) return MPC_ENOMEM;
    hipLaunchKernelGGL(k_bias_part, dim3(nc, BG_CHUNKS), dim3(256), 0, s, d, part, nb, nc, np);
    hipLaunchKernelGGL(k_bias_sum, dim3((nc + 63) / 64), dim3(64), 0, s, part, db, nc, accumulate);
    const bool ok = hipGetLastError() == hipSuccess;
    (void)hipFreeAsync(part, s);
    return ok ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
                            float eps, float wd, void *stream) {
    if (!w || !g || !m || !v || n <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps, wd);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}

extern "C" int mpc_fno_rel_l2(const float *pred, const float *tgt, float *dpred, float *loss, int nb, long nper,
                              void *stream) {
    if (!pred || !tgt || !dpred || !loss || nb <= 0 || nper <= 0) return MPC_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    float *part = nullptr;
    if (hipMallocAsync((void **)&part, sizeof(float) * 2 * nb * RL_CHUNKS, s) != hipSuccess) return MPC_ENOMEM;
    hipLaunchKernelGGL(k_rel_l2_part, dim3(nb, RL_CHUNKS), dim3(256), 0, s, pred, tgt, part, nper);
    hipLaunchKernelGGL(k_rel_l2_sums, dim3((nb + 63) / 64), dim3(64), 0, s, part, loss, nb);
    hipLaunchKernelGGL(k_rel_l2_grad, dim3(grid_for(nper * nb)), dim3(256), 0, s, pred, tgt, dpred, loss, nb, nper);
    const bool ok = hipGetLastError() == hipSuccess;
    (void)hipFreeAsync(part, s);
    return ok ? 0 : MPC_EHIP;
}
/* device buffers for the surrogate's host driver (the default stream) */
extern "C" int mpc_dev_alloc(long bytes, void **ptr) {
    if (!ptr || bytes <= 0) return MPC_EINVAL;
    *ptr = nullptr;
    return hipMalloc(ptr, (size_t)bytes) == hipSuccess ? 0 : MPC_ENOMEM;
}
extern "C" void mpc_dev_free(void *ptr) {
    if (ptr) (void)hipFree(ptr);
}
extern "C" int mpc_dev_upload(void *dst, const void *src, long bytes) {
    if (!dst || !src || bytes < 0) return MPC_EINVAL;
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : MPC_EHIP;
}
extern "C" int mpc_dev_download(void *dst, const void *src, long bytes) {
    if (!dst || !src || bytes < 0) return MPC_EINVAL;
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : MPC_EHIP;
}
extern "C" int mpc_dev_copy(void *dst, const void *src, long bytes, void *stream) {
    if (!dst || !src || bytes < 0) return MPC_EINVAL;
    return hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
               ? 0 : MPC_EHIP;
}
extern "C" int mpc_dev_fill(float *p, long n, float v, void *stream) {
    if (!p || n <= 0) return MPC_EINVAL;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, p, n, v);
    return hipGetLastError() == hipSuccess ? 0 : MPC_EHIP;
}
extern "C" int mpc_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : MPC_EHIP; }
