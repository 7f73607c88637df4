/*
 * fno.hip -- kernels of the Fourier-neural-operator pass surrogate.
 *
 * The reference has no learned part.  The surrogate maps the field at one
 * mirror plane to the field after a pass; every dense contraction in it
 * (the truncated DFTs along t, y and x, the per-mode channel mixing, the
 * pointwise channel maps, and all their transposes in the backward pass) is
 * a batched GEMM run here on the matrix cores: operands are rounded to BF16
 * as they are staged into LDS, products accumulate in FP32
 * (v_mfma_f32_16x16x16_bf16).  The rest are small elementwise kernels:
 * complex recombination after the real-block DFT products, bias + GELU and
 * its derivative, bias reductions and the Adam update.
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

    for (int k0 = kbeg; k0 < kend; k0 += TK) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int ii, kk;
            if (a_krow) { kk = t % TK; ii = t / TK + 8 * r; }
            else        { ii = t % TM; kk = t / TM + 4 * r; }
            const int gi = i0 + ii, gk = k0 + kk;
            float v = 0.f;
            if (gi < g.M && gk < kend) v = A[(long)gi * g.rsA + (long)gk * g.csA];
            As[ii][kk] = to_bf16(v);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int jj, kk;
            if (b_ncol) { jj = t % TN; kk = t / TN + 4 * r; }
            else        { kk = t % TK; jj = t / TK + 8 * r; }
            const int gj = j0 + jj, gk = k0 + kk;
            float v = 0.f;
            if (gj < g.N && gk < kend) v = B[(long)gk * g.rsB + (long)gj * g.csB];
            Bs[jj][kk] = to_bf16(v);
        }
        __syncthreads();
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
