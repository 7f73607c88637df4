/*
 * mpc_fno.h -- kernels of the Fourier-neural-operator pass surrogate.
 *
 * The reference has no learned part and so no counterpart for these entry
 * points.  They operate on device buffers (FP32; mpc_dev_alloc or the
 * caller's own) on a HIP stream (void *stream, NULL = default stream) and
 * return 0 or a negative MPC_E* code (mpc.h).  The network built from them (mpc_amd/fno.py) maps the
 * field at one mirror plane to the field after a pass; the plain split step
 * stays the fallback and the measure of its error.
 */
#ifndef MPC_FNO_H
#define MPC_FNO_H

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Batched GEMM on the matrix cores: C = alpha * A B + beta * C with
 * A[i][k] = A[i*rsA + k*csA], B[k][j] = B[k*rsB + j*csB],
 * C[i][j] = C[i*rsC + j*csC], for every batch index (b1, b2) in
 * [0, nb1) x [0, nb2), each operand offset by b1*s?1 + b2*s?2 elements.
 * A and B are rounded to BF16 as they are staged; products accumulate in
 * FP32.  ksplit > 1 splits K over workgroups: W (ksplit*nb1*nb2*M*N floats)
 * receives the partial sums, then a reduction applies alpha and beta.
 * beta == 0 never reads C.
 */
typedef struct {
    int M, N, K;
    int nb1, nb2;
    int ksplit;
    const float *A;
    const float *B;
    float *C;
    float *W;
    long rsA, csA, rsB, csB, rsC, csC;
    long sA1, sB1, sC1, sA2, sB2, sC2;
    float alpha, beta;
} mpc_gemm;

int mpc_gemm_bf16(const mpc_gemm *g, void *stream);
int mpc_sizeof_gemm(void);   /* sizeof(mpc_gemm), for bindings to check their layout */

/*
 * Complex recombination after a real-block DFT product.  in holds nslice
 * slices of 2m rows (cos-weighted rows, then sin-weighted rows) of nb blocks
 * [re(0..h) | im(0..h)]; out holds m rows:  re = c_re - sign * s_im,
 * im = c_im + sign * s_re  (sign -1: forward transform, +1: inverse).
 * bwd = 1: the transpose map (out = dP from in = dV, same shapes reversed).
 */
int mpc_fno_combine(const float *in, float *out, long nslice, int m, int nb, int h, int sign, int bwd, void *stream);

/*
 * Pointwise channel map with its epilogue (FP32):
 *   z[b][o][p] = sum_i w[o*ldo + i*ldi] x[b*x_bs + i*np + p]
 *                + beta * zin[b*zin_bs + o*np + p] + bias[o]
 *   out[b][o][p] = GELU(z) if act else z   (out, zs: batch stride cout*np)
 * zs (NULL: not stored) receives z; zin is read only if beta != 0; bias may
 * be NULL; out may alias zin or zs; w and bias must not alias an output.
 * cin in {1, 2, 4, 8, 16, 32, 64}, nb <= 65535.
 */
typedef struct {
    int nb, cin, cout;
    long np;
    const float *x;
    long x_bs;
    const float *w;
    long ldo, ldi;
    const float *bias;
    const float *zin;
    long zin_bs;
    float beta;
    float *zs;
    float *out;
    int act;
} mpc_pw;

int mpc_fno_pointwise(const mpc_pw *a, void *stream);
int mpc_sizeof_pw(void);    /* sizeof(mpc_pw), for bindings to check their layout */

/* out (cols x rows) = in (rows x cols)^T; in != out */
int mpc_fno_transpose(const float *in, float *out, int rows, int cols, void *stream);

/*
 * Per-mode complex channel mixing of the spectral convolution (FP32).
 * V, U: [nb][w][nxy][2 mt] floats (re at kt, im at mt + kt of each kxy
 * block); Rr, Ri: [w][w][nxy][mt] (input channel, output channel, mode).
 * mpc_fno_mix:     U[b][o] = sum_i V[b][i] (Rr + i Ri)[i][o]   per mode.
 * mpc_fno_mix_bwd: dV[b][i] = sum_o dU[b][o] conj(R[i][o]) (assigned) and
 *                  dR[i][o] += sum_b conj(V[b][i]) dU[b][o] (accumulated).
 * nb <= 8.
 */
int mpc_fno_mix(const float *V, const float *Rr, const float *Ri, float *U, int nb, int w, int nxy, int mt,
                void *stream);
int mpc_fno_mix_bwd(const float *V, const float *dU, const float *Rr, const float *Ri, float *dV, float *dRr,
                    float *dRi, int nb, int w, int nxy, int mt, void *stream);

/* z[b][c][p] += bias[c] (bias may be NULL), then out = GELU(z) if act else z */
int mpc_fno_bias_act(float *z, const float *bias, float *out, long nb, int nc, long np, int act, void *stream);
/* dz = dout * GELU'(z) if act else dout */
int mpc_fno_act_bwd(const float *z, const float *dout, float *dz, long n, int act, void *stream);
/* db[c] = (accumulate ? db[c] : 0) + sum_{b,p} d[b][c][p] */
int mpc_fno_bias_grad(const float *d, float *db, long nb, int nc, long np, int accumulate, void *stream);
/* one Adam step on n parameters (moments m, v; lr already bias-corrected) */
int mpc_fno_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps,
                 float wd, void *stream);

/* relative-L2 training loss over nb samples of nper floats: loss (1 + 2 nb
 * floats) gets loss[0] = mean_b |pred-tgt|^2 / |tgt|^2 and the per-sample
 * sums; dpred = d loss[0] / d pred */
int mpc_fno_rel_l2(const float *pred, const float *tgt, float *dpred, float *loss, int nb, long nper, void *stream);

/* device buffers for the surrogate's host driver: allocation, host<->device
 * copies (synchronous), device copies and fills on a stream, and a device-wide
 * synchronisation */
int  mpc_dev_alloc(long bytes, void **ptr);
void mpc_dev_free(void *ptr);
int  mpc_dev_upload(void *dst, const void *src, long bytes);
int  mpc_dev_download(void *dst, const void *src, long bytes);
int  mpc_dev_copy(void *dst, const void *src, long bytes, void *stream);
int  mpc_dev_fill(float *p, long n, float v, void *stream);
int  mpc_dev_sync(void);

#ifdef __cplusplus
}
#endif

#endif
