/*
 * noalloc.hip -- setStep's error norm on device-resident fields, with no
 * allocation at all: for callers that cannot allocate device memory (code
 * inside stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).
 *
 * The partial sums live in `scratch`, a file-scope device array of fixed
 * size.  Host code never passes the array's name to a kernel or a copy (in
 * host code it is only a stand-in for the symbol): the device address of each
 * GPU's copy is looked up once per device with hipGetSymbolAddress and kept
 * in g_scratch[device], and that address is what the kernels and copies use.
 *
 * Every call on a GPU shares that GPU's array and leaves its contents behind;
 * each slot a call reads was written earlier in the same call, so nothing
 * needs clearing.  Calls that overlap on one GPU (several host threads, or
 * streams other than the null stream used here) race on the array.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/mpc.h"
#include "internal.h"

/* one partial (two doubles) per chunk of NOALLOC_CHUNK points */
static constexpr long NOALLOC_CHUNK = 1L << 16;
static constexpr int NOALLOC_MAX_PARTS = 4096;                /* fields of at most 2^28 points */
static constexpr int NOALLOC_DOUBLES = 2 * NOALLOC_MAX_PARTS + 2;   /* partials, then the two totals */

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];   /* 64 KB per GPU */

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES];                               /* device address of scratch, per GPU */

/* the device address of this GPU's scratch array (looked up on the first
 * call on each device, with that device current) */
int scratch_of_current_device(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MPC_ENODEV;
    if (dev < 0 || dev >= MAX_DEVICES) return MPC_EUNSUPP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "mpc: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return MPC_EHIP;
        }
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return 0;
}

/* block b: sum |c2 - c1|^2 and |c2|^2 over points [b*chunk, min(n, (b+1)*chunk)) */
__global__ void __launch_bounds__(256) k_err_chunks(const double2 *__restrict__ c1, const double2 *__restrict__ c2,
                                                    long n, long chunk, double *__restrict__ part) {
    __shared__ double s1[256], s2[256];
    const long lo = (long)blockIdx.x * chunk;
    const long hi = lo + chunk < n ? lo + chunk : n;
    double a = 0.0, b = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += 256) {
        const double2 u = c1[i], v = c2[i];
        const double dr = v.x - u.x, di = v.y - u.y;
        a += dr * dr + di * di;
        b += v.x * v.x + v.y * v.y;
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
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

/* one block: the nparts partials in a fixed order into part[2*NOALLOC_MAX_PARTS + 0/1] */
__global__ void __launch_bounds__(256) k_err_total(double *part, int nparts) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) {
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
        part[2 * NOALLOC_MAX_PARTS] = s1[0];
        part[2 * NOALLOC_MAX_PARTS + 1] = s2[0];
    }
}

} // namespace

extern "C" int mpc_setstep_noalloc(const int *dims, double *dzz, const cufftDoubleComplex *C1,
                                   const cufftDoubleComplex *C2, double def_err, double dzmin, const double *y1,
                                   const double *y2, int *bad, double *err) {
    if (!dims || !dzz || !C1 || !C2 || !y1 || !y2 || !bad || !err) return MPC_EINVAL;
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return MPC_EINVAL;
    const long n = (long)dims[0] * dims[1] * dims[2];
    const long nparts = (n + NOALLOC_CHUNK - 1) / NOALLOC_CHUNK;
    if (nparts > NOALLOC_MAX_PARTS) {
        fprintf(stderr, "mpc_setstep_noalloc: %ld points exceed the %ld-point limit of its fixed scratch\n", n,
                (long)NOALLOC_MAX_PARTS * NOALLOC_CHUNK);
        return MPC_EINVAL;
    }
    double *part = nullptr;
    const int rc = scratch_of_current_device(&part);
    if (rc) return rc;
    hipLaunchKernelGGL(k_err_chunks, dim3((unsigned)nparts), dim3(256), 0, nullptr, (const double2 *)C1,
                       (const double2 *)C2, n, NOALLOC_CHUNK, part);
    if (hipGetLastError() != hipSuccess) return MPC_EHIP;
    hipLaunchKernelGGL(k_err_total, dim3(1), dim3(256), 0, nullptr, part, (int)nparts);
    if (hipGetLastError() != hipSuccess) return MPC_EHIP;
    double tot[2];
    if (hipMemcpy(tot, part + 2 * NOALLOC_MAX_PARTS, sizeof tot, hipMemcpyDeviceToHost) != hipSuccess)
        return MPC_EHIP;
    *err = std::sqrt(tot[0]) / std::sqrt(tot[1]);       /* propagation.c:625-631 */
    mpc::setstep_decide(dzz, *err, def_err, dzmin, y1, y2, bad);
    return 0;
}
