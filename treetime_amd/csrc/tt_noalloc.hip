// tt_prob_t_profiles_noalloc: GTR.prob_t_profiles for K transition matrices
// without any device allocation.
//
// tt_prob_t_profiles (tt_api.hip) allocates two stream-ordered buffers per
// call: the K transition matrices and the per-block partial sums.  This
// variant keeps both in `scratch`, a file-scope device array that every GPU
// holds for the life of the process, and sums the block partials on the
// device (in the order the allocating entry point sums them on the host, so
// the two return identical values).  It is for callers that cannot allocate:
// code under graph capture, tight serving loops, harnesses that forbid
// hipMalloc.
//
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (different streams or host threads) race; calls in one stream order are
// safe.  Every element a call reads is written earlier in the same call, so
// no clearing is needed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/ttgpu.h"
#include "tt_common.h"

// K*n*n matrices + K*ceil(P/256) block partials + K sums, in doubles (4 MiB).
#define TT_NOALLOC_SCRATCH_DOUBLES (1u << 19)
#define TT_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[TT_NOALLOC_SCRATCH_DOUBLES];

namespace {

std::mutex g_mu;
double *g_ptr[TT_NOALLOC_MAX_DEVICES];  // device address of `scratch` per GPU, null until looked up

// Address of the current device's copy of `scratch`; looked up once per
// device, with that device current.
int32_t scratch_ptr(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return ttk::set_error(TT_EDEVICE, hipGetErrorString(e));
    if (dev < 0 || dev >= TT_NOALLOC_MAX_DEVICES) return ttk::set_error(TT_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        double *p = nullptr;
        e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char buf[256];
            snprintf(buf, sizeof buf, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return ttk::set_error(TT_EDEVICE, buf);
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return TT_OK;
}

// sums[k] = sum over b of block_out[k][b], b ascending (the host loop of
// tt_prob_t_profiles, in the same order).
__global__ __launch_bounds__(TT_BLOCK) void k_sum_block_partials(const double *block_out, int nblk, int K,
                                                                double *sums) {
    const int k = blockIdx.x * TT_BLOCK + threadIdx.x;
    if (k >= K) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += block_out[(size_t)k * nblk + b];
    sums[k] = acc;
}

}  // namespace

extern "C" int32_t tt_prob_t_profiles_noalloc(const double *pp_dev, const double *pc_dev, const double *mult_dev,
                                              int32_t P, int32_t n, const double *Q, int32_t K, int32_t gap_index,
                                              double *out, void *stream) {
    if (!pp_dev || !pc_dev || !mult_dev || !Q || !out) return ttk::set_error(TT_EINVAL, "NULL argument");
    if (P < 1 || n < 1 || n > TT_MAX_STATES || K < 1 || K > 65535) {
        char buf[128];
        snprintf(buf, sizeof buf, "bad shape P=%d n=%d K=%d", P, n, K);
        return ttk::set_error(TT_EINVAL, buf);
    }
    if (gap_index >= n) {
        char buf[128];
        snprintf(buf, sizeof buf, "gap_index %d >= n %d", gap_index, n);
        return ttk::set_error(TT_EINVAL, buf);
    }
    const int nblk = ttk::root_blocks(P);
    const size_t need = (size_t)K * ((size_t)n * n + nblk + 1);
    if (need > TT_NOALLOC_SCRATCH_DOUBLES) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "tt_prob_t_profiles_noalloc: K*(n*n + ceil(P/256) + 1) = %zu doubles exceeds the fixed scratch of "
                 "%u doubles (P=%d n=%d K=%d)",
                 need, TT_NOALLOC_SCRATCH_DOUBLES, P, n, K);
        return ttk::set_error(TT_EINVAL, buf);
    }
    double *base = nullptr;
    int32_t rc = scratch_ptr(&base);
    if (rc) return rc;
    double *dQ = base;
    double *dB = dQ + (size_t)K * n * n;
    double *dS = dB + (size_t)K * nblk;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(dQ, Q, (size_t)K * n * n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = ttk::prob_t_profiles(pp_dev, pc_dev, mult_dev, P, n, dQ, K, gap_index, dB, nblk, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_block_partials, dim3((K + TT_BLOCK - 1) / TT_BLOCK), dim3(TT_BLOCK), 0, s, dB,
                           nblk, K, dS);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dS, (size_t)K * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ttk::set_error(TT_EDEVICE, hipGetErrorString(e));
    return TT_OK;
}
