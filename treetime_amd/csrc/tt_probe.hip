// Streaming ceilings of the device's HBM, measured with hand-written kernels:
// a read-only sweep, a copy and a write-only sweep over buffers much larger
// than the caches, 16-byte accesses per lane, grid-stride loops over a grid
// that fills every CU.  The bench reports the kernels' roofline fractions
// against these beside the datasheet peak.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ttgpu.h"
#include "tt_common.h"

namespace {

constexpr int kProbeThreads = 256;

// The read sweep folds everything into one value per thread and stores it
// only if it is the impossible pattern below, so the loads cannot be
// dropped and the kernel writes (almost) nothing.
__global__ __launch_bounds__(kProbeThreads) void k_probe_read(const uint4 *__restrict__ src, size_t n16,
                                                              uint4 *__restrict__ sink) {
    const size_t stride = (size_t)gridDim.x * kProbeThreads;
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x; i < n16; i += stride) {
        const uint4 v = src[i];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
    }
    if (acc.x == 0x9e3779b9u && acc.y == 0x7f4a7c15u && acc.z == 0xf39cc060u && acc.w == 0x5cedc834u)
        sink[blockIdx.x * kProbeThreads + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_copy(const uint4 *__restrict__ src, size_t n16,
                                                              uint4 *__restrict__ dst) {
    const size_t stride = (size_t)gridDim.x * kProbeThreads;
    for (size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_write(size_t n16, uint4 *__restrict__ dst) {
    const size_t stride = (size_t)gridDim.x * kProbeThreads;
    const uint4 v = make_uint4(threadIdx.x, blockIdx.x, 0, 0);
    for (size_t i = (size_t)blockIdx.x * kProbeThreads + threadIdx.x; i < n16; i += stride) dst[i] = v;
}

}  // namespace

extern "C" int32_t tt_hbm_probe(int32_t kind, void *dst, const void *src, uint64_t bytes, int32_t reps, double *ms,
                                void *stream) {
    if (!ms || reps < 1 || bytes < 16 || (bytes & 15)) return ttk::set_error(TT_EINVAL, "bad probe arguments");
    if ((kind == 0 || kind == 1) && !src) return ttk::set_error(TT_EINVAL, "NULL source");
    if (!dst) return ttk::set_error(TT_EINVAL, "NULL destination");
    if (kind < 0 || kind > 2) return ttk::set_error(TT_EINVAL, "unknown probe kind");
    if (((uintptr_t)dst | (uintptr_t)src) & 15) return ttk::set_error(TT_EINVAL, "buffers must be 16-byte aligned");
    int dev = 0, n_cu = 0, bpc = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    const void *fn = kind == 0 ? (const void *)k_probe_read : kind == 1 ? (const void *)k_probe_copy
                                                                         : (const void *)k_probe_write;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, kProbeThreads, 0);
    if (e != hipSuccess) return ttk::set_error(TT_EDEVICE, hipGetErrorString(e));
    const unsigned grid = (unsigned)(n_cu * std::max(bpc, 1));
    // the read sink holds one uint4 per thread of the grid: dst must be at
    // least that large for kind 0 (it is never written in practice)
    if (kind == 0 && bytes < (uint64_t)grid * kProbeThreads * 16)
        return ttk::set_error(TT_EINVAL, "read probe needs bytes >= grid threads * 16 for its sink");
    const size_t n16 = bytes / 16;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a = nullptr, b = nullptr;
    e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(kProbeThreads), 0, s, (const uint4 *)src, n16,
                               (uint4 *)dst);
        else if (kind == 1)
            hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(kProbeThreads), 0, s, (const uint4 *)src, n16,
                               (uint4 *)dst);
        else
            hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(kProbeThreads), 0, s, n16, (uint4 *)dst);
    };
    if (e == hipSuccess) {
        launch();  // warm-up
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e != hipSuccess) return ttk::set_error(TT_EDEVICE, hipGetErrorString(e));
    *ms = (double)t / reps;
    return TT_OK;
}
