// pbf_probe.hip -- a plain HBM copy of known size.  It measures the copy
// bandwidth the chip sustains and, run under rocprofv3 --pmc, gives the
// FETCH_SIZE / WRITE_SIZE readings that scale the counters of the PBF
// kernels to bytes.
#include <stdio.h>

#include "pbf_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int pbf_copy_probe(uint64_t bytes, uint32_t reps, double* gbps) {
    if (bytes < 16 || reps == 0) return PBF_E_INVALID;
    const size_t n4 = (size_t)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PBF_OK;
    if (hipMalloc((void**)&a, n4 * 16) != hipSuccess || hipMalloc((void**)&b, n4 * 16) != hipSuccess) rc = PBF_E_NOMEM;
    if (rc == PBF_OK && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = PBF_E_HIP;
    if (rc == PBF_OK && hipMemsetAsync(a, 0, n4 * 16, s) != hipSuccess) rc = PBF_E_HIP;
    const unsigned blocks = 256u * 8u;
    if (rc == PBF_OK) {
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, a, b, n4);  // warm-up
        (void)hipEventRecord(e0, s);
        for (uint32_t r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, a, b, n4);
        (void)hipEventRecord(e1, s);
        float ms = 0.0f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = PBF_E_HIP;
        else if (gbps)
            *gbps = 2.0 * 16.0 * (double)n4 * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
