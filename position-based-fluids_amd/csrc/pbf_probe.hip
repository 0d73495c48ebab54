// pbf_probe.hip -- device self-checks.  A plain HBM copy of known size
// measures the copy bandwidth the chip sustains and, run under rocprofv3
// --pmc, gives the FETCH_SIZE / WRITE_SIZE readings that scale the counters
// of the PBF kernels to bytes.  An exhaustive comparison checks the kernels'
// square root and reciprocal against IEEE binary32.
#include <stdio.h>

#include "pbf_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

// Every positive normal binary32 input (bit patterns lo..hi): the kernels'
// square root against sqrtf from sqrt_lo on (their domain, kTinySq), and their
// reciprocal against 1.0f / x up to rcp_hi (where 1/x is normal).  bad =
// {sqrt count, rcp count}; first = {first sqrt input, first rcp input, last
// sqrt input}.
__global__ __launch_bounds__(256) void k_check_math(uint32_t lo, uint32_t hi, uint32_t sqrt_lo, uint32_t rcp_hi,
                                                    unsigned long long* bad, uint32_t* first) {
    const uint32_t stride = gridDim.x * blockDim.x;
    unsigned long long ns = 0, nr = 0;
    uint32_t fs = 0xFFFFFFFFu, fr = 0xFFFFFFFFu, ls = 0u;
    for (uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x; b <= hi && b >= lo; b += stride) {
        const float x = __uint_as_float(b);
        if (b >= sqrt_lo && pbf::sqrt_normal(x) != sqrtf(x)) {
            ++ns;
            fs = min(fs, b);
            ls = max(ls, b);
        }
        if (b <= rcp_hi && pbf::rcp_normal(x) != 1.0f / x) {
            ++nr;
            fr = min(fr, b);
        }
    }
    if (ns) {
        atomicAdd(bad, ns);
        atomicMin(first, fs);
        atomicMax(first + 2, ls);
    }
    if (nr) {
        atomicAdd(bad + 1, nr);
        atomicMin(first + 1, fr);
    }
}

}  // namespace

extern "C" int pbf_check_fast_math(uint64_t* sqrt_bad, uint64_t* rcp_bad, uint32_t* first_sqrt, uint32_t* first_rcp) {
    unsigned long long* bad = nullptr;
    uint32_t* first = nullptr;
    int rc = PBF_OK;
    if (hipMalloc((void**)&bad, 16) != hipSuccess || hipMalloc((void**)&first, 12) != hipSuccess) rc = PBF_E_NOMEM;
    const uint32_t init_first[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0u};
    const unsigned long long zero[2] = {0ull, 0ull};
    if (rc == PBF_OK && (hipMemcpy(bad, zero, 16, hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemcpy(first, init_first, 12, hipMemcpyHostToDevice) != hipSuccess))
        rc = PBF_E_HIP;
    if (rc == PBF_OK) {
        // normal inputs 0x00800000 .. 0x7F7FFFFF; square roots from 2^-96
        // (0x0F800000); 1/x is normal up to x = 2^126 (0x7E800000)
        static_assert(pbf::kTinySq == 0x1p-96f, "the square root is checked on its domain");
        hipLaunchKernelGGL(k_check_math, dim3(8192), dim3(256), 0, nullptr, 0x00800000u, 0x7F7FFFFFu, 0x0F800000u,
                           0x7E800000u, bad, first);
        unsigned long long h[2] = {0, 0};
        uint32_t f[3] = {0, 0, 0};
        if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(h, bad, 16, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(f, first, 12, hipMemcpyDeviceToHost) != hipSuccess)
            rc = PBF_E_HIP;
        if (sqrt_bad) *sqrt_bad = h[0];
        if (rcp_bad) *rcp_bad = h[1];
        if (first_sqrt) *first_sqrt = h[0] ? f[0] : 0u;
        if (first_rcp) *first_rcp = h[1] ? f[1] : 0u;
    }
    if (bad) (void)hipFree(bad);
    if (first) (void)hipFree(first);
    return rc;
}

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
