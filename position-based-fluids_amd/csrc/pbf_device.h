// pbf_device.h -- device-side helpers shared by the PBF kernels.
//
// Everything here is evaluated in IEEE binary32 with no FMA contraction (the
// library is compiled with -ffp-contract=off), in the operation order of the
// reference kernels, so that results are reproducible across launch shapes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

constexpr int kCircles = 5;          // Params.friendsCircles (src/ParamUtils.cpp:98)
constexpr int kPerCircle = 50;       // Params.particlesPerCircle (src/ParamUtils.cpp:99)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // END_OF_CELL_LIST (src/Simulation.cpp:170)

// Per-step constants, passed by value to every kernel.
struct StepConsts {
    float h, h2, min_r, span;       // span = h - 0.3h (build_friends_list.cl:32,75)
    float dt;
    float g[3];
    float rest_density, epsilon;
    float poly6, spiky;             // the 6-significant-digit factors
    float st_k, poly6_q;            // s_corr constants (compute_delta.cl:81-82,152)
    float viscosity, vorticity;
    float lo[3], hi[3];
    uint32_t walls;
    uint32_t gbs, gbs_mask;         // gbs_mask != 0 when gbs is a power of two
    uint32_t n;
};

// utilities.cl:69-77 on the 32-bit two's-complement pattern.
__device__ __forceinline__ uint32_t expand_bits(int32_t xi) {
    uint32_t x = (uint32_t)xi;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// utilities.cl:79-87.  The Morton number is below 2^30, so the signed modulo
// of the reference equals an unsigned one, and a mask when gbs = 2^k.
__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz, const StepConsts& c) {
    const uint32_t m = expand_bits(cx) | (expand_bits(cy) << 1) | (expand_bits(cz) << 2);
    return c.gbs_mask ? (m & c.gbs_mask) : (m % c.gbs);
}

// convert_int of x/h (radixsort.cl:12): truncation toward zero; NaN -> 0 and
// magnitudes clamped to 2^24 (undefined in the reference).
__device__ __forceinline__ int32_t cell_coord(float q) {
    if (!(q == q)) return 0;
    if (q >= 16777216.0f) return 16777216;
    if (q <= -16777216.0f) return -16777216;
    return (int32_t)q;
}

// Sine used only for the wall edge-offset noise (compute_delta.cl:165-166):
// Cody-Waite reduction by pi with a three-part split, odd Taylor polynomial
// of degree 11 on [-pi/2, pi/2]; |x| >= 65536 or NaN gives 0.
__device__ __forceinline__ float noise_sin(float x) {
    if (!(fabsf(x) < 65536.0f)) return 0.0f;
    const float k = rintf(x * 0x1.45f306p-2f);
    float r = x - k * 0x1.92p+1f;
    r = r - k * 0x1.fb6p-11f;
    r = r - k * -0x1.777a5cp-24f;
    const float r2 = r * r;
    float p = -2.50521083854417e-8f;
    p = p * r2 + 2.75573192239859e-6f;
    p = p * r2 + -1.98412698412698e-4f;
    p = p * r2 + 8.33333333333333e-3f;
    p = p * r2 + -1.66666666666667e-1f;
    p = r + (r * r2) * p;
    return (((int32_t)k) & 1) ? -p : p;
}

// Number of friend circles a sweep processes.  The reference skips circle q
// when (float)processed/total > threshold (compute_scaling.cl:50, 0.6;
// compute_delta.cl:100, 0.5); once one circle is skipped `processed` stops
// growing, so the processed circles form a prefix.  For counts <= 250 the
// fp32 test equals the exact rational one used here.
template <int NUM, int DEN>  // threshold = NUM/DEN
__device__ __forceinline__ int circles_processed(uint32_t packed) {
    int c[kCircles];
    int total = 0;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) { c[q] = (packed >> (6 * q)) & 63; total += c[q]; }
    int done = 0, L = 0;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) {
        if (L == q && !(DEN * done > NUM * total)) { done += c[q]; L = q + 1; }
    }
    return L;
}

__device__ __forceinline__ int circle_count(uint32_t packed, int q) { return (packed >> (6 * q)) & 63; }

// Friend index slot in the reference layout (build_friends_list.cl:18-21):
// indices[circle][f][i] after the counts block.
__device__ __forceinline__ size_t friend_slot(int q, int f, uint32_t i, uint32_t n) {
    return ((size_t)q * kPerCircle + (size_t)f) * n + i;
}

// Box wall planes: BouncePointQuad (compute_delta.cl:1-42) for an infinite
// axis-aligned plane with mask 1.
__device__ __forceinline__ float wall_lo(float f, float prev, float lo, float edge) {
    const float pa = lo + edge;
    const float dp = f - pa;
    if (dp > 0.0f) return f;
    const float nv = prev - f;
    if (fabsf(dp) > edge + nv) return f;
    return pa;
}

__device__ __forceinline__ float wall_hi(float f, float prev, float hi, float edge) {
    const float pa = hi - edge;
    const float dp = f - pa;
    if (dp < 0.0f) return f;
    const float nv = f - prev;
    if (fabsf(dp) > edge + nv) return f;
    return pa;
}

__device__ __forceinline__ void apply_walls(float& fx, float& fy, float& fz, const float4& prev, const StepConsts& c) {
    const float nx = noise_sin(fx * 5.0f / c.h);
    const float ny = noise_sin(fy * 5.0f / c.h);
    const float nz = noise_sin(fz * 5.0f / c.h);
    const float edge = (3.0f + nx + ny + nz) * c.h * 0.03f;
    fx = wall_lo(fx, prev.x, c.lo[0], edge);
    fx = wall_hi(fx, prev.x, c.hi[0], edge);
    fy = wall_lo(fy, prev.y, c.lo[1], edge);
    fy = wall_hi(fy, prev.y, c.hi[1], edge);
    fz = wall_lo(fz, prev.z, c.lo[2], edge);
    fz = wall_hi(fz, prev.z, c.hi[2], edge);
}

}  // namespace pbf
