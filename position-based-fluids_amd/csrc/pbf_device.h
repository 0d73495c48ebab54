// pbf_device.h -- device-side helpers shared by the PBF kernels.
//
// Everything here is evaluated in IEEE binary32 with no implicit FMA
// contraction (the library is compiled with -ffp-contract=off; the fmaf calls
// are the only fused operations), in the operation order of the reference
// kernels, so that results are reproducible across launch shapes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

constexpr int kCircles = 5;          // Params.friendsCircles (src/ParamUtils.cpp:98)
constexpr int kPerCircle = 50;       // Params.particlesPerCircle (src/ParamUtils.cpp:99)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // END_OF_CELL_LIST (src/Simulation.cpp:170)

// Event log (the reference's device logger, logging.cl: a ring in device
// memory with an atomically advanced write index, read by the host after a
// step).  In the error-word allocation: word kLogIndex is the write index
// (monotonic, in words); the ring of kLogRing words starts at kLogRingStart.
// Every message is four words {code, a, b, c} (PBF_LOG_*), so messages never
// straddle the ring's end and the index stays aligned across its wrap.
constexpr uint32_t kLogIndex = 16;
constexpr uint32_t kLogRingStart = 20;
constexpr uint32_t kLogRing = 4096;
constexpr uint32_t kErrAllocWords = kLogRingStart + kLogRing;

__device__ __forceinline__ void log_event(uint32_t* err, uint32_t code, uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t at = atomicAdd(err + kLogIndex, 4u);
    *reinterpret_cast<uint4*>(err + kLogRingStart + (at % kLogRing)) = make_uint4(code, a, b, c);
}

constexpr uint32_t kMaxQuads = 26;   // hit bits 6..31 of a particle's boundary record

// One collision quad with the constants BouncePointQuad derives from it
// (compute_delta.cl:3-9), computed once on the host in binary32: corner B,
// edges E0, E1, unit normal n = cross(E0, E1) / |cross(E0, E1)|, the Gram
// entries aa = E0.E0, bb = E0.E1, cc = E1.E1 and invdet = 1 / (aa cc - bb^2);
// mask rows [mask_y, mask_y + mask_h).
struct QuadConsts {
    float b0, b1, b2, e00, e01, e02, e10, e11, e12, n0, n1, n2;
    float aa, bb, cc, invdet;
    int32_t mask_y, mask_h;
    uint32_t pad0, pad1;
};
static_assert(sizeof(QuadConsts) == 80, "QuadConsts is 20 words");

// Constants of one scene.  Kernels read them through a workgroup-uniform
// index (the block table), so they land in scalar registers.
struct SceneConsts {
    float h, h2, min_r, span;       // span = h - 0.3h (build_friends_list.cl:32,75)
    float dt, g0, g1, g2;
    float rest_density, epsilon, poly6, spiky;  // poly6/spiky: the 6-digit factors
    float st_k, inv_poly6_q, viscosity, vorticity;  // s_corr constants (compute_delta.cl:81-82,151-152)
    float lo0, lo1, lo2, hi0;
    float hi1, hi2;
    uint32_t walls, gbs;
    uint32_t gbs_mask;              // != 0 when gbs is a power of two
    uint32_t cell_base;             // first global cell id of this scene
    uint32_t first, count;          // particle range of this scene
    float circ_t[4];                // r^2 at which the friend circle reaches 1..4
    const QuadConsts* quads;        // collision quads (device), quad_count of them
    const uint32_t* mask;           // quad mask, 1 bit per texel, rows of (mask_w + 31) / 32 words
    uint32_t quad_count, mask_w, mask_rows;
    float edge_base;                // added to the noise edge for the quads
};
static_assert(sizeof(SceneConsts) == 160, "SceneConsts is 40 words");

// utilities.cl:69-77 on the 32-bit two's-complement pattern.
__device__ __forceinline__ uint32_t expand_bits(int32_t xi) {
    uint32_t x = (uint32_t)xi;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// utilities.cl:79-87, offset by the scene's first cell.  The Morton number is
// below 2^30, so the reference's signed modulo equals an unsigned one, and a
// mask when gbs = 2^k.
__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz, const SceneConsts& c) {
    const uint32_t m = expand_bits(cx) | (expand_bits(cy) << 1) | (expand_bits(cz) << 2);
    return c.cell_base + (c.gbs_mask ? (m & c.gbs_mask) : (m % c.gbs));
}

// Hashes of the 3x3x3 neighbourhood of cell (cx, cy, cz), the same values as
// cell_hash(cx + dx, cy + dy, cz + dz).  expand_bits keeps the low 10 bits of
// a coordinate, so the neighbours' dilated coordinates are the centre's
// stepped modulo 2^10 inside the dilated field M: +1 is ((d | ~M) + 1) & M
// (the carry skips the bits outside M) and -1 is (d - 1) & M (the borrow sets
// the bits below the lowest one of d, and M keeps its own).  That is 2-3
// instructions per axis and step instead of a Morton expansion per cell.
struct NeighbourHash {
    static constexpr uint32_t kM = 0x09249249u;
    uint32_t x[3], y[3], z[3];
    __device__ __forceinline__ static void steps(int32_t v, uint32_t* d) {
        const uint32_t e = expand_bits(v);
        d[0] = (e - 1u) & kM;
        d[1] = e;
        d[2] = ((e | ~kM) + 1u) & kM;
    }
    __device__ __forceinline__ NeighbourHash(int32_t cx, int32_t cy, int32_t cz) {
        steps(cx, x);
        steps(cy, y);
        steps(cz, z);
    }
    // d in {-1, 0, 1}; uniform d selects without indexing the arrays
    __device__ __forceinline__ static uint32_t pick(const uint32_t* a, int d) { return d < 0 ? a[0] : (d > 0 ? a[2] : a[1]); }
    __device__ __forceinline__ uint32_t operator()(int dx, int dy, int dz, const SceneConsts& c) const {
        const uint32_t m = pick(x, dx) | (pick(y, dy) << 1) | (pick(z, dz) << 2);
        return c.cell_base + (c.gbs_mask ? (m & c.gbs_mask) : (m % c.gbs));
    }
};

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

// Square root and reciprocal of a normal binary32 number, correctly rounded,
// from the hardware estimates (error below one ulp).  Square root: sqrt(x)
// rounds to one of s-, s, s+ around the estimate s; s- is the rounding when
// sqrt(x) lies below the midpoint of s- and s, i.e. when x <= s- * s (the
// products of two floats and x are multiples of a common unit far above the
// square of the half-ulp, so comparing with s- * s instead of the midpoint's
// square decides the same way), and likewise s+ when x > s+ * s; each
// comparison is the sign of one exact-signed FMA residual.  Reciprocal: one
// Newton-Raphson step y + y (1 - x y) with FMA.  Both are checked against
// IEEE sqrtf and 1.0f / x on the device for every normal input
// (pbf_check_fast_math).
__device__ __forceinline__ float sqrt_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = __uint_as_float(__float_as_uint(s) - 1u);
    const float above = __uint_as_float(__float_as_uint(s) + 1u);
    float r = fmaf(-below, s, x) <= 0.0f ? below : s;
    r = fmaf(-above, s, x) > 0.0f ? above : r;
    return r;
}
__device__ __forceinline__ float rcp_normal(float x) {
    const float y = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, y, 1.0f), y, y);
}

// A pair whose binary32 r^2 is below 2^-96 (|r| < 2^-48) is coincident: its
// length and direction are 0 (h - |r| rounds to h there anyway).  Above it
// the square root estimate is within an ulp, so |r| and 1/|r| are the IEEE
// values.  A length from a squared norm follows the same rule.
constexpr float kTinySq = 0x1p-96f;
struct PairLen {
    float rl, inv;   // |r|, 1/|r|
};
__device__ __forceinline__ PairLen pair_len(float r2) {
    const bool ok = r2 >= kTinySq;
    const float rl = sqrt_normal(r2);
    const float inv = rcp_normal(rl);
    return {ok ? rl : 0.0f, ok ? inv : 0.0f};
}
__device__ __forceinline__ float norm_of_sq(float s2) { return s2 >= kTinySq ? sqrt_normal(s2) : 0.0f; }

// The forward sweeps' form: |r| is not selected, only 1/|r|.  Below 2^-96
// the estimate is below 2^-47, and the sweeps use |r| only in h - |r|, which
// rounds to h there for every h >= 2^-20 (the smallest h pbf_create
// accepts), as with |r| = 0; a NaN r^2 gives a NaN length, whose pair the
// sweeps select out.
__device__ __forceinline__ PairLen pair_len_sweep(float r2) {
    const bool ok = r2 >= kTinySq;
    const float rl = sqrt_normal(r2);
    const float inv = rcp_normal(rl);
    return {rl, ok ? inv : 0.0f};
}

// Selection of a pair's contributions as arithmetic: lane_mask(keep) is all
// ones when keep and zero otherwise, as a value the compiler cannot see
// through (an empty asm statement), and masked(x, m) is x or +0 (also for a
// NaN x).  The pair arithmetic then runs for every pair with no per-pair
// branch around it: a branch that every lane of a wave takes anyway costs
// exec-mask updates, zeroed phi registers and reloads of the scene constants.
__device__ __forceinline__ uint32_t lane_mask(bool keep) {
    uint32_t m = keep ? 0xFFFFFFFFu : 0u;
    asm("" : "+v"(m));
    return m;
}
__device__ __forceinline__ float masked(float x, uint32_t m) { return __uint_as_float(__float_as_uint(x) & m); }
// a where m is all ones, else b (one bit select): a sum that a skipped pair
// leaves as it is.
__device__ __forceinline__ float picked(float a, float b, uint32_t m) {
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// Fused multiply-adds at fixed points, the same in the oracle: the squared
// distance of a pair is fmaf(rz, rz, fmaf(ry, ry, rx * rx)), and the sweeps
// fold each pair's terms into their sums with one fmaf where a product is
// added.  Everything else stays unfused (the build has contraction off).
// The reference compiles its kernels with -cl-mad-enable, which lets its
// compiler fuse any of these.
__device__ __forceinline__ float dist2(float rx, float ry, float rz) { return fmaf(rz, rz, fmaf(ry, ry, rx * rx)); }

__device__ __forceinline__ int circle_count(uint32_t packed, int q) { return (packed >> (6 * q)) & 63; }

// Number of friend circles a sweep processes.  The reference skips circle q
// when (float)processed/total > threshold (compute_scaling.cl:50, 0.6;
// compute_delta.cl:100, 0.5); once one circle is skipped `processed` stops
// growing, so the processed circles form a prefix.  For counts <= 250 the
// fp32 test equals the exact rational one used here.  Returns the prefix
// length; *m receives the number of friends in it.
template <int NUM, int DEN>  // threshold = NUM/DEN
__device__ __forceinline__ int circles_processed(uint32_t packed, int* m) {
    int total = 0;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) total += circle_count(packed, q);
    int done = 0, L = 0;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) {
        if (L == q && !(DEN * done > NUM * total)) { done += circle_count(packed, q); L = q + 1; }
    }
    *m = done;
    return L;
}

// Friends lists: particle i owns, for each of its five circles, kSeg slots
// holding its circle-c friends in the reference's scan order -- the
// reference's layout (counts[5][n], indices[5][50][n],
// build_friends_list.cl:1-21) blocked per 64 consecutive particles and
// grouped in quads: quad q (slots 4q..4q+3) of circle c of the block's 64
// particles is one contiguous 1 KB row [(i>>6)][c][q][lane] of uint4.  The
// build stores a quad when it fills (16-byte stores that land, for all the
// lanes of a wave, in the same few rows) and the regrouping reads it back
// with one load.
constexpr int kSegQuads = (kPerCircle + 3) / 4;   // 13
constexpr int kSeg = 4 * kSegQuads;               // 52 slots per circle
__device__ __host__ __forceinline__ size_t friend_quad(uint32_t i, int circle, int q) {
    return (((size_t)(i >> 6) * kCircles + (size_t)circle) * kSegQuads + (size_t)q) * 64 + (i & 63u);
}
__device__ __host__ __forceinline__ size_t friend_slot(uint32_t i, int circle, int f) {   // in uint32 units
    return friend_quad(i, circle, f >> 2) * 4 + (size_t)(f & 3);
}
constexpr __device__ __host__ size_t friend_slots(size_t n) { return ((n + 63) / 64) * 64 * kCircles * kSeg; }

// Sweep lists: the friends a sweep processes (the 0.6-threshold prefix, in
// the reference's order) regrouped per block of 64 consecutive particles in
// quads: quad q of the block holds friends 4q..4q+3 of each of its 64
// particles as one uint4 per particle, so a wave's load of one quad is a
// contiguous 1 KB and gives every lane four friends.
constexpr int kQuads = (kCircles * kPerCircle + 3) / 4;
__device__ __host__ __forceinline__ size_t sweep_quad(uint32_t i, int q) {
    return ((size_t)(i >> 6) * kQuads + (size_t)q) * 64 + (i & 63u);
}

// Circle 0 is processed by every sweep (0 processed < any threshold), so the
// build writes it straight to the front of the sweep quads; circles 1..4 go
// to the circle quads and only the processed ones are appended after it.
// Friend f of circle c of particle i, from the two (as uint32 arrays):
__device__ __host__ __forceinline__ uint32_t friend_at(const uint32_t* list, const uint32_t* sweep, uint32_t i,
                                                       int c, int f) {
    return c == 0 ? sweep[sweep_quad(i, f >> 2) * 4 + (size_t)(f & 3)] : list[friend_slot(i, c, f)];
}

// Walks the first m entries of particle i's list in quad layout, in two
// phases per quad: load(j) issues the gathers of the four friends, the next
// quad's indices are fetched while they are in flight, then add(c) folds them
// in list order (the reference's order).
template <class Load, class Add>
__device__ __forceinline__ void walk_quads(uint32_t i, int m, const uint4* __restrict__ quads, Load&& load,
                                           Add&& add) {
    if (m <= 0) return;
    const int nfull = m >> 2, rest = m & 3;
    uint4 cur = quads[sweep_quad(i, 0)];
    for (int q = 0; q < nfull; ++q) {
        const auto c0 = load(cur.x);
        const auto c1 = load(cur.y);
        const auto c2 = load(cur.z);
        const auto c3 = load(cur.w);
        // quad q + 1 <= m/4 <= 62 exists in every particle's 63 (unused
        // entries are never gathered)
        cur = quads[sweep_quad(i, q + 1)];
        add(c0);
        add(c1);
        add(c2);
        add(c3);
    }
    if (rest) {
        auto c0 = load(cur.x);
        auto c1 = c0, c2 = c0;
        if (rest > 1) c1 = load(cur.y);
        if (rest > 2) c2 = load(cur.z);
        add(c0);
        if (rest > 1) add(c1);
        if (rest > 2) add(c2);
    }
}

// The same walk one quad deeper: the gathers of quad q + 1 are issued before
// quad q is folded in, so a wave keeps its next gathers in flight while it
// computes.  Entries past m are replaced by i, so every gather stays in
// bounds.  add(g, valid) is called for all four entries of every quad, in
// list order, with valid false for the entries past m: the accumulators fold
// those in as +0, so the loop body has no per-entry branch.  The loop is
// unrolled by two so that the two quads in flight keep their registers (no
// copies between iterations).
template <class Load, class Add>
__device__ __forceinline__ void walk_quads_ahead(uint32_t i, int m, const uint4* __restrict__ quads, Load&& load,
                                                 Add&& add) {
    if (m <= 0) return;
    const int nq = (m + 3) >> 2;
    auto own = [&](uint4 u, int k) -> uint4 {
        const int b = 4 * k;
        return make_uint4(b < m ? u.x : i, b + 1 < m ? u.y : i, b + 2 < m ? u.z : i, b + 3 < m ? u.w : i);
    };
    const uint4 first = own(quads[sweep_quad(i, 0)], 0);
    auto a0 = load(first.x);
    auto a1 = load(first.y);
    auto a2 = load(first.z);
    auto a3 = load(first.w);
    uint4 next = quads[sweep_quad(i, 1)];
    int q = 0;
    for (;;) {
        // quad q is in a0..a3, the indices of quad q + 1 in next
        const uint4 u = own(next, q + 1);
        const auto b0 = load(u.x);
        const auto b1 = load(u.y);
        const auto b2 = load(u.z);
        const auto b3 = load(u.w);
        next = quads[sweep_quad(i, min(q + 2, kQuads - 1))];   // a particle's 63 quads
        add(a0, true);
        add(a1, 4 * q + 1 < m);
        add(a2, 4 * q + 2 < m);
        add(a3, 4 * q + 3 < m);
        if (++q >= nq) break;
        const uint4 w = own(next, q + 1);
        a0 = load(w.x);
        a1 = load(w.y);
        a2 = load(w.z);
        a3 = load(w.w);
        next = quads[sweep_quad(i, min(q + 2, kQuads - 1))];
        add(b0, true);
        add(b1, 4 * q + 1 < m);
        add(b2, 4 * q + 2 < m);
        add(b3, 4 * q + 3 < m);
        if (++q >= nq) break;
    }
}

template <int L_NUM, int L_DEN, bool AHEAD = false, class Load, class Add>
__device__ __forceinline__ void walk_friends(uint32_t i, uint32_t packed, const uint4* __restrict__ quads, Load&& load,
                                             Add&& add) {
    int m;
    (void)circles_processed<L_NUM, L_DEN>(packed, &m);
    if constexpr (AHEAD)
        walk_quads_ahead(i, m, quads, load, add);
    else
        walk_quads(i, m, quads, load, add);
}

// A tile's neighbourhood in LDS.  In the cell-sorted (Morton) order about
// four in five friends of the particles of a 256-particle tile lie within
// kNearPad entries of the tile; a sweep copies that window of the arrays it
// gathers into LDS, and a gather inside the window reads LDS instead of the
// L1 cache.  The values are the same either way.  The compiler turns "LDS
// inside the window, global outside" into one generic (flat) load through a
// selected address, which serves a wave's mixed lanes in one instruction; an
// LDS read followed by an exec-masked global read measured 10-40% slower.
constexpr uint32_t kNearPad = 256;
// Windows pay from this many tiles per CU on.
constexpr uint32_t kNearMinTilesPerCu = 4;
template <int BS>
struct NearWindow {
    static constexpr uint32_t kLen = BS + 2 * kNearPad;
    uint32_t base = 0, len = 0;
    // entries [base, base + len) of the tile be (the array holds n); an
    // empty window when `on` is 0: on a grid of one or two workgroups per CU
    // the copy only adds latency (a 65,536-particle scene ran 5% slower)
    __device__ __forceinline__ NearWindow(const uint4& be, uint32_t n, uint32_t on) {
        base = be.y > kNearPad ? be.y - kNearPad : 0u;
        len = on ? min(kLen, n - base) : 0u;
    }
    __device__ __forceinline__ void stage(float4* __restrict__ dst, const float4* __restrict__ src) const {
        for (uint32_t t = threadIdx.x; t < len; t += BS) dst[t] = src[base + t];
    }
    // entry j of src, from its copy win when the window holds it
    __device__ __forceinline__ float4 get(const float4* win, const float4* __restrict__ src, uint32_t j) const {
        const uint32_t o = j - base;
        if (o < len) return win[o];
        return src[j];
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs of the chip; this
// maps dispatch index b of nb workgroups to a tile so that each XCD takes one
// contiguous run of tiles.  Neighbouring particles of a tile then sit in the
// same XCD's L2 as the tile.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Box wall planes: BouncePointQuad (compute_delta.cl:1-42) for an infinite
// axis-aligned plane with mask 1.  `hit` records whether the plane moved the
// point (the adjoint replays the same branch).
__device__ __forceinline__ float wall_lo(float f, float prev, float lo, float edge, bool& hit) {
    const float pa = lo + edge;
    const float dp = f - pa;
    if (dp > 0.0f) return f;
    const float nv = prev - f;
    if (fabsf(dp) > edge + nv) return f;
    hit = true;
    return pa;
}

__device__ __forceinline__ float wall_hi(float f, float prev, float hi, float edge, bool& hit) {
    const float pa = hi - edge;
    const float dp = f - pa;
    if (dp < 0.0f) return f;
    const float nv = f - prev;
    if (fabsf(dp) > edge + nv) return f;
    hit = true;
    return pa;
}

// Edge offset of the force zone (compute_delta.cl:165-166 without its
// scene-scale "1.0 +" term).
__device__ __forceinline__ float wall_edge(float fx, float fy, float fz, const SceneConsts& c) {
    const float nx = noise_sin(fx * 5.0f / c.h);
    const float ny = noise_sin(fy * 5.0f / c.h);
    const float nz = noise_sin(fz * 5.0f / c.h);
    return (3.0f + nx + ny + nz) * c.h * 0.03f;
}

// BouncePointQuad (compute_delta.cl:1-42): a point inside the quad's
// parallelogram (s, t in [0, 1]), behind its plane offset by `edge` along n
// and within the force zone (|F - plane point| <= edge + n.(prev - F)) moves
// onto the offset plane where the mask texel ((int)(t W), mask_y +
// (int)(s mask_h)), clamped to the image as CLK_ADDRESS_CLAMP_TO_EDGE
// (utilities.cl:89), is set.  Returns whether it moved the point.
__device__ __forceinline__ bool bounce_quad(const QuadConsts& q, const float4& prev, float& fx, float& fy, float& fz,
                                            float edge, const SceneConsts& c) {
    const float Dx = q.b0 - fx, Dy = q.b1 - fy, Dz = q.b2 - fz;
    const float e = q.e10 * Dx + q.e11 * Dy + q.e12 * Dz;
    const float d = q.e00 * Dx + q.e01 * Dy + q.e02 * Dz;
    const float s = q.invdet * (q.bb * e - q.cc * d);
    const float t = q.invdet * (q.bb * d - q.aa * e);
    if ((s < 0.0f) || (s > 1.0f) || (t < 0.0f) || (t > 1.0f)) return false;
    const float px = q.b0 + q.e00 * s + q.e10 * t + q.n0 * edge;
    const float py = q.b1 + q.e01 * s + q.e11 * t + q.n1 * edge;
    const float pz = q.b2 + q.e02 * s + q.e12 * t + q.n2 * edge;
    const float dx = fx - px, dy = fy - py, dz = fz - pz;
    if (q.n0 * dx + q.n1 * dy + q.n2 * dz > 0.0f) return false;
    const float nv = q.n0 * (prev.x - fx) + q.n1 * (prev.y - fy) + q.n2 * (prev.z - fz);
    if (sqrtf(dx * dx + dy * dy + dz * dz) > edge + nv) return false;
    int32_t mx = (int32_t)(t * (float)c.mask_w);
    int32_t my = (int32_t)((float)q.mask_y + s * (float)q.mask_h);
    mx = mx < 0 ? 0 : (mx >= (int32_t)c.mask_w ? (int32_t)c.mask_w - 1 : mx);
    my = my < 0 ? 0 : (my >= (int32_t)c.mask_rows ? (int32_t)c.mask_rows - 1 : my);
    const uint32_t stride = (c.mask_w + 31u) >> 5;
    if (!((c.mask[(size_t)my * stride + ((uint32_t)mx >> 5)] >> ((uint32_t)mx & 31u)) & 1u)) return false;
    fx = px;
    fy = py;
    fz = pz;
    return true;
}

// The delta stage's boundaries (compute_delta.cl:163-182): the force-zone
// edge from the unconstrained position, the box planes in the order x-lo,
// x-hi, y-lo, y-hi, z-lo, z-hi, then the collision quads in order with edge
// edge_base + edge.  Returns which moved the point: bit 2a (lo) and 2a+1 (hi)
// of axis a, bit 6+k for quad k.
__device__ __forceinline__ uint32_t apply_walls(float& fx, float& fy, float& fz, const float4& prev,
                                                const SceneConsts& c) {
    const float edge = wall_edge(fx, fy, fz, c);
    uint32_t hits = 0;
    if (c.walls) {
        bool h0 = false, h1 = false, h2 = false, h3 = false, h4 = false, h5 = false;
        fx = wall_lo(fx, prev.x, c.lo0, edge, h0);
        fx = wall_hi(fx, prev.x, c.hi0, edge, h1);
        fy = wall_lo(fy, prev.y, c.lo1, edge, h2);
        fy = wall_hi(fy, prev.y, c.hi1, edge, h3);
        fz = wall_lo(fz, prev.z, c.lo2, edge, h4);
        fz = wall_hi(fz, prev.z, c.hi2, edge, h5);
        hits = (uint32_t)h0 | ((uint32_t)h1 << 1) | ((uint32_t)h2 << 2) | ((uint32_t)h3 << 3) |
                    ((uint32_t)h4 << 4) | ((uint32_t)h5 << 5);
    }
    const float eq = c.edge_base + edge;
    for (uint32_t k = 0; k < c.quad_count; ++k)
        if (bounce_quad(c.quads[k], prev, fx, fy, fz, eq, c)) hits |= 1u << (6 + k);
    return hits;
}

}  // namespace pbf
