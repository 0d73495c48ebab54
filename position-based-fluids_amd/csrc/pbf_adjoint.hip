// pbf_adjoint.hip -- reverse-mode adjoint of the substep (include/pbf/pbf_adjoint.h).
//
// Forward of one substep s, in the cell-sorted order of s (perm_s), with the
// friend lists L05/L06 (the 0.5 / 0.6 circle prefixes) frozen:
//   P0      = X + dt (V + dt g)                      (predictPositions, gathered)
//   k < K:  lambda_k = Lambda(P_k)                   (computeScaling)
//           P_k+1    = Wall(P_k + Dp(P_k, lambda_k)) (computeDelta + walls)
//   Vu = (P_K - Xs) / dt                             (updateVelocities)
//   Vx, Om = Xsph(P_K, Vu; c)                        (applyViscosity, Jacobi)
//   V' = Vx + dt eps N(P_K, Om) x Om                 (applyVorticity)
//   X' = P_K
// The backward walks it in reverse with one kernel per sweep.  A pair term
// f(p_i - p_j) of particle i's sum gives gradient to both i and j; instead of
// scattering to j with atomics, every kernel gathers for its particle m both
// its own pairs (its list) and the pairs in which m is the friend (the
// transposed list T_m: the particles i whose processed prefix holds m).  Each
// thread then owns its outputs and the sums are deterministic.
//
// Transposed lists.  Without friend drops (the 50 cap) and without duplicate
// bucket visits, i lists m in circle c iff m lists i in circle c: r^2 is
// computed from an exactly antisymmetric difference and the 27-cell
// neighbourhoods are symmetric.  So T_m = {i in m's full list, circle c :
// c < circles processed by i}, built when the substep is recorded
// (k_transpose).  A substep that dropped friends or reached a bucket twice
// is transposed by searching instead (k_transpose_all: the candidates in m's
// neighbour cells whose prefixes list m); the backward fails rather than
// differentiate a different function if that search misses an entry or a
// list outgrows its slot.
//
// Per-pair derivatives, r = p_i - p_j, q = |r|, d = h^2 - q^2:
//   psi(r) = r (h-q)^2 / q            (the spiky direction without SPIKY)
//   J(r) u = w u + (w'/q)(r.u) r,  w = (h-q)^2/q, w' = -(h-q)(h+q)/q^2
//   grad d^3 = -6 d^2 r
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <vector>

#include "../../include/pbf/pbf_adjoint.h"
#include "pbf_internal.h"

namespace pbf {
namespace {

constexpr int kAdjThreads = 256;   // workgroup of every adjoint kernel (block table bt[2])
constexpr int kParts = 8;          // per-tile partial sums: [0] viscosity, [1] vorticity, [2..4] gravity
constexpr uint32_t kDupFlag = 1u;  // a particle's 27-cell neighbourhood visits one bucket twice
constexpr uint32_t kCapFlag = 2u;  // a transposed list would exceed the slot's 4 * kQuads entries
// flags words: [0] kDupFlag | kCapFlag, [2..3] entries placed by k_transpose_all,
// [4..5] entries of the prefixes it transposed (u64 each)
constexpr int kFlagWords = 8;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(const float4& a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
__device__ __forceinline__ V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float4 f4(V3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }

// Geometry of a pair, evaluated whether or not it is in range (the sweeps
// below are branch-free: out-of-range pairs contribute a selected zero).
struct Pair {
    V3 r;
    float r2, q, inv, hr, d;
    bool in;
};

__device__ __forceinline__ Pair pair_of(const float4& a, const float4& b, const SceneConsts& c) {
    Pair g;
    g.r = {a.x - b.x, a.y - b.y, a.z - b.z};
    g.r2 = dist2(g.r.x, g.r.y, g.r.z);   // as the forward sweeps
    g.in = g.r2 < c.h2;
    const PairLen L = pair_len(g.r2);   // as the forward sweeps: 0 and 0 for a coincident pair
    g.q = L.rl;
    g.inv = L.inv;
    g.hr = c.h - g.q;
    g.d = c.h2 - g.r2;
    return g;
}

__device__ __forceinline__ V3 psi(const Pair& g) { return g.r * (g.hr * g.hr * g.inv); }

__device__ __forceinline__ V3 jpsi(const Pair& g, V3 u, const SceneConsts& c) {
    const float w = g.hr * g.hr * g.inv;
    const float wq = -(g.hr * (c.h + g.q)) * (g.inv * g.inv * g.inv);   // w'/q
    return u * w + g.r * (wq * dot(g.r, u));
}

__device__ __forceinline__ V3 sel(bool m, V3 v) {
    return {m ? v.x : 0.0f, m ? v.y : 0.0f, m ? v.z : 0.0f};
}

struct G2 {
    float4 a, b;
};
struct G4 {
    float4 a, b, c, d;
};
struct G2f {
    float4 a, b;
    float f;
};

// Transposed entries: count of the 0.6 list in the low half, of its 0.5
// prefix in the high half.
__device__ __forceinline__ int t06(uint32_t tc) { return (int)(tc & 0xFFFFu); }
__device__ __forceinline__ int t05(uint32_t tc) { return (int)(tc >> 16); }

// Sum over the workgroup of NV values per thread (fixed order), written by
// thread 0 to out[0..NV).
template <int NV>
__device__ __forceinline__ void block_sum(float (&v)[NV], float* out) {
    __shared__ float s[kAdjThreads / 64][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            float a = 0.0f;
            for (int q = 0; q < kAdjThreads / 64; ++q) a += s[q][k];
            out[k] = a;
        }
    }
}

// Maximum over the workgroup (NaN propagates), written by thread 0 to *out.
__device__ __forceinline__ void block_max(float v, float* out) {
    __shared__ float s[kAdjThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_xor(v, o, 64);
        v = (t > v || t != t) ? t : v;
    }
    if (lane == 0) s[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f;
        for (int q = 0; q < kAdjThreads / 64; ++q) a = (s[q] > a || s[q] != s[q]) ? s[q] : a;
        *out = a;
    }
}

// Transpose of the boundaries (compute_delta.cl:1-42,163-182 as applied in
// apply_walls), in reverse order.  A quad k that moved the point maps the
// post-wall position F' to B + E0 s + E1 t + n (edge_base + edge(F)), with
// s, t the linear parameters of F' (so g passes through the projection's
// transpose, and n.g goes to the edge); a coordinate a plane moved becomes
// lo + edge(F) or hi - edge(F), with edge(F) = (3 + sum_c sin(5 F_c / h))
// 0.03 h; the others pass through.
__device__ __forceinline__ V3 wall_vjp(const float4& fut, V3 g, const SceneConsts& c) {
    const uint32_t hits = __float_as_uint(fut.w);
    if (!hits) return g;
    float eg = 0.0f;   // dL/d edge through the quads
    for (uint32_t k = c.quad_count; k-- > 0;) {
        if (!((hits >> (6 + k)) & 1u)) continue;
        const QuadConsts& q = c.quads[k];
        const float gs = q.e00 * g.x + q.e01 * g.y + q.e02 * g.z;
        const float gt = q.e10 * g.x + q.e11 * g.y + q.e12 * g.z;
        eg += q.n0 * g.x + q.n1 * g.y + q.n2 * g.z;
        const float ge = q.invdet * (q.bb * gs - q.aa * gt);   // dL/de, e = E1.(B - F')
        const float gd = q.invdet * (q.bb * gt - q.cc * gs);   // dL/dd, d = E0.(B - F')
        g = {-(q.e10 * ge + q.e00 * gd), -(q.e11 * ge + q.e01 * gd), -(q.e12 * ge + q.e02 * gd)};
    }
    const float sx = (hits & 2u) ? -1.0f : ((hits & 1u) ? 1.0f : 0.0f);
    const float sy = (hits & 8u) ? -1.0f : ((hits & 4u) ? 1.0f : 0.0f);
    const float sz = (hits & 32u) ? -1.0f : ((hits & 16u) ? 1.0f : 0.0f);
    V3 o = {sx != 0.0f ? 0.0f : g.x, sy != 0.0f ? 0.0f : g.y, sz != 0.0f ? 0.0f : g.z};
    const float e = sx * g.x + sy * g.y + sz * g.z + eg;
    const float k = 0.15f * e;   // d edge / dF_c = 0.03 h * 5/h * cos(5 F_c / h)
    o.x += k * cosf(fut.x * 5.0f / c.h);
    o.y += k * cosf(fut.y * 5.0f / c.h);
    o.z += k * cosf(fut.z * 5.0f / c.h);
    return o;
}

// ---------------------------------------------------------------------------
// Recording
// ---------------------------------------------------------------------------

// T_m from m's full friend list (all circles: circle 0 from the front of
// the sweep quads, the others from the circle quads): the
// friends i whose processed prefix contains m's circle, 0.5 members first.
__global__ __launch_bounds__(kAdjThreads) void k_transpose(const SceneConsts* __restrict__ scenes,
                                                          const uint4* __restrict__ blocks,
                                                          const float4* __restrict__ xs,
                                                          const uint32_t* __restrict__ cnt,
                                                          const uint32_t* __restrict__ nb_list,
                                                          const uint32_t* __restrict__ sweep,
                                                          const uint2* __restrict__ cells,
                                                          uint32_t* __restrict__ tcnt, uint4* __restrict__ tquads,
                                                          uint32_t* flags) {
    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];
    const SceneConsts& c = scenes[be.x];
    const uint32_t m = be.y + threadIdx.x;
    if (m >= be.z) return;
    const uint32_t pm = cnt[m];
    uint32_t acc[4] = {0u, 0u, 0u, 0u};
    int na = 0, qo = 0, n05 = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = 0; q < kCircles; ++q) {
            const int cq = circle_count(pm, q);
            for (int f = 0; f < cq; ++f) {
                const uint32_t i = friend_at(nb_list, sweep, m, q, f);
                const uint32_t pi = cnt[i];
                int dummy;
                const int l05 = circles_processed<1, 2>(pi, &dummy);
                const int l06 = circles_processed<3, 5>(pi, &dummy);
                const bool take = pass == 0 ? (q < l05) : (q >= l05 && q < l06);
                if (take) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] = (u == na) ? i : acc[u];
                    if (++na == 4) {
                        tquads[sweep_quad(m, qo++)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
                        na = 0;
                    }
                }
            }
        }
        if (pass == 0) n05 = 4 * qo + na;
    }
    if (na) tquads[sweep_quad(m, qo)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
    tcnt[m] = (uint32_t)(4 * qo + na) | ((uint32_t)n05 << 16);

    // Symmetry needs every bucket visited once.  With gbs = 2^b, b >= 6, the
    // 27 neighbour cells always hash apart (their coordinates differ mod 4).
    if (!(c.gbs_mask != 0u && c.gbs >= 64u)) {
        const float4 p = xs[m];
        const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);
        uint32_t keys[27];
        int nk = 0;
        bool dup = false;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const uint32_t k = cell_hash(cx + dx, cy + dy, cz + dz, c);
                    if (cells[k].x == kEmpty) continue;
                    for (int t = 0; t < nk; ++t) dup |= keys[t] == k;
                    keys[nk++] = k;
                }
        if (dup) atomicOr(flags, kDupFlag);
    }
}

// Transposition of a substep whose lists need not be symmetric (friends
// dropped at the 50 cap, or a bucket reached twice): T_m holds every i whose
// processed prefix lists m, as often as it lists m, the 0.5-prefix entries
// first.  A particle that lists m is within h of it, so it lies in one of m's
// 27 neighbour cells: m scans their distinct buckets in order, keeps the
// candidates that pass the friends build's own binary32 test (r^2 < h^2, r =
// p_i - p_m) and counts m in each one's prefix.  Runs only when this
// substep's dropped-friends snapshot moved (note[-1], note[0]) or a repeated
// bucket was seen, and then replaces k_transpose's lists.  It adds the
// entries it placed and the prefix lengths of all particles to flags[2..5];
// the two agree unless a particle outside the neighbour cells listed one
// (only possible at the rounding of a cell boundary), and a list longer
// than the slot sets kCapFlag.
__global__ __launch_bounds__(kAdjThreads) void k_transpose_all(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ xs,
    const uint32_t* __restrict__ cnt, const uint4* __restrict__ sweep, const uint2* __restrict__ cells, uint32_t n,
    uint32_t* __restrict__ tcnt, uint4* __restrict__ tquads, const unsigned long long* __restrict__ note,
    uint32_t* flags) {
    if (note[0] == note[-1] && !(flags[0] & kDupFlag)) return;
    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];
    const SceneConsts& c = scenes[be.x];
    const uint32_t m = be.y + threadIdx.x;
    unsigned long long placed = 0, listed = 0;
    if (m < be.z) {
        const float4 p = xs[m];
        int own;
        circles_processed<3, 5>(cnt[m], &own);
        listed = (unsigned long long)own;
        const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);
        uint32_t keys[27];
        int nk = 0;
        for (int k = 0; k < 27; ++k) {
            const uint32_t key = cell_hash(cx + k / 9 - 1, cy + (k / 3) % 3 - 1, cz + k % 3 - 1, c);
            bool seen = false;
            for (int t = 0; t < nk; ++t) seen |= keys[t] == key;
            if (!seen && cells[key].x != kEmpty) keys[nk++] = key;
        }
        uint32_t acc[4] = {0u, 0u, 0u, 0u};
        int na = 0, qo = 0, n05 = 0;
        bool full = false;
        for (int pass = 0; pass < 2; ++pass) {
            for (int t = 0; t < nk; ++t) {
                const uint2 b = cells[keys[t]];
                const uint32_t e = min(b.y, n - 1);
                for (uint32_t i = b.x; i <= e; ++i) {
                    if (i == m) continue;
                    const float4 q = xs[i];
                    // the forward build's friend test (the sign of r does not change r^2)
                    if (!(dist2(q.x - p.x, q.y - p.y, q.z - p.z) < c.h2)) continue;
                    int m05, m06;
                    const uint32_t pi = cnt[i];
                    circles_processed<1, 2>(pi, &m05);
                    circles_processed<3, 5>(pi, &m06);
                    const int lo = pass == 0 ? 0 : m05, hi = pass == 0 ? m05 : m06;
                    for (int k = lo & ~3; k < hi; k += 4) {
                        const uint4 qd = sweep[sweep_quad(i, k >> 2)];
                        const uint32_t v4[4] = {qd.x, qd.y, qd.z, qd.w};
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (k + u < lo || k + u >= hi || v4[u] != m) continue;
                            if (4 * qo + na >= 4 * kQuads) {
                                full = true;
                                continue;
                            }
#pragma unroll
                            for (int w = 0; w < 4; ++w) acc[w] = (w == na) ? i : acc[w];
                            ++placed;
                            if (++na == 4) {
                                tquads[sweep_quad(m, qo++)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
                                na = 0;
                            }
                        }
                    }
                }
            }
            if (pass == 0) n05 = 4 * qo + na;
        }
        if (na) tquads[sweep_quad(m, qo)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
        tcnt[m] = (uint32_t)(4 * qo + na) | ((uint32_t)n05 << 16);
        if (full) atomicOr(flags, kCapFlag);
    }
    // one add per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        placed += __shfl_xor(placed, o, 64);
        listed += __shfl_xor(listed, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && (placed | listed)) {
        atomicAdd(reinterpret_cast<unsigned long long*>(flags + 2), placed);
        atomicAdd(reinterpret_cast<unsigned long long*>(flags + 4), listed);
    }
}

// Snapshot of the dropped-friends counter (err words 2..3).
__global__ void k_note_drops(const uint32_t* __restrict__ err, unsigned long long* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *out = (unsigned long long)err[2] | ((unsigned long long)err[3] << 32);
}

__global__ void k_iota(uint32_t* __restrict__ ids, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ids[i] = i;
}

// ids after a substep: the identity of the particle at sorted slot i.
__global__ void k_ids_gather(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ ids_in,
                             uint32_t* __restrict__ ids_out, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint32_t p = perm[i];
        ids_out[i] = p < n ? ids_in[p] : 0u;
    }
}

// Sorted state -> identity order.
__global__ void k_scatter_state(const uint32_t* __restrict__ ids, const float4* __restrict__ x,
                                const float4* __restrict__ v, float4* __restrict__ xo, float4* __restrict__ vo,
                                uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = ids[i];
    if (d >= n) return;
    if (xo) xo[d] = x[i];
    if (vo) vo[d] = v[i];
}

// Identity-order gradients -> sorted order of the last substep (w cleared).
__global__ void k_gather_grads(const uint32_t* __restrict__ ids, const float4* __restrict__ gx_in,
                               const float4* __restrict__ gv_in, float4* __restrict__ gx, float4* __restrict__ gv,
                               uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = ids[i];
    const float4 a = (gx_in && d < n) ? gx_in[d] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 b = (gv_in && d < n) ? gv_in[d] : make_float4(0.f, 0.f, 0.f, 0.f);
    gx[i] = make_float4(a.x, a.y, a.z, 0.0f);
    gv[i] = make_float4(b.x, b.y, b.z, 0.0f);
}

__global__ void k_fill(float* __restrict__ p, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Backward kernels, in the order they run for one substep
// ---------------------------------------------------------------------------

// applyVorticity, own part (apply_vorticity.cl:1-87): recomputes eta and
// N = eta/|eta| of particle i, gives dL/deps, the direct dL/dOm_i, and the
// pair coefficient zeta_i = -SPIKY dL/deta_i.
__global__ __launch_bounds__(kAdjThreads) void k_vort_b1(const SceneConsts* __restrict__ scenes,
                                                        const uint4* __restrict__ blocks,
                                                        const float4* __restrict__ P, const float4* __restrict__ om,
                                                        const uint32_t* __restrict__ cnt,
                                                        const uint4* __restrict__ quads,
                                                        const float4* __restrict__ gv, const float* __restrict__ scale,
                                                        float4* __restrict__ z, float4* __restrict__ gom,
                                                        float* __restrict__ part, uint32_t n, uint32_t near_on) {
    __shared__ float4 s_P[NearWindow<kAdjThreads>::kLen], s_om[NearWindow<kAdjThreads>::kLen];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint4 be = blocks[tile];
    const SceneConsts& c = scenes[be.x];
    const NearWindow<kAdjThreads> near(be, n, near_on);   // the tile's neighbourhood in LDS
    near.stage(s_P, P);
    near.stage(s_om, om);
    __syncthreads();
    const uint32_t i = be.y + threadIdx.x;
    float ge[1] = {0.0f};
    if (i < be.z) {
        const float4 p = P[i];
        V3 eta = {0.f, 0.f, 0.f};
        walk_friends<1, 2>(
            i, cnt[i], quads, [&](uint32_t j) { return G2{near.get(s_P, P, j), near.get(s_om, om, j)}; },
            [&](const G2& t) {
                const Pair g = pair_of(p, t.a, c);
                const V3 oj = v3(t.b);
                eta += sel(g.in && fabsf(t.a.w) > 1e-8f, psi(g) * norm_of_sq(dot(oj, oj)));
            });
        eta = eta * -c.spiky;
        const float l = sqrtf(dot(eta, eta));
        const V3 N = l > 0.0f ? eta * (1.0f / l) : eta;
        const V3 oi = v3(om[i]);
        const V3 g = v3(gv[i]) * scale[be.x];
        ge[0] = c.dt * dot(cross(N, oi), g);
        const float u = c.vorticity * c.dt;
        const V3 gN = cross(oi, g) * u;
        const V3 gOd = cross(g, N) * u;
        const V3 geta = l > 0.0f ? (gN - N * dot(N, gN)) * (1.0f / l) : gN;
        z[i] = f4(geta * -c.spiky);
        gom[i] = f4(gOd);
    }
    block_sum<1>(ge, part + (size_t)tile * kParts + 1);
}

// applyVorticity, pair part: dL/dP_K and dL/dOm from the pairs of m's list
// and of its transposed list.
__global__ __launch_bounds__(kAdjThreads) void k_vort_b2(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ P,
    const float4* __restrict__ om, const uint32_t* __restrict__ cnt, const uint4* __restrict__ quads,
    const uint32_t* __restrict__ tcnt, const uint4* __restrict__ tquads, const float4* __restrict__ z,
    const float* __restrict__ scale, float4* __restrict__ gom, float4* __restrict__ gx, uint32_t n, uint32_t near_on) {
    __shared__ float4 s_P[NearWindow<kAdjThreads>::kLen], s_om[NearWindow<kAdjThreads>::kLen];
    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];
    const SceneConsts& c = scenes[be.x];
    const NearWindow<kAdjThreads> near(be, n, near_on);   // the tile's neighbourhood in LDS
    near.stage(s_P, P);
    near.stage(s_om, om);
    __syncthreads();
    const uint32_t m = be.y + threadIdx.x;
    if (m >= be.z) return;
    const float4 p = P[m];
    const V3 zm = v3(z[m]);
    V3 gP = {0.f, 0.f, 0.f};
    walk_friends<1, 2>(
        m, cnt[m], quads, [&](uint32_t j) { return G2{near.get(s_P, P, j), near.get(s_om, om, j)}; },
        [&](const G2& t) {
            const Pair g = pair_of(p, t.a, c);
            const V3 oj = v3(t.b);
            gP += sel(g.in && fabsf(t.a.w) > 1e-8f, jpsi(g, zm, c) * norm_of_sq(dot(oj, oj)));
        });
    if (fabsf(p.w) > 1e-8f) {
        const V3 omm = v3(om[m]);
        const float wl = norm_of_sq(dot(omm, omm));
        float gs = 0.0f;
        walk_quads(
            m, t05(tcnt[m]), tquads, [&](uint32_t i) { return G2{near.get(s_P, P, i), z[i]}; },
            [&](const G2& t) {
                const Pair g = pair_of(t.a, p, c);
                const V3 zi = v3(t.b);
                gP -= sel(g.in, jpsi(g, zi, c) * wl);
                gs += g.in ? dot(psi(g), zi) : 0.0f;
            });
        if (wl > 0.0f) gom[m] = f4(v3(gom[m]) + omm * (gs / wl));
    }
    gx[m] = f4(v3(gx[m]) * scale[be.x] + gP);
}

// applyViscosity (apply_viscosity.cl:1-87, Jacobi) and updateVelocities
// (update_velocities.cl:13): dL/dVu, dL/dP_K, dL/dc; then the walls of the
// last iteration: gx <- dL/dF_{K-1}, z <- kappa = -SPIKY/rho0 dL/dF.
__global__ __launch_bounds__(kAdjThreads) void k_xsph_b(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ P,
    const float4* __restrict__ vu, const uint32_t* __restrict__ cnt, const uint4* __restrict__ quads,
    const uint32_t* __restrict__ tcnt, const uint4* __restrict__ tquads, const float4* __restrict__ gv,
    const float* __restrict__ scale, const float4* __restrict__ gom, float4* __restrict__ gx,
    float4* __restrict__ gxs, float4* __restrict__ z, const float4* __restrict__ fut, float* __restrict__ part,
    uint32_t n, uint32_t near_on) {
    __shared__ float4 s_P[NearWindow<kAdjThreads>::kLen], s_vu[NearWindow<kAdjThreads>::kLen];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint4 be = blocks[tile];
    const SceneConsts& c = scenes[be.x];
    const NearWindow<kAdjThreads> near(be, n, near_on);   // the tile's neighbourhood in LDS
    near.stage(s_P, P);
    near.stage(s_vu, vu);
    __syncthreads();
    const uint32_t m = be.y + threadIdx.x;
    float gc[1] = {0.0f};
    if (m < be.z) {
        const float4 p = P[m];
        const V3 vm = v3(vu[m]);
        const float sc = scale[be.x];
        const V3 gvm = v3(gv[m]) * sc;
        const float cp6 = c.viscosity * c.poly6;
        const V3 am = gvm * cp6;
        const V3 gam = v3(gom[m]) * -c.spiky;
        V3 S = {0.f, 0.f, 0.f}, gVu = {0.f, 0.f, 0.f}, gP = {0.f, 0.f, 0.f};
        walk_friends<1, 2>(
            m, cnt[m], quads, [&](uint32_t j) { return G2{near.get(s_P, P, j), near.get(s_vu, vu, j)}; },
            [&](const G2& t) {
                const Pair g = pair_of(p, t.a, c);
                const bool in = g.in && fabsf(t.a.w) > 1e-8f;
                const V3 u = v3(t.b) - vm;
                const float W = g.d * g.d * g.d;
                S += sel(in, u * W);
                gVu -= sel(in, am * W + cross(psi(g), gam));
                gP += sel(in, g.r * (-6.0f * g.d * g.d * dot(u, am)) + jpsi(g, cross(gam, u), c));
            });
        gc[0] = dot(S * c.poly6, gvm);
        if (fabsf(p.w) > 1e-8f) {
            walk_quads(
                m, t05(tcnt[m]), tquads,
                [&](uint32_t i) { return G4{near.get(s_P, P, i), near.get(s_vu, vu, i), gv[i], gom[i]}; },
                [&](const G4& t) {
                    const Pair g = pair_of(t.a, p, c);
                    const V3 u = vm - v3(t.b);
                    const V3 ai = (v3(t.c) * sc) * cp6;
                    const V3 gai = v3(t.d) * -c.spiky;
                    const float W = g.d * g.d * g.d;
                    gVu += sel(g.in, ai * W + cross(psi(g), gai));
                    gP -= sel(g.in, g.r * (-6.0f * g.d * g.d * dot(u, ai)) + jpsi(g, cross(gai, u), c));
                });
        }
        gVu += gvm;
        const V3 gPK = v3(gx[m]) + gP + gVu * (1.0f / c.dt);
        gxs[m] = f4(gVu * (-1.0f / c.dt));
        const V3 gF = wall_vjp(fut[m], gPK, c);
        gx[m] = f4(gF);
        z[m] = f4(gF * (-c.spiky / c.rest_density));
    }
    block_sum<1>(gc, part + (size_t)tile * kParts + 0);
}

// computeDelta (compute_delta.cl:44-186) of iteration k: dL/dlambda_k and
// the pair part of dL/dP_k; then the pointwise coefficients of
// computeScaling's adjoint: with lambda = -C/Den,
//   A = gl POLY6 / (rho0 Den),  B = gl 2 SPIKY C / (rho0^2 Den^2),  H = B G.
__global__ __launch_bounds__(kAdjThreads) void k_delta_b(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ P,
    const float4* __restrict__ aux, const uint32_t* __restrict__ cnt, const uint4* __restrict__ quads,
    const uint32_t* __restrict__ tcnt, const uint4* __restrict__ tquads, const float4* __restrict__ z,
    float4* __restrict__ gx, float4* __restrict__ z2, float* __restrict__ zb, uint32_t n, uint32_t near_on) {
    __shared__ float4 s_P[NearWindow<kAdjThreads>::kLen], s_z[NearWindow<kAdjThreads>::kLen];
    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];
    const SceneConsts& c = scenes[be.x];
    const NearWindow<kAdjThreads> near(be, n, near_on);   // the tile's neighbourhood in LDS
    near.stage(s_P, P);
    near.stage(s_z, z);
    __syncthreads();
    const uint32_t m = be.y + threadIdx.x;
    if (m >= be.z) return;
    const float4 p = P[m];
    const V3 km = v3(z[m]);
    const float ks = -24.0f * c.st_k * c.inv_poly6_q;
    float gl = 0.0f;
    V3 gP = {0.f, 0.f, 0.f};
    walk_friends<1, 2>(
        m, cnt[m], quads, [&](uint32_t j) { return near.get(s_P, P, j); },
        [&](const float4& pj) {
            const Pair g = pair_of(p, pj, c);
            const V3 ps = psi(g);
            const float ratio = g.d * g.d * g.d * c.inv_poly6_q;
            const float r3 = ratio * ratio * ratio;
            const float coef = p.w + pj.w + c.st_k * r3 * ratio;
            const float t = dot(ps, km);
            gl += g.in ? t : 0.0f;
            gP += sel(g.in, jpsi(g, km, c) * coef + g.r * (t * ks * r3 * g.d * g.d));
        });
    walk_quads(
        m, t05(tcnt[m]), tquads, [&](uint32_t i) { return G2{near.get(s_P, P, i), near.get(s_z, z, i)}; },
        [&](const G2& t) {
            const Pair g = pair_of(t.a, p, c);
            const V3 ki = v3(t.b);
            const V3 ps = psi(g);
            const float ratio = g.d * g.d * g.d * c.inv_poly6_q;
            const float r3 = ratio * ratio * ratio;
            const float coef = t.a.w + p.w + c.st_k * r3 * ratio;
            const float tt = dot(ps, ki);
            gl += g.in ? tt : 0.0f;
            gP -= sel(g.in, jpsi(g, ki, c) * coef + g.r * (tt * ks * r3 * g.d * g.d));
        });
    gx[m] = f4(v3(gx[m]) + gP);
    const float4 a4 = aux[m];
    const float den = a4.w;
    const float C = -p.w * den;
    const float rd = c.rest_density;
    const float A = gl * c.poly6 / (rd * den);
    const float B = gl * 2.0f * c.spiky * C / (rd * rd * den * den);
    z2[m] = make_float4(B * a4.x, B * a4.y, B * a4.z, A);
    zb[m] = B;
}

// computeScaling (compute_scaling.cl:1-113) of iteration k: the pair part of
// dL/dP_k, e(r) = 6A d^2 r + J(r)(B SPIKY psi(r) + H).  Then either the
// walls of iteration k-1 (gx <- dL/dF_{k-1}, z <- kappa), or for k = 0 the
// reorder and predictPositions: dL/dX[perm] = dL/dXs + dL/dP0,
// dL/dV[perm] = dt dL/dP0, dL/dg += dt^2 dL/dP0.
template <bool FIRST>
__global__ __launch_bounds__(kAdjThreads) void k_lambda_b(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ P,
    const uint32_t* __restrict__ cnt, const uint4* __restrict__ quads, const uint32_t* __restrict__ tcnt,
    const uint4* __restrict__ tquads, const float4* __restrict__ z2, const float* __restrict__ zb,
    float4* __restrict__ gx, const float4* __restrict__ fut_prev, float4* __restrict__ z,
    const uint32_t* __restrict__ perm, const float4* __restrict__ gxs, float4* __restrict__ gx_out,
    float4* __restrict__ gv_out, float* __restrict__ part, uint32_t n, uint32_t near_on) {
    __shared__ float4 s_P[NearWindow<kAdjThreads>::kLen], s_z2[NearWindow<kAdjThreads>::kLen];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint4 be = blocks[tile];
    const SceneConsts& c = scenes[be.x];
    const NearWindow<kAdjThreads> near(be, n, near_on);   // the tile's neighbourhood in LDS
    near.stage(s_P, P);
    near.stage(s_z2, z2);
    __syncthreads();
    const uint32_t m = be.y + threadIdx.x;
    float gg[3] = {0.0f, 0.0f, 0.0f};
    float mx = 0.0f;
    if (m < be.z) {
        const float4 p = P[m];
        const float4 qm = z2[m];
        const V3 Hm = v3(qm);
        const float Am = qm.w, Bm = zb[m];
        V3 gP = {0.f, 0.f, 0.f};
        walk_friends<3, 5>(
            m, cnt[m], quads, [&](uint32_t j) { return near.get(s_P, P, j); },
            [&](const float4& pj) {
                const Pair g = pair_of(p, pj, c);
                const V3 ps = psi(g);
                gP += sel(g.in, g.r * (6.0f * Am * g.d * g.d) + jpsi(g, ps * (Bm * c.spiky) + Hm, c));
            });
        walk_quads(
            m, t06(tcnt[m]), tquads,
            [&](uint32_t i) { return G2f{near.get(s_P, P, i), near.get(s_z2, z2, i), zb[i]}; },
            [&](const G2f& t) {
                const Pair g = pair_of(t.a, p, c);
                const V3 ps = psi(g);
                gP -= sel(g.in, g.r * (6.0f * t.b.w * g.d * g.d) + jpsi(g, ps * (t.f * c.spiky) + v3(t.b), c));
            });
        const V3 gPk = v3(gx[m]) + gP;
        if (!FIRST) {
            const V3 gF = wall_vjp(fut_prev[m], gPk, c);
            gx[m] = f4(gF);
            z[m] = f4(gF * (-c.spiky / c.rest_density));
        } else {
            const uint32_t pp = perm[m];
            const V3 ox = v3(gxs[m]) + gPk;
            const V3 ov = gPk * c.dt;
            if (pp < n) {
                gx_out[pp] = f4(ox);
                gv_out[pp] = f4(ov);
            }
            mx = fmaxf(fmaxf(fmaxf(fabsf(ox.x), fabsf(ox.y)), fmaxf(fabsf(ox.z), fabsf(ov.x))),
                       fmaxf(fabsf(ov.y), fabsf(ov.z)));
            const float d2 = c.dt * c.dt;
            gg[0] = gPk.x * d2;
            gg[1] = gPk.y * d2;
            gg[2] = gPk.z * d2;
        }
    }
    if (FIRST) {
        block_sum<3>(gg, part + (size_t)tile * kParts + 2);
        block_max(mx, part + (size_t)tile * kParts + 5);
    }
}

// Per-scene sums of the per-tile partials (tiles of a scene are contiguous),
// added to the running coefficient totals, and the power-of-two rescale of
// the scene's adjoint state for the next substep.  Sensitivities of these
// flows grow geometrically with the horizon; the state (dL/dx, dL/dv) is kept
// as true * 2^expo[s] with its largest component in [2^-16, 2^16], so that
// no intermediate product overflows binary32.  The totals are kept in the
// same units as the state (gparam = true * 2^expo[s]) and rescaled with it,
// so they stay finite where their true value exceeds binary32.  Scaling by a
// power of two is exact: the results equal the unscaled arithmetic wherever
// that stays finite and normal.  One wave per scene.
__global__ __launch_bounds__(64) void k_param_reduce(const uint2* __restrict__ tile_range,
                                                    const float* __restrict__ part, float* __restrict__ gparam,
                                                    float* __restrict__ scale, int* __restrict__ expo) {
    const uint32_t s = blockIdx.x;
    const uint2 tr = tile_range[s];
    const int lane = threadIdx.x;
    const int E = expo[s];
    for (int k = 0; k < 5; ++k) {
        float a = 0.0f;
        for (uint32_t t = tr.x + lane; t < tr.y; t += 64) a += part[(size_t)t * kParts + k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if (lane == 0) gparam[s * 5 + k] += a;
    }
    float mx = 0.0f;
    for (uint32_t t = tr.x + lane; t < tr.y; t += 64) {
        const float v = part[(size_t)t * kParts + 5];
        mx = (v > mx || v != v) ? v : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_xor(mx, o, 64);
        mx = (t > mx || t != t) ? t : mx;
    }
    if (lane == 0) {
        int k = 0;
        if (mx > 0.0f && mx <= 3.0e38f) {
            int e;
            (void)frexpf(mx, &e);   // mx = f 2^e, f in [0.5, 1)
            if (e > 16 || e < -16) k = -e;
        }
        scale[s] = ldexpf(1.0f, k);
        expo[s] = E + k;
        for (int q = 0; q < 5; ++q) gparam[s * 5 + q] = ldexpf(gparam[s * 5 + q], k);
    }
}

// Coefficient gradients in true units, gparam * 2^-expo; a finite total whose
// true value exceeds binary32 comes out as +-inf and sets bit 0 of the
// scene's overflow word.
__global__ __launch_bounds__(64) void k_coef_out(const float* __restrict__ gparam, const int* __restrict__ expo,
                                                 float* __restrict__ gcoef, uint32_t* __restrict__ ovf, uint32_t S) {
    const uint32_t q = blockIdx.x * 64 + threadIdx.x;
    if (q >= 5 * S) return;
    const uint32_t s = q / 5;
    const float m = gparam[q];
    const float t = ldexpf(m, -expo[s]);
    gcoef[q] = t;
    if (isfinite(m) && !isfinite(t)) atomicOr(ovf + s, 1u);
}

// The same totals as mantissa and exponent: gparam itself (the exponent,
// -expo[s], is read on the host).
__global__ __launch_bounds__(64) void k_coef_scaled(const float* __restrict__ gparam, float* __restrict__ gcoef,
                                                    uint32_t S) {
    const uint32_t q = blockIdx.x * 64 + threadIdx.x;
    if (q < 5 * S) gcoef[q] = gparam[q];
}

__global__ __launch_bounds__(64) void k_scale_init(float* __restrict__ scale, int* __restrict__ expo, uint32_t S) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s < S) {
        scale[s] = 1.0f;
        expo[s] = 0;
    }
}

// Gradients in true units (identity order: the rollout's first substep).
// The state is still in the units of the last substep: the rescale chosen
// after it (scale = 2^k, already counted in expo) has not been applied.
// SCALED writes the mantissas instead, state * scale: the true values are
// those times 2^-expo.  A finite component whose true value exceeds binary32
// comes out as +-inf and sets bit 1 of the scene's overflow word.
template <bool SCALED>
__global__ __launch_bounds__(kAdjThreads) void k_unscale(const uint4* __restrict__ blocks,
                                                        const float* __restrict__ scale, const int* __restrict__ expo,
                                                        const float4* __restrict__ gx, const float4* __restrict__ gv,
                                                        float4* __restrict__ gx0, float4* __restrict__ gv0,
                                                        uint32_t* __restrict__ ovf) {
    const uint4 be = blocks[blockIdx.x];
    const uint32_t i = be.y + threadIdx.x;
    if (i >= be.z) return;
    const int e = SCALED ? ilogbf(scale[be.x]) : ilogbf(scale[be.x]) - expo[be.x];
    const float4 a = gx[i], b = gv[i];
    const float4 ta = make_float4(ldexpf(a.x, e), ldexpf(a.y, e), ldexpf(a.z, e), 0.0f);
    const float4 tb = make_float4(ldexpf(b.x, e), ldexpf(b.y, e), ldexpf(b.z, e), 0.0f);
    if (gx0) gx0[i] = ta;
    if (gv0) gv0[i] = tb;
    if (!SCALED) {
        const bool before = isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(b.x) && isfinite(b.y) &&
                            isfinite(b.z);
        const bool after = isfinite(ta.x) && isfinite(ta.y) && isfinite(ta.z) && isfinite(tb.x) && isfinite(tb.y) &&
                           isfinite(tb.z);
        if (before && !after) atomicOr(ovf + be.x, 2u);
    }
}

// One recorded substep (device arrays).
struct TapeSlot {
    uint32_t* perm = nullptr;
    uint32_t* cnt = nullptr;
    uint4* quads = nullptr;
    uint32_t* tcnt = nullptr;
    uint4* tquads = nullptr;
    float4* P = nullptr;     // (K+1) x n: iterates, w = lambda_k (k < K) / density (k = K)
    float4* aux = nullptr;   // K x n: (sum of gradients, denominator) of computeScaling
    float4* fut = nullptr;   // K x n: pre-wall position, w = planes that moved it
    float4* vu = nullptr;    // velocity from updateVelocities
    float4* om = nullptr;    // omega from applyViscosity
};

size_t quad_words(size_t n) { return ((n + 63) / 64) * 64 * (size_t)kQuads; }

uint64_t slot_bytes(size_t n, size_t K) {
    return 4 * n * 3 + 16 * quad_words(n) * 2 + 16 * n * ((K + 1) + 2 * K + 2) + 256 * 8;
}

}  // namespace
}  // namespace pbf

struct pbf_adjoint {
    pbf_ctx* ctx = nullptr;
    uint32_t T = 0, C = 0, segs = 0, K = 0;
    std::vector<pbf::TapeSlot> slots;
    std::vector<void*> allocs;
    float4* ckpt = nullptr;            // segs x {x, v} x n
    uint32_t *ids = nullptr, *ids_tmp = nullptr;
    float4 *gx = nullptr, *gv = nullptr, *gx2 = nullptr, *gv2 = nullptr, *gom = nullptr, *gxs = nullptr;
    float4 *z = nullptr, *z2 = nullptr;
    float* zb = nullptr;
    float* part = nullptr;
    float* gparam = nullptr;               // per scene: coefficient gradients, true * 2^expo
    float* gout = nullptr;                 // per scene: coefficient gradients in true units
    float* scale = nullptr;                // per scene: rescale of the adjoint state for the next substep
    int* expo = nullptr;                   // per scene: state = true * 2^expo
    uint2* tile_range = nullptr;
    unsigned long long* notes = nullptr;   // C + 1 snapshots of the dropped-friends counter
    uint32_t* flags = nullptr;             // kFlagWords: dup / capacity flags, transposition counts
    uint32_t* ovf = nullptr;               // per scene: bit 0 coefficient, bit 1 x0/v0 gradient beyond binary32
    std::vector<uint32_t> overflow;        // ovf of the last backward (host copy)
    bool have_forward = false;
    bool have_backward = false;            // gx, gv, gparam, expo hold the last backward's results
    int recorded = -1;                     // segment whose substeps the slots hold
    double fwd_ms = 0.0, bwd_ms = 0.0, replay_ms = 0.0;
};

namespace pbf {
namespace {

template <class T>
hipError_t aalloc(pbf_adjoint* a, T** p, size_t count) {
    const hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
    if (e == hipSuccess) a->allocs.push_back((void*)*p);
    return e;
}

void free_adjoint(pbf_adjoint* a) {
    for (void* p : a->allocs) (void)hipFree(p);
    a->allocs.clear();
}

dim3 grid_of(uint32_t n) { return dim3((n + 255) / 256); }

// One substep with the forward kernels, recording into slot t.  The context's
// buffers are pointed at the slot for the arrays the slot keeps and restored
// afterwards; the rotation of x and v ends as after a plain step.
int tape_step(pbf_adjoint* a, const TapeSlot& t, unsigned long long* note) {
    pbf_ctx* ctx = a->ctx;
    const size_t n = ctx->n;
    const int K = (int)a->K;
    uint32_t* const perm0 = ctx->perm;
    uint32_t* const cnt0 = ctx->nb_cnt;
    uint4* const quads0 = ctx->nb_quads;
    float4* const xs0 = ctx->xs[0];
    float4* const xs1 = ctx->xs[1];
    float4* const om0 = ctx->omega;
    float4* const v0 = ctx->v;
    auto restore = [&]() {
        ctx->perm = perm0;
        ctx->nb_cnt = cnt0;
        ctx->nb_quads = quads0;
        ctx->xs[0] = xs0;
        ctx->xs[1] = xs1;
        ctx->omega = om0;
    };
    ctx->pause_now = 0;
    ctx->cells_fresh = true;
    hipError_t e = launch_predict_hash(ctx, ctx->cells_hold_keys);   // the sort workspace is zero between steps
    ctx->perm = t.perm;
    if (e == hipSuccess) e = launch_sort(ctx);
    ctx->xs[0] = t.P;
    if (e == hipSuccess) e = launch_gather(ctx);   // + cells, and zeroes the sort workspace again
    ctx->cells_hold_keys = true;
    ctx->order_sorted = true;
    ctx->nb_cnt = t.cnt;
    ctx->nb_quads = t.quads;
    if (e == hipSuccess) e = launch_friends(ctx);
    if (e == hipSuccess) {
        const BlockTable& bt = ctx->bt[2];
        hipLaunchKernelGGL(k_transpose, dim3(bt.count), dim3(kAdjThreads), 0, ctx->stream, ctx->sc_d, bt.d, t.P,
                           t.cnt, ctx->nb_list, reinterpret_cast<const uint32_t*>(t.quads),
                           reinterpret_cast<const uint2*>(ctx->cells), t.tcnt, t.tquads, a->flags);
        hipLaunchKernelGGL(k_note_drops, dim3(1), dim3(64), 0, ctx->stream, ctx->err, note);
        hipLaunchKernelGGL(k_transpose_all, dim3(bt.count), dim3(kAdjThreads), 0, ctx->stream, ctx->sc_d, bt.d, t.P,
                           t.cnt, t.quads, reinterpret_cast<const uint2*>(ctx->cells), ctx->n, t.tcnt, t.tquads, note,
                           a->flags);
        e = hipGetLastError();
    }
    for (int k = 0; k < K && e == hipSuccess; ++k) {
        ctx->xs[ctx->xs_cur] = t.P + (size_t)k * n;
        ctx->xs[1 - ctx->xs_cur] = t.P + (size_t)(k + 1) * n;
        e = launch_lambda(ctx, k, t.aux + (size_t)k * n);
        if (k == K - 1) ctx->v = t.vu;   // updateVelocities writes Vu into the slot
        if (e == hipSuccess) e = launch_delta(ctx, k, k == K - 1, t.fut + (size_t)k * n);
    }
    ctx->omega = t.om;
    if (e == hipSuccess) e = launch_viscosity(ctx);   // v <- v_alt (owned), v_alt <- slot
    if (e == hipSuccess) e = launch_vorticity(ctx);
    ctx->v_alt = v0;
    restore();
    if (e != hipSuccess) return fail(ctx, PBF_E_HIP, "recording a substep: %s", hipGetErrorString(e));
    return PBF_OK;
}

int ids_step(pbf_adjoint* a, const uint32_t* perm) {
    pbf_ctx* ctx = a->ctx;
    hipLaunchKernelGGL(k_ids_gather, grid_of(ctx->n), dim3(256), 0, ctx->stream, perm, a->ids, a->ids_tmp, ctx->n);
    std::swap(a->ids, a->ids_tmp);
    HIP_TRY(ctx, hipGetLastError());
    return PBF_OK;
}

// Replays segment g from its checkpoint, recording every substep.
int record_segment(pbf_adjoint* a, uint32_t g) {
    pbf_ctx* ctx = a->ctx;
    const size_t n = ctx->n;
    const uint32_t s0 = g * a->C, s1 = std::min(a->T, s0 + a->C);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->x, a->ckpt + (size_t)g * 2 * n, 16 * n, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->v, a->ckpt + ((size_t)g * 2 + 1) * n, 16 * n, hipMemcpyDeviceToDevice,
                                ctx->stream));
    hipLaunchKernelGGL(k_note_drops, dim3(1), dim3(64), 0, ctx->stream, ctx->err, a->notes);
    for (uint32_t s = s0; s < s1; ++s) {
        const int rc = tape_step(a, a->slots[s - s0], a->notes + (s - s0) + 1);
        if (rc) return rc;
    }
    a->recorded = (int)g;
    return PBF_OK;
}

// The transposed lists recorded so far are complete (checked on the host
// after the device work has finished).
int check_segment(pbf_adjoint* a, uint32_t g) {
    pbf_ctx* ctx = a->ctx;
    uint32_t w[kFlagWords];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(w, a->flags, sizeof w, hipMemcpyDeviceToHost));
    unsigned long long placed, listed;
    memcpy(&placed, w + 2, 8);
    memcpy(&listed, w + 4, 8);
    if (w[0] & kCapFlag)
        return fail(ctx, PBF_E_STATE,
                    "segment %u: a particle is listed by more than %d others (a dense cluster at the 50-per-circle "
                    "cap); its transposed list does not fit",
                    g, 4 * kQuads);
    if (placed != listed)
        return fail(ctx, PBF_E_STATE,
                    "segment %u: %llu of %llu friend entries have no transposed entry (a friend outside the "
                    "27 neighbour cells)",
                    g, listed - placed, listed);
    return PBF_OK;
}

int backward_substep(pbf_adjoint* a, const TapeSlot& t) {
    pbf_ctx* ctx = a->ctx;
    const size_t n = ctx->n;
    const int K = (int)a->K;
    const BlockTable& bt = ctx->bt[2];
    const dim3 grid(bt.count), block(kAdjThreads);
    hipStream_t st = ctx->stream;
    const float4* PK = t.P + (size_t)K * n;
    // the tiles' neighbourhoods in LDS on grids of kNearMinTilesPerCu tiles per CU or more
    const uint32_t near = bt.count >= kNearMinTilesPerCu * (uint32_t)std::max(ctx->cu_count, 1) ? 1u : 0u;
    hipLaunchKernelGGL(k_vort_b1, grid, block, 0, st, ctx->sc_d, bt.d, PK, t.om, t.cnt, t.quads, a->gv, a->scale,
                       a->z, a->gom, a->part, ctx->n, near);
    hipLaunchKernelGGL(k_vort_b2, grid, block, 0, st, ctx->sc_d, bt.d, PK, t.om, t.cnt, t.quads, t.tcnt, t.tquads,
                       a->z, a->scale, a->gom, a->gx, ctx->n, near);
    hipLaunchKernelGGL(k_xsph_b, grid, block, 0, st, ctx->sc_d, bt.d, PK, t.vu, t.cnt, t.quads, t.tcnt, t.tquads,
                       a->gv, a->scale, a->gom, a->gx, a->gxs, a->z, t.fut + (size_t)(K - 1) * n, a->part, ctx->n,
                       near);
    for (int k = K - 1; k >= 0; --k) {
        const float4* Pk = t.P + (size_t)k * n;
        hipLaunchKernelGGL(k_delta_b, grid, block, 0, st, ctx->sc_d, bt.d, Pk, t.aux + (size_t)k * n, t.cnt,
                           t.quads, t.tcnt, t.tquads, a->z, a->gx, a->z2, a->zb, ctx->n, near);
        if (k > 0) {
            hipLaunchKernelGGL(k_lambda_b<false>, grid, block, 0, st, ctx->sc_d, bt.d, Pk, t.cnt, t.quads, t.tcnt,
                               t.tquads, a->z2, a->zb, a->gx, t.fut + (size_t)(k - 1) * n, a->z, t.perm, a->gxs,
                               a->gx2, a->gv2, a->part, ctx->n, near);
        } else {
            hipLaunchKernelGGL(k_lambda_b<true>, grid, block, 0, st, ctx->sc_d, bt.d, Pk, t.cnt, t.quads, t.tcnt,
                               t.tquads, a->z2, a->zb, a->gx, nullptr, a->z, t.perm, a->gxs, a->gx2, a->gv2,
                               a->part, ctx->n, near);
        }
    }
    hipLaunchKernelGGL(k_param_reduce, dim3(ctx->S), dim3(64), 0, st, a->tile_range, a->part, a->gparam, a->scale,
                       a->expo);
    HIP_TRY(ctx, hipGetLastError());
    std::swap(a->gx, a->gx2);
    std::swap(a->gv, a->gv2);
    return PBF_OK;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace pbf

extern "C" {

int pbf_set_scene_coefficients(pbf_ctx* ctx, uint32_t scene, float viscosity, float vorticity,
                               const float gravity[3]) {
    if (!ctx) return PBF_E_INVALID;
    if (scene >= ctx->S) return pbf::fail(ctx, PBF_E_INVALID, "scene %u out of range (%u scenes)", scene, ctx->S);
    if (!gravity) return pbf::fail(ctx, PBF_E_INVALID, "gravity is NULL");
    pbf::SceneConsts& c = ctx->sc[scene];
    c.viscosity = viscosity;
    c.vorticity = vorticity;
    c.g0 = gravity[0];
    c.g1 = gravity[1];
    c.g2 = gravity[2];
    ctx->P[scene].viscosityFactor = viscosity;
    ctx->P[scene].vorticityFactor = vorticity;
    ctx->opt[scene].use_gravity_vector = 1;
    for (int k = 0; k < 3; ++k) ctx->opt[scene].gravity[k] = gravity[k];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(ctx->sc_d + scene, &c, sizeof c, hipMemcpyHostToDevice));
    return PBF_OK;
}

int pbf_get_scene_coefficients(pbf_ctx* ctx, uint32_t scene, float* out5) {
    if (!ctx || !out5) return PBF_E_INVALID;
    if (scene >= ctx->S) return pbf::fail(ctx, PBF_E_INVALID, "scene %u out of range (%u scenes)", scene, ctx->S);
    const pbf::SceneConsts& c = ctx->sc[scene];
    out5[0] = c.viscosity;
    out5[1] = c.vorticity;
    out5[2] = c.g0;
    out5[3] = c.g1;
    out5[4] = c.g2;
    return PBF_OK;
}

int pbf_adjoint_tape_bytes(pbf_ctx* ctx, uint64_t* per_substep, uint64_t* per_checkpoint) {
    if (!ctx) return PBF_E_INVALID;
    if (per_substep) *per_substep = pbf::slot_bytes(ctx->n, ctx->iters);
    if (per_checkpoint) *per_checkpoint = 32ull * ctx->n;
    return PBF_OK;
}

int pbf_adjoint_plan(pbf_ctx* ctx, uint32_t substeps, uint64_t budget_bytes, uint32_t* interval) {
    if (!ctx || !interval || substeps == 0) return PBF_E_INVALID;
    const uint64_t slot = pbf::slot_bytes(ctx->n, ctx->iters), ck = 32ull * ctx->n;
    const uint64_t fixed = 4ull * 2 * ctx->n + 16ull * 8 * ctx->n + 4ull * ctx->n;   // ids, gradients, coefficients
    // Largest interval that fits: the fewest intervals.  The forward records
    // the last interval itself, so the backward replays all substeps but
    // those of the last one; among the intervals giving the fewest intervals,
    // the smallest makes the last one longest.
    uint32_t cmax = 0;
    for (uint32_t C = substeps; C >= 1; --C) {
        const uint64_t segs = (substeps + C - 1) / C;
        if (fixed + C * slot + segs * ck <= budget_bytes) {
            cmax = C;
            break;
        }
    }
    if (cmax == 0)
        return pbf::fail(ctx, PBF_E_NOMEM, "one recorded substep (%llu B) does not fit in %llu B",
                         (unsigned long long)slot, (unsigned long long)budget_bytes);
    const uint32_t segs = (substeps + cmax - 1) / cmax;
    *interval = (substeps + segs - 1) / segs;
    return PBF_OK;
}

int pbf_adjoint_create(pbf_ctx* ctx, uint32_t substeps, uint32_t interval, pbf_adjoint** out) {
    if (!ctx || !out) return PBF_E_INVALID;
    *out = nullptr;
    if (substeps == 0 || interval == 0) return pbf::fail(ctx, PBF_E_INVALID, "substeps and interval must be >= 1");
    pbf_adjoint* a = new (std::nothrow) pbf_adjoint();
    if (!a) return pbf::fail(ctx, PBF_E_NOMEM, "host allocation failed");
    a->ctx = ctx;
    a->T = substeps;
    a->C = std::min(interval, substeps);
    a->segs = (a->T + a->C - 1) / a->C;
    a->K = ctx->iters;
    const size_t n = ctx->n, K = a->K, qw = pbf::quad_words(n);
    auto bail = [&](hipError_t e) {
        const int code = e == hipErrorOutOfMemory ? PBF_E_NOMEM : PBF_E_HIP;
        pbf::fail(ctx, code, "adjoint allocation: %s", hipGetErrorString(e));
        pbf::free_adjoint(a);
        delete a;
        return code;
    };
#define ADJ_TRY(call)                          \
    do {                                       \
        const hipError_t e_ = (call);          \
        if (e_ != hipSuccess) return bail(e_); \
    } while (0)
    a->slots.resize(a->C);
    for (pbf::TapeSlot& t : a->slots) {
        ADJ_TRY(pbf::aalloc(a, &t.perm, n));
        ADJ_TRY(pbf::aalloc(a, &t.cnt, n));
        ADJ_TRY(pbf::aalloc(a, &t.tcnt, n));
        ADJ_TRY(pbf::aalloc(a, &t.quads, qw));
        ADJ_TRY(pbf::aalloc(a, &t.tquads, qw));
        ADJ_TRY(pbf::aalloc(a, &t.P, (K + 1) * n));
        ADJ_TRY(pbf::aalloc(a, &t.aux, K * n));
        ADJ_TRY(pbf::aalloc(a, &t.fut, K * n));
        ADJ_TRY(pbf::aalloc(a, &t.vu, n));
        ADJ_TRY(pbf::aalloc(a, &t.om, n));
        // updateVelocities writes only xyz of the recorded velocity: w stays 0
        ADJ_TRY(hipMemset(t.vu, 0, 16 * std::max<size_t>(n, 1)));
    }
    ADJ_TRY(pbf::aalloc(a, &a->ckpt, (size_t)a->segs * 2 * n));
    ADJ_TRY(pbf::aalloc(a, &a->ids, n));
    ADJ_TRY(pbf::aalloc(a, &a->ids_tmp, n));
    ADJ_TRY(pbf::aalloc(a, &a->gx, n));
    ADJ_TRY(pbf::aalloc(a, &a->gv, n));
    ADJ_TRY(pbf::aalloc(a, &a->gx2, n));
    ADJ_TRY(pbf::aalloc(a, &a->gv2, n));
    ADJ_TRY(pbf::aalloc(a, &a->gom, n));
    ADJ_TRY(pbf::aalloc(a, &a->gxs, n));
    ADJ_TRY(pbf::aalloc(a, &a->z, n));
    ADJ_TRY(pbf::aalloc(a, &a->z2, n));
    ADJ_TRY(pbf::aalloc(a, &a->zb, n));
    const pbf::BlockTable& bt = ctx->bt[2];
    ADJ_TRY(pbf::aalloc(a, &a->part, (size_t)bt.count * pbf::kParts));
    ADJ_TRY(pbf::aalloc(a, &a->gparam, (size_t)ctx->S * 5));
    ADJ_TRY(pbf::aalloc(a, &a->gout, (size_t)ctx->S * 5));
    ADJ_TRY(pbf::aalloc(a, &a->scale, ctx->S));
    ADJ_TRY(pbf::aalloc(a, &a->expo, ctx->S));
    ADJ_TRY(pbf::aalloc(a, &a->tile_range, ctx->S));
    ADJ_TRY(pbf::aalloc(a, &a->notes, (size_t)a->C + 1));
    ADJ_TRY(pbf::aalloc(a, &a->flags, pbf::kFlagWords));
    ADJ_TRY(pbf::aalloc(a, &a->ovf, ctx->S));
    // tiles of the 256-particle block table per scene
    std::vector<uint2> tr(ctx->S);
    uint32_t tile = 0;
    for (uint32_t s = 0; s < ctx->S; ++s) {
        const uint32_t tiles = (ctx->sc[s].count + 255) / 256;
        tr[s] = make_uint2(tile, tile + tiles);
        tile += tiles;
    }
    if (tile != bt.count) {
        pbf::fail(ctx, PBF_E_INVALID, "internal: block table has %u tiles, expected %u", bt.count, tile);
        pbf::free_adjoint(a);
        delete a;
        return PBF_E_INVALID;
    }
    ADJ_TRY(hipMemcpy(a->tile_range, tr.data(), sizeof(uint2) * ctx->S, hipMemcpyHostToDevice));
    ADJ_TRY(hipMemset(a->flags, 0, 4 * pbf::kFlagWords));
#undef ADJ_TRY
    *out = a;
    return PBF_OK;
}

int pbf_adjoint_forward(pbf_adjoint* a, const float* x0, const float* v0, float* xT, float* vT) {
    if (!a) return PBF_E_INVALID;
    pbf_ctx* ctx = a->ctx;
    if (!x0 || !v0) return pbf::fail(ctx, PBF_E_INVALID, "x0 and v0 are required");
    const double t0 = pbf::now_ms();
    const size_t n = ctx->n;
    a->have_forward = false;
    a->have_backward = false;
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->x, x0, 16 * n, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->v, v0, 16 * n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    ctx->have_state = true;
    ctx->order_sorted = false;
    if (ctx->x_base == nullptr) {
        ctx->x_base = ctx->x;
        ctx->v_base = ctx->v;
    }
    if (n == 0) return PBF_OK;
    hipLaunchKernelGGL(pbf::k_iota, pbf::grid_of(ctx->n), dim3(256), 0, ctx->stream, a->ids, ctx->n);
    HIP_TRY(ctx, hipMemsetAsync(a->flags, 0, 4 * pbf::kFlagWords, ctx->stream));
    for (uint32_t g = 0; g < a->segs; ++g) {
        HIP_TRY(ctx, hipMemcpyAsync(a->ckpt + (size_t)g * 2 * n, ctx->x, 16 * n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(a->ckpt + ((size_t)g * 2 + 1) * n, ctx->v, 16 * n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        const uint32_t s0 = g * a->C, s1 = std::min(a->T, s0 + a->C);
        if (g + 1 == a->segs) {
            hipLaunchKernelGGL(pbf::k_note_drops, dim3(1), dim3(64), 0, ctx->stream, ctx->err, a->notes);
            for (uint32_t s = s0; s < s1; ++s) {
                int rc = pbf::tape_step(a, a->slots[s - s0], a->notes + (s - s0) + 1);
                if (!rc) rc = pbf::ids_step(a, a->slots[s - s0].perm);
                if (rc) return rc;
            }
            a->recorded = (int)g;
        } else {
            for (uint32_t s = s0; s < s1; ++s) {
                int rc = pbf_step(ctx, 0);
                if (!rc) rc = pbf::ids_step(a, ctx->perm);
                if (rc) return rc;
            }
        }
    }
    if (xT || vT)
        hipLaunchKernelGGL(pbf::k_scatter_state, pbf::grid_of(ctx->n), dim3(256), 0, ctx->stream, a->ids, ctx->x,
                           ctx->v, reinterpret_cast<float4*>(xT), reinterpret_cast<float4*>(vT), ctx->n);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    a->have_forward = true;
    a->fwd_ms = pbf::now_ms() - t0;
    return PBF_OK;
}

int pbf_adjoint_backward(pbf_adjoint* a, const float* gxT, const float* gvT, float* gx0, float* gv0,
                         float* gcoef) {
    if (!a) return PBF_E_INVALID;
    pbf_ctx* ctx = a->ctx;
    if (!a->have_forward) return pbf::fail(ctx, PBF_E_STATE, "pbf_adjoint_backward before pbf_adjoint_forward");
    const double t0 = pbf::now_ms();
    double replay = 0.0;
    const size_t n = ctx->n;
    if (n == 0) return PBF_OK;
    hipLaunchKernelGGL(pbf::k_gather_grads, pbf::grid_of(ctx->n), dim3(256), 0, ctx->stream, a->ids,
                       reinterpret_cast<const float4*>(gxT), reinterpret_cast<const float4*>(gvT), a->gx, a->gv,
                       ctx->n);
    hipLaunchKernelGGL(pbf::k_fill, dim3((ctx->S * 5 + 255) / 256), dim3(256), 0, ctx->stream, a->gparam,
                       (size_t)ctx->S * 5);
    hipLaunchKernelGGL(pbf::k_scale_init, dim3((ctx->S + 63) / 64), dim3(64), 0, ctx->stream, a->scale, a->expo,
                       ctx->S);
    HIP_TRY(ctx, hipMemsetAsync(a->ovf, 0, 4 * (size_t)ctx->S, ctx->stream));
    HIP_TRY(ctx, hipGetLastError());
    a->have_backward = false;
    for (int g = (int)a->segs - 1; g >= 0; --g) {
        if (a->recorded != g) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            const double r0 = pbf::now_ms();
            const int rc = pbf::record_segment(a, (uint32_t)g);
            if (rc) return rc;
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            replay += pbf::now_ms() - r0;
        }
        int rc = pbf::check_segment(a, (uint32_t)g);
        if (rc) return rc;
        const uint32_t s0 = (uint32_t)g * a->C, s1 = std::min(a->T, s0 + a->C);
        for (uint32_t s = s1; s-- > s0;) {
            rc = pbf::backward_substep(a, a->slots[s - s0]);
            if (rc) return rc;
        }
    }
    {
        // true units, flagging what exceeds binary32 (the x0/v0 check runs
        // even when the caller takes neither)
        const pbf::BlockTable& bt = ctx->bt[2];
        hipLaunchKernelGGL(pbf::k_unscale<false>, dim3(bt.count), dim3(pbf::kAdjThreads), 0, ctx->stream, bt.d,
                           a->scale, a->expo, a->gx, a->gv, reinterpret_cast<float4*>(gx0),
                           reinterpret_cast<float4*>(gv0), a->ovf);
        hipLaunchKernelGGL(pbf::k_coef_out, dim3((5 * ctx->S + 63) / 64), dim3(64), 0, ctx->stream, a->gparam,
                           a->expo, a->gout, a->ovf, ctx->S);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (gcoef)
        HIP_TRY(ctx, hipMemcpyAsync(gcoef, a->gout, 4 * 5 * (size_t)ctx->S, hipMemcpyDeviceToDevice, ctx->stream));
    a->overflow.assign(ctx->S, 0u);
    HIP_TRY(ctx, hipMemcpyAsync(a->overflow.data(), a->ovf, 4 * (size_t)ctx->S, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    a->have_backward = true;
    a->bwd_ms = pbf::now_ms() - t0;
    a->replay_ms = replay;
    return PBF_OK;
}

int pbf_adjoint_overflow(pbf_adjoint* a, uint32_t* scenes, uint32_t* per_scene) {
    if (!a || !scenes) return PBF_E_INVALID;
    if (!a->have_backward) return pbf::fail(a->ctx, PBF_E_STATE, "pbf_adjoint_overflow before pbf_adjoint_backward");
    uint32_t k = 0;
    for (size_t s = 0; s < a->overflow.size(); ++s) {
        if (a->overflow[s]) ++k;
        if (per_scene) per_scene[s] = a->overflow[s];
    }
    *scenes = k;
    return PBF_OK;
}

int pbf_adjoint_scaled_gradients(pbf_adjoint* a, float* gx0, float* gv0, float* gcoef, int32_t* expo) {
    if (!a) return PBF_E_INVALID;
    pbf_ctx* ctx = a->ctx;
    if (!a->have_backward)
        return pbf::fail(ctx, PBF_E_STATE, "pbf_adjoint_scaled_gradients before pbf_adjoint_backward");
    if (ctx->n == 0) return PBF_OK;
    if (gx0 || gv0) {
        const pbf::BlockTable& bt = ctx->bt[2];
        hipLaunchKernelGGL(pbf::k_unscale<true>, dim3(bt.count), dim3(pbf::kAdjThreads), 0, ctx->stream, bt.d,
                           a->scale, a->expo, a->gx, a->gv, reinterpret_cast<float4*>(gx0),
                           reinterpret_cast<float4*>(gv0), a->ovf);
    }
    if (gcoef)
        hipLaunchKernelGGL(pbf::k_coef_scaled, dim3((5 * ctx->S + 63) / 64), dim3(64), 0, ctx->stream, a->gparam,
                           gcoef, ctx->S);
    HIP_TRY(ctx, hipGetLastError());
    if (expo) {
        HIP_TRY(ctx, hipMemcpyAsync(expo, a->expo, 4 * (size_t)ctx->S, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (expo)
        for (uint32_t s = 0; s < ctx->S; ++s) expo[s] = -expo[s];
    return PBF_OK;
}

int pbf_adjoint_times(pbf_adjoint* a, double* forward_ms, double* backward_ms, double* replay_ms) {
    if (!a) return PBF_E_INVALID;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (backward_ms) *backward_ms = a->bwd_ms;
    if (replay_ms) *replay_ms = a->replay_ms;
    return PBF_OK;
}

void pbf_adjoint_destroy(pbf_adjoint* a) {
    if (!a) return;
    if (a->ctx && a->ctx->stream) (void)hipStreamSynchronize(a->ctx->stream);
    pbf::free_adjoint(a);
    delete a;
}

}  // extern "C"
