// pbf_forward.hip -- the per-substep kernels other than the sort.
//
// Kernel chain of one substep (reference: src/Simulation.cpp:554-621):
//   k_predict_hash   predictPositions + computeKeys + digit histograms for the
//                    sort + reset of the previous step's cells
//   (onesweep passes, pbf_sort.hip)
//   k_gather_cells   sortParticles (recomputes x* from the gathered x, v) +
//                    updateCells; zeroes the sort workspace for the next step
//   k_friends_direct buildFriendsList, candidates read from global memory
//   / k_friends      (or from the workgroup's neighbourhood staged in LDS)
//   K x { k_lambda   computeScaling + packData(lambda)
//         k_delta    computeDelta (+ walls and collision quads) +
//                    updatePredicted; on the last iteration also
//                    packData(density) + updateVelocities }
//   k_viscosity      applyViscosity (Jacobi: reads a snapshot of v)
//   k_vorticity      applyVorticity
//
// Every per-particle kernel runs over a block table (one {scene, first, end}
// per workgroup), so one launch covers all scenes of a batch and the scene
// constants of a workgroup are uniform (scalar loads).  Particles live in
// cell-sorted order as float4 arrays.  The friends build writes each
// particle's circle lists (counts packed 5 x 6 bits) and the list a sweep
// processes, grouped per 64 consecutive particles in 16-byte quads, so a
// wave's load of one quad is a contiguous 1 KB (pbf_device.h).  The sweeps
// walk a particle's friends a quad at a time, one quad ahead: the next
// quad's four gathers are in flight while the current quad's four updates
// run in the reference's order.
#include <algorithm>

#include "pbf_internal.h"

namespace pbf {

namespace {


__device__ __forceinline__ float4 predicted_from(const float4 x, const float4 v, const SceneConsts& c, uint32_t pause) {
    // predict_positions.cl:12-17
    float4 vg = v;
    if (pause == 0u) {
        vg.x = v.x + c.dt * c.g0;
        vg.y = v.y + c.dt * c.g1;
        vg.z = v.z + c.dt * c.g2;
    }
    float4 o;
    o.x = x.x + c.dt * vg.x;
    o.y = x.y + c.dt * vg.y;
    o.z = x.z + c.dt * vg.z;
    o.w = x.w + c.dt * vg.w;
    return o;
}

__device__ __forceinline__ uint32_t key_of(const float4 p, const SceneConsts& c) {
    // radixsort.cl:11-13
    return cell_hash(cell_coord(p.x / c.h), cell_coord(p.y / c.h), cell_coord(p.z / c.h), c);
}

__global__ __launch_bounds__(kPredictThreads) void k_predict_hash(
    const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks, const float4* __restrict__ x,
    const float4* __restrict__ v, uint32_t* __restrict__ keys, const uint32_t* __restrict__ old_keys,
    uint2* __restrict__ cells, uint32_t reset_old, uint32_t total_cells, uint32_t* __restrict__ hist,
    uint32_t passes, uint32_t pause, uint32_t* err) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    __shared__ uint32_t s_bad[2];   // first non-finite particle of the workgroup, their count
    const uint4 be = blocks[blockIdx.x];
    const SceneConsts& c = scenes[be.x];
    for (int q = threadIdx.x; q < (int)passes * kRadix; q += kPredictThreads) s_hist[q] = 0;
    if (threadIdx.x == 0) {
        s_bad[0] = 0xFFFFFFFFu;
        s_bad[1] = 0u;
    }
    __syncthreads();
    // every load of the thread's particles first, so that they are in flight
    // together (computed item by item, each waited for its own loads)
    float4 xa[kPredictItems], va[kPredictItems];
    uint32_t ok[kPredictItems], okp[kPredictItems];
#pragma unroll
    for (int k = 0; k < kPredictItems; ++k) {
        // clamped into the tile (be.y < be.z): no branch between the loads
        const uint32_t i = min(be.y + (uint32_t)(k * kPredictThreads) + threadIdx.x, be.z - 1u);
        xa[k] = x[i];
        va[k] = v[i];
    }
    if (reset_old) {
#pragma unroll
        for (int k = 0; k < kPredictItems; ++k) {
            const uint32_t i = min(be.y + (uint32_t)(k * kPredictThreads) + threadIdx.x, be.z - 1u);
            ok[k] = old_keys[i];
            okp[k] = old_keys[i > 0 ? i - 1 : 0];
        }
    }
#pragma unroll
    for (int k = 0; k < kPredictItems; ++k) {
        const uint32_t i = be.y + (uint32_t)(k * kPredictThreads) + threadIdx.x;
        if (i < be.z) {
            const float4 p = predicted_from(xa[k], va[k], c, pause);
            const uint32_t key = key_of(p, c);
            keys[i] = key;
            for (uint32_t q = 0; q < passes; ++q)
                atomicAdd(&s_hist[q * kRadix + ((key >> (kRadixBits * q)) & (kRadix - 1))], 1u);
            // reset_grid.cl: empty the cells of the previous step (the first
            // particle of each old cell; i = 0 always starts one)
            if (reset_old && (i == 0 || okp[k] != ok[k]) && ok[k] < total_cells)
                cells[ok[k]] = make_uint2(kEmpty, kEmpty);
            if (!(fabsf(p.x) <= 3.4e38f && fabsf(p.y) <= 3.4e38f && fabsf(p.z) <= 3.4e38f)) {
                atomicMin(&s_bad[0], i);
                atomicAdd(&s_bad[1], 1u);
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (int)passes * kRadix; q += kPredictThreads) {
        const uint32_t h = s_hist[q];
        if (h) atomicAdd(&hist[q], h);
    }
    if (threadIdx.x == 0 && s_bad[1]) {
        atomicOr(err, (uint32_t)PBF_FLAG_NAN);
        log_event(err, PBF_LOG_NAN, s_bad[0], s_bad[1], be.x);
    }
}

// Paused step with a valid sort: predictPositions only (no re-sort).
__global__ __launch_bounds__(256) void k_predict_only(const SceneConsts* __restrict__ scenes,
                                                     const uint4* __restrict__ blocks,
                                                     const float4* __restrict__ x, const float4* __restrict__ v,
                                                     float4* __restrict__ xs, uint32_t pause) {
    const uint4 be = blocks[blockIdx.x];
    const SceneConsts& c = scenes[be.x];
    const uint32_t i = be.y + threadIdx.x;
    if (i >= be.z) return;
    xs[i] = predicted_from(x[i], v[i], c, pause);
}

// radixsort.cl:22-37 (sortParticles) and update_cells.cl:1-26 (updateCells)
// in one pass over the sorted order.  x* is recomputed from the gathered x
// and v with the same operations as k_predict_hash, so it is bitwise the
// value that was hashed; v itself is not permuted (it is rewritten before it
// is read).  The cells come from the sorted global keys (the last particle
// always closes its cell).  The launch also zeroes the sort workspace
// (histograms, tile counters, look-back status), whose last reader was this
// step's final sort pass, for the next step's histograms.
__global__ __launch_bounds__(256) void k_gather_cells(const SceneConsts* __restrict__ scenes,
                                                     const uint4* __restrict__ blocks,
                                                     const uint32_t* __restrict__ perm, const float4* __restrict__ x,
                                                     const float4* __restrict__ v, float4* __restrict__ x_out,
                                                     float4* __restrict__ xs_out, const uint32_t* __restrict__ ks,
                                                     uint32_t* __restrict__ cells, uint32_t total_cells,
                                                     uint32_t* __restrict__ ws, size_t ws_words, uint32_t pause,
                                                     uint32_t n, uint32_t* err) {
    const size_t stride = (size_t)gridDim.x * 256u;
    for (size_t w = (size_t)blockIdx.x * 256u + threadIdx.x; w < ws_words; w += stride) ws[w] = 0u;
    const uint4 be = blocks[blockIdx.x];
    const SceneConsts& c = scenes[be.x];
    const uint32_t i = be.y + threadIdx.x;
    if (i >= be.z) return;
    uint32_t p = perm[i];
    if (p >= n) {  // only after a failed sort (flagged there)
        atomicOr(err, (uint32_t)PBF_FLAG_INTERNAL);
        log_event(err, PBF_LOG_INTERNAL, 1u, i, p);
        p = i;
    }
    const float4 xp = x[p];
    x_out[i] = xp;
    xs_out[i] = predicted_from(xp, v[p], c, pause);

    const uint32_t k = ks[i];
    if (k >= total_cells || (i > 0 && ks[i - 1] >= total_cells)) {  // only after a failed sort
        atomicOr(err, (uint32_t)PBF_FLAG_INTERNAL);
        log_event(err, PBF_LOG_INTERNAL, 2u, i, k);
        return;
    }
    if (i == 0) {
        cells[2 * (size_t)k] = 0;
    } else {
        const uint32_t pk = ks[i - 1];
        if (pk != k) {
            cells[2 * (size_t)k] = i;
            cells[2 * (size_t)pk + 1] = i - 1;
        }
    }
    if (i == n - 1) cells[2 * (size_t)k + 1] = i;
}

// compute_scaling.cl:1-113 + pack_data.cl (lambda into x*.w, in place: no
// thread of this kernel reads another particle's w).
struct LambdaAcc {
    float dens = 0.0f, gsum = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    // Branch-free: the contributions of an out-of-range pair, and of a list
    // entry past the walk's end (valid false), are selected to +0 (no sum
    // here is ever -0, so the result is bitwise that of skipping it, also
    // when the friend's position is not finite) and no branch joins the
    // loop, which keeps the next gathers in flight.
    __device__ __forceinline__ void add(const float4& p, const float4& pj, const SceneConsts& c, bool valid) {
        const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
        const float r2 = dist2(rx, ry, rz);
        const uint32_t in = lane_mask(valid && r2 < c.h2);
        const PairLen L = pair_len_sweep(r2);   // r/|r| = r * (1/|r|); 0 for a coincident pair
        const float hr = c.h - L.rl;
        const float s = c.spiky * hr * hr;
        const float t = s * L.inv;
        const float ax = masked(rx * t, in), ay = masked(ry * t, in), az = masked(rz * t, in);
        const float d = masked(c.h2 - r2, in);
        dens = fmaf(d * d, d, dens);
        gsum = fmaf(az, az, fmaf(ay, ay, fmaf(ax, ax, gsum)));
        gx += ax;
        gy += ay;
        gz += az;
    }
};

// computeScaling for particle i from its sweep list (packed circle counts
// and quads): density, lambda, and lambda into x*.w in place (no thread of a
// launch reads another particle's w).  load(j) gathers x*_j (only its x, y,
// z are used).  TAPE also stores (sum of gradients,
// denominator) for the adjoint.
template <bool TAPE, class Load>
__device__ __forceinline__ void scaling_of(uint32_t i, uint32_t packed, float4* __restrict__ xs,
                                           const uint4* __restrict__ quads, float* __restrict__ rho,
                                           float* __restrict__ lambda, float4* __restrict__ aux, const SceneConsts& c,
                                           Load&& load) {
    const float4 p = xs[i];
    LambdaAcc a;
    walk_friends<3, 5, true>(
        i, packed, quads, load,
        [&](const float4& q, bool valid) { a.add(p, q, c, valid); });
    const float dens = a.dens * c.poly6;
    rho[i] = dens;
    const float gsum = a.gsum + (a.gx * a.gx + a.gy * a.gy + a.gz * a.gz);
    const float C = (dens / c.rest_density) - 1.0f;
    const float den = gsum / (c.rest_density * c.rest_density) + c.epsilon;
    const float lam = -1.0f * C / den;
    lambda[i] = lam;
    reinterpret_cast<float*>(xs + i)[3] = lam;
    if (TAPE) aux[i] = make_float4(a.gx, a.gy, a.gz, den);
}

// ---------------------------------------------------------------------------
// buildFriendsList (build_friends_list.cl:23-103).  27 cells in x, y, z order,
// ascending j; circle = min(int(a^3 * 5), 4), a = max(0, (r - 0.3h) / 0.7h);
// at most 50 friends per circle, later ones dropped (and counted).  A thread
// appends each friend to its particle's circle segment as it finds it.
//
// Two variants give the same lists.  k_friends_direct reads candidates from
// global memory, four at a time; the neighbouring tiles are in the same
// XCD's L2.  k_friends stages its workgroup's neighbourhood in LDS: the BS
// consecutive sorted particles of a tile span a small cell box (Morton order
// keeps a run of particles compact); the box grown by one cell is staged
// (every box cell's hash bucket, start and count, and the bucket's particles
// in bucket order), and each thread scans its 27 cells from LDS.  Hash-folded
// cells stage their bucket once per alias, which reproduces the reference's
// repeated visits.  A tile whose box or particles exceed the LDS budget scans
// global memory.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Circle of a candidate in range, or -1 when r^2 >= h^2 or r^2 is NaN (a
// friend needs r^2 < h^2; the reference's relaxed-math build leaves NaN
// comparisons undefined, and skipping NaN keeps the cell culling exact).  The
// reference's
// circle, min(int(a^3 * 5), 4) with a = max(0, (sqrt(r^2) - 0.3h) / 0.7h)
// (build_friends_list.cl:75-76), is a non-decreasing step function of the
// binary32 r^2 (every operation in it is monotone), so it equals the number
// of thresholds circ_t[k] <= r^2; the thresholds are found on the host by
// evaluating the same expression.
__device__ __forceinline__ int friend_circle(const float4& p, const float4& q, const SceneConsts& c) {
    const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
    const int circ =
        (int)(r2 >= c.circ_t[0]) + (int)(r2 >= c.circ_t[1]) + (int)(r2 >= c.circ_t[2]) + (int)(r2 >= c.circ_t[3]);
    return (int)((uint32_t)circ | ~lane_mask(r2 < c.h2));   // -1 out of range, computed without a branch
}

// Appends friend j of circle `circ` to particle i's list under the 50 cap.
// Each lane stages the quad it is filling for every circle in its own LDS
// column (entry (c, s) at stage[(4c + s) * BS]) and stores the quad when it
// fills; finish() stores the partial ones (entries past a count are never
// read).
template <int BS>
struct FriendWriter {
    uint32_t packed = 0, dropped = 0;
    uint32_t* stage = nullptr;
    uint4* qbase0 = nullptr;   // circle 0: the particle's sweep quads, quad q at q * 64
    uint4* qbase1 = nullptr;   // circles 1..4: its circle quads, quad q of circle c at ((c - 1) * 13 + q) * 64
    __device__ __forceinline__ void bind(uint32_t i, uint4* __restrict__ list, uint4* __restrict__ sweep) {
        qbase0 = sweep + sweep_quad(i, 0);
        qbase1 = list + friend_quad(i, 1, 0);
    }
    __device__ __forceinline__ static uint4* quad_of(uint32_t i, int circ, int q, uint4* __restrict__ list,
                                                      uint4* __restrict__ sweep) {
        return circ == 0 ? sweep + sweep_quad(i, q) : list + friend_quad(i, circ, q);
    }
    __device__ __forceinline__ void take(uint32_t i, uint32_t j, int circ, uint4* __restrict__ list,
                                         uint4* __restrict__ sweep) {
        if (circ < 0 || j == i) return;
        // circ is 0..4 here: 24-bit multiplies run at the full VALU rate,
        // 32-bit ones at a quarter of it
        const uint32_t sh = __umul24((uint32_t)circ, 6u);
        const uint32_t f = (packed >> sh) & 63u;
        if (f >= (uint32_t)kPerCircle) {
            ++dropped;
            return;
        }
        uint32_t* col = stage + 4 * circ * BS;
        col[(f & 3u) * BS] = j;
        if ((f & 3u) == 3u) {
            // quad f / 4 of the circle's segment, addressed without a branch on
            // the circle: circle 0 in the sweep quads, circle c > 0 at quad
            // (c - 1) * 13 of the circle quads, read from 6-bit fields at the
            // same shift as the counts.  Quads of one particle lie 64 apart,
            // 1 KiB, so the offset (at most 64 quads) fits 32 bits.
            constexpr uint32_t kSegBase = (1u * kSegQuads << 12) | (2u * kSegQuads << 18) | (3u * kSegQuads << 24);
            const uint32_t qo = ((kSegBase >> sh) & 63u) + (f >> 2);
            char* base = reinterpret_cast<char*>(circ == 0 ? qbase0 : qbase1);
            *reinterpret_cast<uint4*>(base + (qo << 10)) = make_uint4(col[0], col[BS], col[2 * BS], col[3 * BS]);
        }
        packed += 1u << sh;
    }
    __device__ __forceinline__ void finish(uint32_t i, uint32_t* __restrict__ nb_cnt, uint4* __restrict__ list,
                                           uint4* __restrict__ sweep, uint32_t* err) {
#pragma unroll
        for (int q = 0; q < kCircles; ++q) {
            const int f = circle_count(packed, q);
            if (f & 3) {
                const uint32_t* col = stage + 4 * q * BS;
                *quad_of(i, q, f >> 2, list, sweep) = make_uint4(col[0], col[BS], col[2 * BS], col[3 * BS]);
            }
        }
        nb_cnt[i] = packed;
        if (dropped) {
            atomicOr(err, (uint32_t)PBF_FLAG_FRIENDS_OVERFLOW);
            atomicAdd((unsigned long long*)(err + 2), (unsigned long long)dropped);
            log_event(err, PBF_LOG_FRIENDS_OVERFLOW, i, dropped, packed);
        }
    }
};

// Appends particle i's processed circles after circle 0 (already in place)
// in the sweep quads.  Run by the thread that has just written the circle
// quads, so they are read back from cache; with circle 0 alone processed
// there is nothing to do.
__device__ __forceinline__ void pack_friends(uint32_t i, uint32_t packed, const uint4* __restrict__ list,
                                             uint4* __restrict__ quads) {
    int m;
    const int L = circles_processed<3, 5>(packed, &m);
    if (L <= 1) return;
    const int n0 = circle_count(packed, 0);
    uint32_t acc[4] = {0u, 0u, 0u, 0u};
    int na = n0 & 3, qo = n0 >> 2;
    if (na) {
        const uint4 t = quads[sweep_quad(i, qo)];
        acc[0] = t.x; acc[1] = t.y; acc[2] = t.z; acc[3] = t.w;
    }
    for (int q = 1; q < L; ++q) {
        const int cq = circle_count(packed, q);
        for (int f = 0; f < cq; f += 4) {
            const uint4 jj = list[friend_quad(i, q, f >> 2)];
            const uint32_t v4[4] = {jj.x, jj.y, jj.z, jj.w};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (f + t < cq) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] = (u == na) ? v4[t] : acc[u];
                    if (++na == 4) {
                        quads[sweep_quad(i, qo++)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
                        na = 0;
                    }
                }
            }
        }
    }
    if (na) quads[sweep_quad(i, qo)] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
}

// Scans the global bucket [b.x, b.y] for particle i, four candidates at a time.
template <int BS>
__device__ __forceinline__ void scan_bucket_global(const float4& p, uint32_t i, uint2 b, uint32_t n,
                                                   const float4* __restrict__ xs, const SceneConsts& c,
                                                   uint4* __restrict__ nb_list, uint4* __restrict__ sweep,
                                                   FriendWriter<BS>& w) {
    if (b.x == kEmpty) return;
    const uint32_t e = min(b.y, n - 1);
    uint32_t j = b.x;
    for (; j + 3 <= e; j += 4) {
        const float4* __restrict__ g = xs + j;   // one address, the other three at immediate offsets
        const float4 q0 = g[0], q1 = g[1], q2 = g[2], q3 = g[3];
        const int c0 = friend_circle(p, q0, c), c1 = friend_circle(p, q1, c);
        const int c2 = friend_circle(p, q2, c), c3 = friend_circle(p, q3, c);
        w.take(i, j, c0, nb_list, sweep); w.take(i, j + 1, c1, nb_list, sweep);
        w.take(i, j + 2, c2, nb_list, sweep); w.take(i, j + 3, c3, nb_list, sweep);
    }
    // the last 1..3 candidates as one group: the loads clamped to e stay in
    // the bucket and the slots past e are marked out of range, so the wave
    // waits for one round of loads instead of up to three in turn
    if (j <= e) {
        const uint32_t r = e - j;   // 0..2
        const float4* __restrict__ g = xs + j;
        const float4 q0 = g[0], q1 = g[min(1u, r)], q2 = g[min(2u, r)];
        const int c0 = friend_circle(p, q0, c);
        const int c1 = (int)((uint32_t)friend_circle(p, q1, c) | ~lane_mask(r >= 1u));
        const int c2 = (int)((uint32_t)friend_circle(p, q2, c) | ~lane_mask(r >= 2u));
        w.take(i, j, c0, nb_list, sweep); w.take(i, j + 1, c1, nb_list, sweep); w.take(i, j + 2, c2, nb_list, sweep);
    }
}

// Neighbour cells that cannot hold a friend.  Cell k covers [k h, (k+1) h)
// for k > 0, (-h, h) for k = 0 and ((k-1) h, k h] for k < 0 (truncating
// conversion); a neighbour whose box is farther than h from the particle
// holds no particle within h.  Skipping it leaves the list unchanged provided
// its hash bucket is not also reached through another of the 27 offsets (the
// reference then lists that bucket's friends twice): with gridBufSize = 2^b,
// b >= 6, the 27 keys always differ (the cell coordinates differ mod 4), and
// particles of far cells folded into a bucket are never within h.  Returns
// bit (dx+1)*9 + (dy+1)*3 + (dz+1) set for each cell to scan; the 1e-4
// margin covers the rounding of x/h, k*h and the gap sums.
__device__ __forceinline__ float cell_lo(int k, float h) {
    return k > 0 ? (float)k * h : (k == 0 ? -h : (float)(k - 1) * h);
}
__device__ __forceinline__ float cell_hi(int k, float h) { return k >= 0 ? (float)(k + 1) * h : (float)k * h; }

__device__ __forceinline__ uint32_t cells_to_scan(const float4& p, int32_t cx, int32_t cy, int32_t cz,
                                                  const SceneConsts& c, uint32_t fresh) {
    // a paused step keeps the previous step's buckets (no sort): particles
    // need not lie in the box of the cell they are listed under
    if (!fresh || !(c.gbs_mask != 0u && c.gbs >= 64u)) return (1u << 27) - 1u;
    const float h = c.h;
    float g[3][3];
    const float pc[3] = {p.x, p.y, p.z};
    const int32_t cc[3] = {cx, cy, cz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = fmaxf(pc[a] - cell_hi(cc[a] - 1, h), 0.0f);
        const float hi = fmaxf(cell_lo(cc[a] + 1, h) - pc[a], 0.0f);
        g[a][0] = lo * lo;
        g[a][1] = 0.0f;
        g[a][2] = hi * hi;
    }
    const float lim = c.h2 * 1.0001f;
    uint32_t mask = 0u;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int z = 0; z < 3; ++z)
                mask |= (uint32_t)(!(g[0][x] + g[1][y] + g[2][z] > lim)) << (x * 9 + y * 3 + z);
    return mask;
}

// LAMBDA0 also runs the first computeScaling of the substep for the
// particle, from the list its thread has just written (scaling_of).
template <int BS, bool LAMBDA0, bool STRIDED>
__global__ __launch_bounds__(BS) void k_friends_direct(const SceneConsts* __restrict__ scenes,
                                                      const uint4* __restrict__ blocks, float4* __restrict__ xs,
                                                      const uint2* __restrict__ cells, uint32_t* __restrict__ nb_cnt,
                                                      uint32_t* __restrict__ nb_list, uint4* __restrict__ quads,
                                                      float* __restrict__ rho, float* __restrict__ lambda,
                                                      uint32_t n, uint32_t fresh, uint32_t* err,
                                                      uint32_t ntiles) {
    __shared__ uint32_t s_stage[4 * kCircles * BS];
    for (uint32_t tv = blockIdx.x; tv < ntiles; tv += gridDim.x) {   // tiles b, b + grid, ...
        const uint4 be = blocks[xcd_tile(tv, ntiles)];
        // a copy in registers: the list stores could alias the constants'
        // memory as far as the compiler knows, which would reload them
        const SceneConsts c = scenes[be.x];
        const uint32_t i = be.y + threadIdx.x;
        if (i < be.z) {
            uint4* list = reinterpret_cast<uint4*>(nb_list);
            const float4 p = xs[i];
            const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);
            FriendWriter<BS> w;
            w.stage = s_stage + threadIdx.x;
            w.bind(i, list, quads);
            const uint32_t scan = cells_to_scan(p, cx, cy, cz, c, fresh);
            const NeighbourHash nh(cx, cy, cz);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                        if (scan & (1u << ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1))))
                            scan_bucket_global(p, i, cells[nh(dx, dy, dz, c)], n, xs, c, list, quads, w);
            w.finish(i, nb_cnt, list, quads, err);
            pack_friends(i, w.packed, list, quads);
            if (LAMBDA0) scaling_of<false>(i, w.packed, xs, quads, rho, lambda, nullptr, c,
                                         [&](uint32_t j) { return xs[j]; });
        }
        if (!STRIDED) break;   // one workgroup per tile: a single pass
    }
}

// Five per-circle counters in registers, read and advanced at a circle
// chosen per lane (selects, not an indexed array).
struct CircleCounts {
    uint32_t v[kCircles] = {0u, 0u, 0u, 0u, 0u};
    __device__ __forceinline__ uint32_t at(int q) const {
        uint32_t r = v[0];
#pragma unroll
        for (int k = 1; k < kCircles; ++k) r = q == k ? v[k] : r;
        return r;
    }
    __device__ __forceinline__ void inc(int q) {
#pragma unroll
        for (int k = 0; k < kCircles; ++k) v[k] += q == k ? 1u : 0u;
    }
};

// Bucket ranges {first, last} of a lane's offsets [k0, k0 + U) (U at most
// MAXOFF; a culled or missing one is empty), loaded together: the neighbour's
// hash stepped from the centre's dilated coordinates (NeighbourHash's steps)
// chosen per lane (k differs between a particle's lanes; an indexed choice
// would go through memory), each range one 8-byte load.
template <int MAXOFF>
__device__ __forceinline__ void load_ranges(uint32_t scan, const NeighbourHash& nh, int k0, int k1,
                                            const uint2* __restrict__ cells, const SceneConsts& c,
                                            uint64_t* rng) {
    auto step = [](uint32_t e, int a) {
        constexpr uint32_t kM = NeighbourHash::kM;
        return a == 0 ? ((e - 1u) & kM) : (a == 1 ? e : (((e | ~kM) + 1u) & kM));
    };
#pragma unroll
    for (int u = 0; u < MAXOFF; ++u) {
        const int k = k0 + u;
        rng[u] = 0xFFFFFFFFFFFFFFFFull;
        if (k < k1 && (scan & (1u << k))) {
            const uint32_t m = step(nh.x[1], k / 9) | (step(nh.y[1], (k / 3) % 3) << 1) | (step(nh.z[1], k % 3) << 2);
            rng[u] = reinterpret_cast<const uint64_t*>(cells)[c.cell_base + (c.gbs_mask ? (m & c.gbs_mask)
                                                                                          : (m % c.gbs))];
        }
    }
}

// The buckets of a lane's ranges in order, candidates four at a time; each
// friend of circle q advances cnt[q] and, with STORE, is stored at place
// base[q] + its index among the lane's circle-q friends (if under the 50
// cap).
template <int MAXOFF, bool STORE>
__device__ __forceinline__ void scan_ranges(const float4& p, uint32_t i, const uint64_t* rng,
                                            const float4* __restrict__ xs, uint32_t n, const SceneConsts& c,
                                            CircleCounts& cnt, const CircleCounts base, uint32_t* __restrict__ list,
                                            uint32_t* __restrict__ sweep) {
#pragma unroll
    for (int u = 0; u < MAXOFF; ++u) {
        const uint32_t bs = (uint32_t)rng[u], e = min((uint32_t)(rng[u] >> 32), n - 1);
        if (bs == kEmpty) continue;
        for (uint32_t j = bs; j <= e; j += 4) {
            const uint32_t r = min(e - j, 3u);
            const float4* __restrict__ g = xs + j;
            const float4 q[4] = {g[0], g[min(1u, r)], g[min(2u, r)], g[min(3u, r)]};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int circ = friend_circle(p, q[t], c);
                if ((uint32_t)t > r || circ < 0 || j + t == i) continue;
                if (STORE) {
                    const uint32_t at = base.at(circ) + cnt.at(circ);
                    if (at < (uint32_t)kPerCircle) {
                        if (circ == 0)
                            sweep[sweep_quad(i, (int)(at >> 2)) * 4 + (at & 3u)] = j + t;
                        else
                            list[friend_slot(i, circ, (int)at)] = j + t;
                    }
                }
                cnt.inc(circ);
            }
        }
    }
}

// The friends build with T lanes per particle, for scenes too small to fill
// the device with one thread per particle (the build's latency is then that
// of one thread scanning its 27 cells).  Each workgroup takes a 64-particle
// tile with 64 T threads; lane l of a particle scans its offsets
// [27 l / T, 27 (l + 1) / T) in the x, y, z order, so the particle's list is
// its lanes' lists one after another.  The lanes first count their friends
// per circle, then take exclusive prefix sums of the counts across the
// particle's lanes, then scan again and store each friend at its place in
// the circle (friends past the 50 cap are dropped, as in the one-lane
// build).  The first lane of the particle then stores the counts, regroups
// the processed circles (pack_friends) and runs the first computeScaling.
template <int T, bool LAMBDA0>
__global__ __launch_bounds__(64 * T) void k_friends_lanes(const SceneConsts* __restrict__ scenes,
                                                         const uint4* __restrict__ blocks, float4* __restrict__ xs,
                                                         const uint2* __restrict__ cells,
                                                         uint32_t* __restrict__ nb_cnt, uint32_t* __restrict__ nb_list,
                                                         uint4* __restrict__ quads, float* __restrict__ rho,
                                                         float* __restrict__ lambda, uint32_t n, uint32_t fresh,
                                                         uint32_t* err) {
    static_assert(T >= 2 && T <= 16 && (T & (T - 1)) == 0, "T lanes per particle: a power of two up to 16");
    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];   // a tile of at most 64 particles
    const SceneConsts c = scenes[be.x];
    const int l = (int)(threadIdx.x % T);
    const uint32_t i = be.y + threadIdx.x / T;
    // a particle's T lanes are consecutive in one wave; lanes of particles
    // past the tile take part in the shuffles with nothing to scan
    const bool valid = i < be.z;
    const uint32_t ii = valid ? i : be.y;
    const float4 p = xs[ii];
    const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);
    const uint32_t scan = valid ? cells_to_scan(p, cx, cy, cz, c, fresh) : 0u;
    const NeighbourHash nh(cx, cy, cz);
    const int k0 = 27 * l / T, k1 = 27 * (l + 1) / T;
    constexpr int kMaxOff = (27 + T - 1) / T;
    uint64_t rng[kMaxOff];
    load_ranges<kMaxOff>(scan, nh, k0, k1, cells, c, rng);
    CircleCounts cnt;
    scan_ranges<kMaxOff, false>(p, ii, rng, xs, n, c, cnt, cnt, nullptr, nullptr);
    // friends of the particle's earlier lanes, per circle, and its totals
    CircleCounts base, total;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) {
        uint32_t incl = cnt.v[q];
#pragma unroll
        for (int o = 1; o < T; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, T);
            if (l >= o) incl += u;
        }
        base.v[q] = incl - cnt.v[q];
        total.v[q] = __shfl(incl, T - 1, T);
        cnt.v[q] = 0u;
    }
    uint32_t* list = nb_list;
    uint32_t* sweep = reinterpret_cast<uint32_t*>(quads);
    scan_ranges<kMaxOff, true>(p, ii, rng, xs, n, c, cnt, base, list, sweep);
    // the lanes' stores are complete before the first lane reads the
    // circles back (the lanes share a wave and its L1)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (!valid || l != 0) return;
    uint32_t packed = 0, dropped = 0;
#pragma unroll
    for (int q = 0; q < kCircles; ++q) {
        const uint32_t kept = min(total.v[q], (uint32_t)kPerCircle);
        packed |= kept << (6 * q);
        dropped += total.v[q] - kept;
    }
    nb_cnt[i] = packed;
    if (dropped) {
        atomicOr(err, (uint32_t)PBF_FLAG_FRIENDS_OVERFLOW);
        atomicAdd((unsigned long long*)(err + 2), (unsigned long long)dropped);
        log_event(err, PBF_LOG_FRIENDS_OVERFLOW, i, dropped, packed);
    }
    pack_friends(i, packed, reinterpret_cast<const uint4*>(list), quads);
    if (LAMBDA0) scaling_of<false>(i, packed, xs, quads, rho, lambda, nullptr, c, [&](uint32_t j) { return xs[j]; });
}

// Staging budget of the LDS variant per workgroup size: box cells, particles.
template <int BS> struct FriendsBudget;
template <> struct FriendsBudget<64> { static constexpr int cells = 128, staged = 1024; };
template <> struct FriendsBudget<128> { static constexpr int cells = 256, staged = 1536; };
template <> struct FriendsBudget<256> { static constexpr int cells = 512, staged = 2560; };
template <> struct FriendsBudget<512> { static constexpr int cells = 1024, staged = 4096; };

template <int BS, bool LAMBDA0>
__global__ __launch_bounds__(BS) void k_friends(const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks,
                                               float4* __restrict__ xs, const uint2* __restrict__ cells,
                                               uint32_t* __restrict__ nb_cnt, uint32_t* __restrict__ nb_list,
                                               uint4* __restrict__ quads, float* __restrict__ rho,
                                               float* __restrict__ lambda, uint32_t n, uint32_t fresh, uint32_t* err) {
    constexpr int kWaves = BS / 64;
    constexpr int kMaxCells = FriendsBudget<BS>::cells;
    constexpr int kMaxStaged = FriendsBudget<BS>::staged;
    static_assert(kMaxCells <= 2 * BS, "the cell scan gives each thread two cells");
    __shared__ float4 s_pos[kMaxStaged];
    __shared__ uint32_t s_start[kMaxCells];
    __shared__ uint32_t s_off[kMaxCells + 1];
    __shared__ int s_red[6][kWaves];
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint32_t s_stage[4 * kCircles * BS];

    const uint4 be = blocks[xcd_tile(blockIdx.x, gridDim.x)];
    const SceneConsts c = scenes[be.x];   // in registers (see k_friends_direct)
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const uint32_t i = be.y + tid;
    const bool valid = i < be.z;
    const float4 p = valid ? xs[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);

    // cell box of the tile
    const int v0 = wave_min(valid ? cx : INT32_MAX), v1 = wave_min(valid ? cy : INT32_MAX);
    const int v2 = wave_min(valid ? cz : INT32_MAX), v3 = wave_max(valid ? cx : INT32_MIN);
    const int v4 = wave_max(valid ? cy : INT32_MIN), v5 = wave_max(valid ? cz : INT32_MIN);
    if (lane == 0) {
        s_red[0][wv] = v0; s_red[1][wv] = v1; s_red[2][wv] = v2;
        s_red[3][wv] = v3; s_red[4][wv] = v4; s_red[5][wv] = v5;
    }
    __syncthreads();
    int lo0 = INT32_MAX, lo1 = INT32_MAX, lo2 = INT32_MAX, hi0 = INT32_MIN, hi1 = INT32_MIN, hi2 = INT32_MIN;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
        lo0 = min(lo0, s_red[0][q]); lo1 = min(lo1, s_red[1][q]); lo2 = min(lo2, s_red[2][q]);
        hi0 = max(hi0, s_red[3][q]); hi1 = max(hi1, s_red[4][q]); hi2 = max(hi2, s_red[5][q]);
    }
    const int64_t dxs = (int64_t)hi0 - lo0 + 3, dys = (int64_t)hi1 - lo1 + 3, dzs = (int64_t)hi2 - lo2 + 3;
    const int64_t ncell64 = dxs * dys * dzs;
    bool staged = ncell64 <= kMaxCells;
    const int ncell = staged ? (int)ncell64 : 0;
    const int bx = lo0 - 1, by = lo1 - 1, bz = lo2 - 1;
    const int ddx = (int)dxs, ddy = (int)dys;

    if (staged) {
        // bucket of every box cell; counts in s_off, prefix-summed below
        for (int q = tid; q < ncell; q += BS) {
            const int ix = q % ddx, iy = (q / ddx) % ddy, iz = q / (ddx * ddy);
            const uint2 b = cells[cell_hash(bx + ix, by + iy, bz + iz, c)];
            uint32_t cnt = 0;
            if (b.x != kEmpty) {
                const uint32_t e = min(b.y, n - 1);
                cnt = e >= b.x ? e - b.x + 1 : 0;
            }
            s_start[q] = b.x;
            s_off[q] = cnt;
        }
        __syncthreads();
        // exclusive scan of s_off[0..ncell): each thread owns 2 consecutive cells
        const int q0 = 2 * tid, q1 = 2 * tid + 1;
        const uint32_t a0 = q0 < ncell ? s_off[q0] : 0u, a1 = q1 < ncell ? s_off[q1] : 0u;
        uint32_t incl = a0 + a1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) s_scan[wv] = incl;
        __syncthreads();
        uint32_t base = 0, total = 0;
        for (int q = 0; q < kWaves; ++q) {
            if (q < wv) base += s_scan[q];
            total += s_scan[q];
        }
        const uint32_t excl = base + incl - (a0 + a1);
        if (q0 < ncell) s_off[q0] = excl;
        if (q1 < ncell) s_off[q1] = excl + a0;
        if (tid == 0) s_off[ncell] = total;
        staged = total <= (uint32_t)kMaxStaged;
        __syncthreads();
        if (staged) {
            for (uint32_t s = tid; s < total; s += BS) {
                // box cell holding slot s: the last cell whose offset is <= s
                int lo_c = 0, hi_c = ncell - 1;
                while (lo_c < hi_c) {
                    const int mid = (lo_c + hi_c + 1) >> 1;
                    if (s_off[mid] <= s) lo_c = mid; else hi_c = mid - 1;
                }
                s_pos[s] = xs[s_start[lo_c] + (s - s_off[lo_c])];
            }
            __syncthreads();
        }
    }
    if (!staged && tid == 0) atomicAdd(err + 4, 1u);
    if (!valid) return;

    uint4* list = reinterpret_cast<uint4*>(nb_list);
    FriendWriter<BS> w;
    w.stage = s_stage + tid;
    w.bind(i, list, quads);
    const uint32_t scan = cells_to_scan(p, cx, cy, cz, c, fresh);
    const NeighbourHash nh(cx, cy, cz);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (!(scan & (1u << ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1))))) continue;
                if (staged) {
                    const int q = (cx + dx - bx) + ddx * ((cy + dy - by) + ddy * (cz + dz - bz));
                    const uint32_t off = s_off[q], cnt = s_off[q + 1] - off, gs = s_start[q];
                    uint32_t k = 0;
                    for (; k + 4 <= cnt; k += 4) {
                        const float4 q0 = s_pos[off + k], q1 = s_pos[off + k + 1];
                        const float4 q2 = s_pos[off + k + 2], q3 = s_pos[off + k + 3];
                        const int c0 = friend_circle(p, q0, c), c1 = friend_circle(p, q1, c);
                        const int c2 = friend_circle(p, q2, c), c3 = friend_circle(p, q3, c);
                        w.take(i, gs + k, c0, list, quads); w.take(i, gs + k + 1, c1, list, quads);
                        w.take(i, gs + k + 2, c2, list, quads); w.take(i, gs + k + 3, c3, list, quads);
                    }
                    for (; k < cnt; ++k) w.take(i, gs + k, friend_circle(p, s_pos[off + k], c), list, quads);
                } else {
                    scan_bucket_global(p, i, cells[nh(dx, dy, dz, c)], n, xs, c, list, quads, w);
                }
            }
    w.finish(i, nb_cnt, list, quads, err);
    pack_friends(i, w.packed, list, quads);
    if (LAMBDA0) scaling_of<false>(i, w.packed, xs, quads, rho, lambda, nullptr, c,
                                         [&](uint32_t j) { return xs[j]; });
}

// ---------------------------------------------------------------------------
// Friend sweeps.  Each walks the friends of the circles it processes in the
// reference's order (circle-major, then scan order), four per 16-byte load
// (walk_friends, pbf_device.h): the gathers of the next quad are issued
// before the updates of the current one run in order.
// ---------------------------------------------------------------------------

// Two float4 of one friend (position and velocity or vorticity).
struct Pair4 {
    float4 a, b;
};

// TAPE also stores (sum of gradients, denominator) of every particle for the
// adjoint (aux).
template <int BS, bool TAPE, bool STRIDED>
__global__ __launch_bounds__(BS) void k_lambda(const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks,
                                              float4* __restrict__ xs, const uint32_t* __restrict__ nb_cnt,
                                              const uint4* __restrict__ nb_list, float* __restrict__ rho,
                                              float* __restrict__ lambda, float4* __restrict__ aux, uint32_t n,
                                              uint32_t near_on, uint32_t ntiles) {
    __shared__ float4 s_x[NearWindow<BS>::kLen];
    for (uint32_t tv = blockIdx.x; tv < ntiles; tv += gridDim.x) {   // tiles b, b + grid, ...
        const uint4 be = blocks[xcd_tile(tv, ntiles)];
        const SceneConsts& c = scenes[be.x];
        const NearWindow<BS> near(be, n, near_on);
        near.stage(s_x, xs);   // w is rewritten below, but only x, y, z are gathered
        __syncthreads();
        const uint32_t i = be.y + threadIdx.x;
        if (i < be.z)
            scaling_of<TAPE>(i, nb_cnt[i], xs, nb_list, rho, lambda, aux, c,
                             [&](uint32_t j) { return near.get(s_x, xs, j); });
        if (!STRIDED) break;   // one workgroup per tile: a single pass
        __syncthreads();       // the window of the next tile
    }
}

// compute_delta.cl:44-186 + update_predicted.cl; on the last iteration also
// pack_data.cl (density into x*.w) and update_velocities.cl.
struct DeltaAcc {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    __device__ __forceinline__ void add(const float4& p, const float4& pj, const SceneConsts& c, bool valid) {
        const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
        const float r2 = dist2(rx, ry, rz);
        const uint32_t in = lane_mask(valid && r2 < c.h2);   // left out otherwise (as in LambdaAcc)
        const PairLen L = pair_len_sweep(r2);
        const float hr = c.h - L.rl;
        const float t = L.inv * hr * hr;
        const float ax = rx * t, ay = ry * t, az = rz * t;
        const float d = c.h2 - r2;
        const float p6 = d * d * d;
        const float ratio = p6 * c.inv_poly6_q;
        const float r4 = ratio * ratio;
        const float coef = fmaf(c.st_k, r4 * r4, p.w + pj.w);   // lambda_i + lambda_j + s_corr
        sx = picked(fmaf(coef, ax, sx), sx, in);
        sy = picked(fmaf(coef, ay, sy), sy, in);
        sz = picked(fmaf(coef, az, sz), sz, in);
    }
};

// TAPE also stores the position before the walls and which planes and quads
// moved it (bits in w, as apply_walls returns them) for the adjoint (fut).
template <int BS, bool LAST, bool TAPE, bool STRIDED>
__global__ __launch_bounds__(BS) void k_delta(const SceneConsts* __restrict__ scenes, const uint4* __restrict__ blocks,
                                             const float4* __restrict__ xs, float4* __restrict__ xs_next,
                                             float4* __restrict__ x, float4* __restrict__ v,
                                             const float* __restrict__ rho, const uint32_t* __restrict__ nb_cnt,
                                             const uint4* __restrict__ nb_list, float4* __restrict__ fut, uint32_t n,
                                             uint32_t near_on, uint32_t ntiles) {
    __shared__ float4 s_x[NearWindow<BS>::kLen];
    for (uint32_t tv = blockIdx.x; tv < ntiles; tv += gridDim.x) {   // tiles b, b + grid, ...
        const uint4 be = blocks[xcd_tile(tv, ntiles)];
        const SceneConsts& c = scenes[be.x];
        const NearWindow<BS> near(be, n, near_on);
        near.stage(s_x, xs);
        __syncthreads();
        const uint32_t i = be.y + threadIdx.x;
        if (i < be.z) {
            const float4 p = xs[i];
            DeltaAcc a;
            walk_friends<1, 2, true>(
                i, nb_cnt[i], nb_list, [&](uint32_t j) { return near.get(s_x, xs, j); },
                [&](const float4& q, bool valid) { a.add(p, q, c, valid); });
            float fx = p.x + (-c.spiky * a.sx) / c.rest_density;
            float fy = p.y + (-c.spiky * a.sy) / c.rest_density;
            float fz = p.z + (-c.spiky * a.sz) / c.rest_density;
            const float4 prev = x[i];
            const float ux = fx, uy = fy, uz = fz;
            uint32_t hits = 0;
            if (c.walls | c.quad_count) hits = apply_walls(fx, fy, fz, prev, c);
            if (TAPE) fut[i] = make_float4(ux, uy, uz, __uint_as_float(hits));
            const float dx = fx - p.x, dy = fy - p.y, dz = fz - p.z;
            float4 o;
            o.x = p.x + dx;
            o.y = p.y + dy;
            o.z = p.z + dz;
            if (!LAST) {
                o.w = p.w;
                xs_next[i] = o;
            } else {
                o.w = rho[i];
                xs_next[i] = o;
                const float vx = (o.x - prev.x) / c.dt, vy = (o.y - prev.y) / c.dt, vz = (o.z - prev.z) / c.dt;
                float* vp = reinterpret_cast<float*>(v + i);
                vp[0] = vx;
                vp[1] = vy;
                vp[2] = vz;
                x[i] = make_float4(o.x, o.y, o.z, sqrtf(vx * vx + vy * vy + vz * vz));
            }
        }
        if (!STRIDED) break;   // one workgroup per tile: a single pass
        __syncthreads();       // the window of the next tile
    }
}

// apply_viscosity.cl:1-87, Jacobi: v_j from the snapshot v_in.
struct ViscAcc {
    float vsx = 0.0f, vsy = 0.0f, vsz = 0.0f, ox = 0.0f, oy = 0.0f, oz = 0.0f;
    __device__ __forceinline__ void add(const float4& p, const float4& vi, const float4& pj, const float4& vj,
                                        const SceneConsts& c, bool valid) {
        const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
        const float r2 = dist2(rx, ry, rz);
        // a skipped pair leaves the sums as they are (as in LambdaAcc)
        const uint32_t in = lane_mask(valid && r2 < c.h2 && fabsf(pj.w) > 1e-8f);
        const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
        const float d = c.h2 - r2;
        const PairLen L = pair_len_sweep(r2);
        const float hr = c.h - L.rl;
        const float t = L.inv * hr * hr;
        const float ax = rx * t, ay = ry * t, az = rz * t;
        ox = picked(ox + fmaf(uy, az, -(uz * ay)), ox, in);   // (v_j - v_i) x grad
        oy = picked(oy + fmaf(uz, ax, -(ux * az)), oy, in);
        oz = picked(oz + fmaf(ux, ay, -(uy * ax)), oz, in);
        const float p6 = d * d * d;
        vsx = picked(fmaf(ux, p6, vsx), vsx, in);
        vsy = picked(fmaf(uy, p6, vsy), vsy, in);
        vsz = picked(fmaf(uz, p6, vsz), vsz, in);
    }
};

template <int BS, bool STRIDED>
__global__ __launch_bounds__(BS) void k_viscosity(const SceneConsts* __restrict__ scenes,
                                                 const uint4* __restrict__ blocks, const float4* __restrict__ xs,
                                                 const float4* __restrict__ v_in, float4* __restrict__ v_out,
                                                 float4* __restrict__ omega, const uint32_t* __restrict__ nb_cnt,
                                                 const uint4* __restrict__ nb_list, uint32_t n, uint32_t near_on,
                                                 uint32_t ntiles) {
    __shared__ float4 s_x[NearWindow<BS>::kLen], s_v[NearWindow<BS>::kLen];
    for (uint32_t tv = blockIdx.x; tv < ntiles; tv += gridDim.x) {   // tiles b, b + grid, ...
        const uint4 be = blocks[xcd_tile(tv, ntiles)];
        const SceneConsts& c = scenes[be.x];
        const NearWindow<BS> near(be, n, near_on);
        near.stage(s_x, xs);
        near.stage(s_v, v_in);
        __syncthreads();
        const uint32_t i = be.y + threadIdx.x;
        if (i < be.z) {
            const float4 p = xs[i];
            const float4 vi = v_in[i];
            ViscAcc a;
            walk_friends<1, 2, true>(
                i, nb_cnt[i], nb_list,
                [&](uint32_t j) { return Pair4{near.get(s_x, xs, j), near.get(s_v, v_in, j)}; },
                [&](const Pair4& t, bool valid) { a.add(p, vi, t.a, t.b, c, valid); });
            float4 vo;
            vo.x = vi.x + c.viscosity * (a.vsx * c.poly6);
            vo.y = vi.y + c.viscosity * (a.vsy * c.poly6);
            vo.z = vi.z + c.viscosity * (a.vsz * c.poly6);
            vo.w = vi.w + c.viscosity * 0.0f;
            v_out[i] = vo;
            // w carries |omega| (the reference stores 0, apply_viscosity.cl:86),
            // so the vorticity sweep reads each neighbour's length instead of
            // taking a square root per pair
            const float ox = a.ox * -c.spiky, oy = a.oy * -c.spiky, oz = a.oz * -c.spiky;
            omega[i] = make_float4(ox, oy, oz, norm_of_sq(ox * ox + oy * oy + oz * oz));
        }
        if (!STRIDED) break;   // one workgroup per tile: a single pass
        __syncthreads();       // the window of the next tile
    }
}

// apply_vorticity.cl:1-87
struct VortAcc {
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    __device__ __forceinline__ void add(const float4& p, const float4& pj, const float4& oj, const SceneConsts& c,
                                        bool valid) {
        const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
        const float r2 = dist2(rx, ry, rz);
        const uint32_t in = lane_mask(valid && r2 < c.h2 && fabsf(pj.w) > 1e-8f);
        const PairLen L = pair_len_sweep(r2);
        const float hr = c.h - L.rl;
        const float t = L.inv * hr * hr;
        const float ax = rx * t, ay = ry * t, az = rz * t;
        const float wl = oj.w;   // |omega_j| (fast_length, apply_vorticity.cl:63), stored by k_viscosity
        ex = picked(fmaf(wl, ax, ex), ex, in);
        ey = picked(fmaf(wl, ay, ey), ey, in);
        ez = picked(fmaf(wl, az, ez), ez, in);
    }
};

template <int BS, bool STRIDED>
__global__ __launch_bounds__(BS) void k_vorticity(const SceneConsts* __restrict__ scenes,
                                                 const uint4* __restrict__ blocks, const float4* __restrict__ xs,
                                                 float4* __restrict__ v, const float4* __restrict__ omega,
                                                 const uint32_t* __restrict__ nb_cnt,
                                                 const uint4* __restrict__ nb_list, uint32_t n, uint32_t near_on,
                                                 uint32_t ntiles) {
    __shared__ float4 s_x[NearWindow<BS>::kLen], s_o[NearWindow<BS>::kLen];
    for (uint32_t tv = blockIdx.x; tv < ntiles; tv += gridDim.x) {   // tiles b, b + grid, ...
        const uint4 be = blocks[xcd_tile(tv, ntiles)];
        const SceneConsts& c = scenes[be.x];
        const NearWindow<BS> near(be, n, near_on);
        near.stage(s_x, xs);
        near.stage(s_o, omega);
        __syncthreads();
        const uint32_t i = be.y + threadIdx.x;
        if (i < be.z) {
            const float4 p = xs[i];
            VortAcc a;
            walk_friends<1, 2, true>(
                i, nb_cnt[i], nb_list,
                [&](uint32_t j) { return Pair4{near.get(s_x, xs, j), near.get(s_o, omega, j)}; },
                [&](const Pair4& t, bool valid) { a.add(p, t.a, t.b, c, valid); });
            float ex = a.ex * -c.spiky, ey = a.ey * -c.spiky, ez = a.ez * -c.spiky;
            const float l = sqrtf(ex * ex + ey * ey + ez * ez);
            if (l > 0.0f) { ex /= l; ey /= l; ez /= l; }
            const float4 oi = omega[i];
            const float cx = ey * oi.z - ez * oi.y;
            const float cy = ez * oi.x - ex * oi.z;
            const float cz = ex * oi.y - ey * oi.x;
            float4 vi = v[i];
            vi.x += (c.vorticity * cx) * c.dt;
            vi.y += (c.vorticity * cy) * c.dt;
            vi.z += (c.vorticity * cz) * c.dt;
            v[i] = vi;
        }
        if (!STRIDED) break;   // one workgroup per tile: a single pass
        __syncthreads();       // the window of the next tile
    }
}

}  // namespace

// Launch a kernel template at the workgroup size chosen for its family.
#define PBF_DISPATCH_BS(LOG2, LAUNCH)                        \
    switch (LOG2) {                                          \
        case 6: { constexpr int BS = 64; LAUNCH; } break;    \
        case 7: { constexpr int BS = 128; LAUNCH; } break;   \
        case 9: { constexpr int BS = 512; LAUNCH; } break;   \
        default: { constexpr int BS = 256; LAUNCH; } break;  \
    }

// Workgroups of a launch over `tiles` tiles for family f: one per tile, or,
// when the family runs k > 1 tiles per workgroup (or persistently, k = 0),
// fewer: a multiple of the XCD count, so that the dispatch indices a
// workgroup strides over (b, b + grid, ...) stay on its XCD's run of tiles
// (xcd_tile).
uint32_t launch_grid(const pbf_ctx* ctx, Family f, uint32_t tiles) {
    const uint32_t k = ctx->tiles_per_wg[f];
    if (k == 1 || tiles <= kXcds) return tiles;
    uint32_t g = k == 0 ? ctx->resident[f][ctx->bs_log2[f] - 6] : (tiles + k - 1) / k;
    g = std::max<uint32_t>(kXcds, g / kXcds * kXcds);
    return std::min(g, tiles);
}

template <class K>
uint32_t resident_blocks(const pbf_ctx* ctx, K kernel, int bs) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, 0) != hipSuccess) {
        (void)hipGetLastError();   // clear it: the launchers report hipGetLastError
        per_cu = 1;
    }
    return (uint32_t)std::max(per_cu, 1) * (uint32_t)std::max(ctx->cu_count, 1);
}

hipError_t query_residency(pbf_ctx* ctx) {
    for (int l = 6; l <= 9; ++l) {
        PBF_DISPATCH_BS(l, {
            ctx->resident[F_FRIENDS][l - 6] = resident_blocks(ctx, k_friends_direct<BS, true, true>, BS);
            ctx->resident[F_LAMBDA][l - 6] = resident_blocks(ctx, k_lambda<BS, false, true>, BS);
            ctx->resident[F_DELTA][l - 6] = resident_blocks(ctx, k_delta<BS, false, false, true>, BS);
            ctx->resident[F_VISC][l - 6] = resident_blocks(ctx, k_viscosity<BS, true>, BS);
            ctx->resident[F_VORT][l - 6] = resident_blocks(ctx, k_vorticity<BS, true>, BS);
        });
    }
    return hipSuccess;   // a failed occupancy query falls back to one workgroup per CU
}

// Whether the sweeps copy their tiles' neighbourhoods into LDS (NearWindow):
// only on grids of kNearMinTilesPerCu tiles per CU or more.
static uint32_t near_on(const pbf_ctx* ctx, uint32_t tiles) {
    return tiles >= kNearMinTilesPerCu * (uint32_t)std::max(ctx->cu_count, 1) ? 1u : 0u;
}

hipError_t launch_predict_hash(pbf_ctx* ctx, bool reset_old) {
    track_begin(ctx, "predictPositions");
    PBF_KERNEL(k_predict_hash, dim3(ctx->bt_predict.count), dim3(kPredictThreads), 0, ctx->stream, ctx->sc_d,
                       ctx->bt_predict.d, ctx->x, ctx->v, ctx->keys_in, ctx->keys_sorted,
                       reinterpret_cast<uint2*>(ctx->cells), reset_old ? 1u : 0u, ctx->total_cells, ctx->sort.hist,
                       ctx->sort.passes, ctx->pause_now, ctx->err);
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_predict_only(pbf_ctx* ctx, uint32_t pause) {
    track_begin(ctx, "predictPositions");
    const BlockTable& t = ctx->bt[2];
    PBF_KERNEL(k_predict_only, dim3(t.count), dim3(256), 0, ctx->stream, ctx->sc_d, t.d, ctx->x, ctx->v,
                       ctx->xs[0], pause);
    track_end(ctx);
    ctx->xs_cur = 0;
    return hipGetLastError();
}

hipError_t launch_gather(pbf_ctx* ctx) {
    track_begin(ctx, "sortParticles");
    const BlockTable& t = ctx->bt[2];
    PBF_KERNEL(k_gather_cells, dim3(t.count), dim3(256), 0, ctx->stream, ctx->sc_d, t.d, ctx->perm, ctx->x,
                       ctx->v, ctx->x_alt, ctx->xs[0], ctx->keys_sorted, ctx->cells, ctx->total_cells, ctx->sort.ws,
                       ctx->sort.ws_words, ctx->pause_now, ctx->n, ctx->err);
    track_end(ctx);
    float4* tmp = ctx->x;
    ctx->x = ctx->x_alt;
    ctx->x_alt = tmp;
    ctx->xs_cur = 0;
    return hipGetLastError();
}

// One launch of a tiled kernel: KERNEL<..., STRIDED> over the family's
// block table, single-pass when the grid has one workgroup per tile.  KERNEL
// is the template name and its leading arguments without the closing '>'
// (COMMA separates them).
#define COMMA ,
#define PBF_LAUNCH_TILED(FAM, KERNEL, ...)                                                                     \
    do {                                                                                                       \
        const uint32_t grid_ = launch_grid(ctx, FAM, t.count);                                                 \
        PBF_DISPATCH_BS(ctx->bs_log2[FAM], {                                                                   \
            if (grid_ < t.count)                                                                               \
                PBF_KERNEL((KERNEL, true>), dim3(grid_), dim3(BS), 0, ctx->stream, __VA_ARGS__, t.count);  \
            else                                                                                               \
                PBF_KERNEL((KERNEL, false>), dim3(grid_), dim3(BS), 0, ctx->stream, __VA_ARGS__, t.count); \
        });                                                                                                    \
    } while (0)

uint32_t friends_lanes_for(const pbf_ctx* ctx) {
    if (ctx->friends_lanes) return ctx->friends_lanes;
    // The lanes build scans twice and stores friends one by one, so it pays
    // only while one lane per particle leaves most of the device idle: 16
    // lanes up to 2 waves per CU (32,768 particles on 256 CUs; measured at
    // 4,096 and 16,384 particles), one lane above (slower with any count of
    // lanes from 65,536 particles on).
    const uint64_t waves = ((uint64_t)ctx->n + 63) / 64;
    return waves <= 2ull * (uint64_t)std::max(ctx->cu_count, 1) ? 16u : 1u;
}

hipError_t launch_friends(pbf_ctx* ctx, bool lambda0) {
    track_begin(ctx, "buildFriendsList");
    const uint32_t fresh = ctx->cells_fresh ? 1u : 0u;
    const BlockTable& t = table_for(ctx, F_FRIENDS);
    float4* xs = ctx->xs[ctx->xs_cur];
    const uint2* cells = reinterpret_cast<const uint2*>(ctx->cells);
    const uint32_t lanes = ctx->friends_staged ? 1u : friends_lanes_for(ctx);
    if (lanes > 1) {
        // 64-particle tiles, 64 * lanes threads each
        const BlockTable& t64 = ctx->bt[0];
#define PBF_LAUNCH_LANES(T)                                                                                        \
    do {                                                                                                           \
        if (lambda0)                                                                                               \
            PBF_KERNEL((k_friends_lanes<T, true>), dim3(t64.count), dim3(64 * T), 0, ctx->stream, ctx->sc_d, t64.d, \
                       xs, cells, ctx->nb_cnt, ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n, fresh,    \
                       ctx->err);                                                                                  \
        else                                                                                                       \
            PBF_KERNEL((k_friends_lanes<T, false>), dim3(t64.count), dim3(64 * T), 0, ctx->stream, ctx->sc_d,      \
                       t64.d, xs, cells, ctx->nb_cnt, ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n,  \
                       fresh, ctx->err);                                                                           \
    } while (0)
        switch (lanes) {
            case 2: PBF_LAUNCH_LANES(2); break;
            case 4: PBF_LAUNCH_LANES(4); break;
            case 8: PBF_LAUNCH_LANES(8); break;
            default: PBF_LAUNCH_LANES(16); break;
        }
#undef PBF_LAUNCH_LANES
    } else if (ctx->friends_staged) {
        PBF_DISPATCH_BS(ctx->bs_log2[F_FRIENDS], {
            if (lambda0)
                PBF_KERNEL((k_friends<BS, true>), dim3(t.count), dim3(BS), 0, ctx->stream, ctx->sc_d, t.d, xs,
                                   cells, ctx->nb_cnt, ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n,
                                   fresh, ctx->err);
            else
                PBF_KERNEL((k_friends<BS, false>), dim3(t.count), dim3(BS), 0, ctx->stream, ctx->sc_d, t.d, xs,
                                   cells, ctx->nb_cnt, ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n,
                                   fresh, ctx->err);
        });
    } else if (lambda0) {
        PBF_LAUNCH_TILED(F_FRIENDS, k_friends_direct<BS COMMA true, ctx->sc_d, t.d, xs, cells, ctx->nb_cnt,
                         ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n, fresh, ctx->err);
    } else {
        PBF_LAUNCH_TILED(F_FRIENDS, k_friends_direct<BS COMMA false, ctx->sc_d, t.d, xs, cells, ctx->nb_cnt,
                         ctx->nb_list, ctx->nb_quads, ctx->rho, ctx->lambda, ctx->n, fresh, ctx->err);
    }
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_lambda(pbf_ctx* ctx, int iter, float4* aux) {
    track_begin(ctx, "computeScaling", iter);
    const BlockTable& t = table_for(ctx, F_LAMBDA);
    if (aux)
        PBF_LAUNCH_TILED(F_LAMBDA, k_lambda<BS COMMA true, ctx->sc_d, t.d, ctx->xs[ctx->xs_cur], ctx->nb_cnt,
                         ctx->nb_quads, ctx->rho, ctx->lambda, aux, ctx->n, near_on(ctx, t.count));
    else
        PBF_LAUNCH_TILED(F_LAMBDA, k_lambda<BS COMMA false, ctx->sc_d, t.d, ctx->xs[ctx->xs_cur], ctx->nb_cnt,
                         ctx->nb_quads, ctx->rho, ctx->lambda, aux, ctx->n, near_on(ctx, t.count));
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_delta(pbf_ctx* ctx, int iter, bool last, float4* fut) {
    track_begin(ctx, "computeDelta", iter);
    const int cur = ctx->xs_cur;
    const BlockTable& t = table_for(ctx, F_DELTA);
#define PBF_LAUNCH_DELTA(LASTV, TAPEV)                                                                           \
    PBF_LAUNCH_TILED(F_DELTA, k_delta<BS COMMA LASTV COMMA TAPEV, ctx->sc_d, t.d, ctx->xs[cur], ctx->xs[1 - cur], \
                     ctx->x, ctx->v, ctx->rho, ctx->nb_cnt, ctx->nb_quads, fut, ctx->n, near_on(ctx, t.count))
    if (last && fut) {
        PBF_LAUNCH_DELTA(true, true);
    } else if (last) {
        PBF_LAUNCH_DELTA(true, false);
    } else if (fut) {
        PBF_LAUNCH_DELTA(false, true);
    } else {
        PBF_LAUNCH_DELTA(false, false);
    }
#undef PBF_LAUNCH_DELTA
    track_end(ctx);
    ctx->xs_cur = 1 - cur;
    return hipGetLastError();
}

hipError_t launch_viscosity(pbf_ctx* ctx) {
    track_begin(ctx, "applyViscosity");
    const BlockTable& t = table_for(ctx, F_VISC);
    PBF_LAUNCH_TILED(F_VISC, k_viscosity<BS, ctx->sc_d, t.d, ctx->xs[ctx->xs_cur], ctx->v, ctx->v_alt, ctx->omega,
                     ctx->nb_cnt, ctx->nb_quads, ctx->n, near_on(ctx, t.count));
    track_end(ctx);
    float4* tmp = ctx->v;
    ctx->v = ctx->v_alt;
    ctx->v_alt = tmp;
    return hipGetLastError();
}

hipError_t launch_vorticity(pbf_ctx* ctx) {
    track_begin(ctx, "applyVorticity");
    const BlockTable& t = table_for(ctx, F_VORT);
    PBF_LAUNCH_TILED(F_VORT, k_vorticity<BS, ctx->sc_d, t.d, ctx->xs[ctx->xs_cur], ctx->v, ctx->omega, ctx->nb_cnt,
                     ctx->nb_quads, ctx->n, near_on(ctx, t.count));
    track_end(ctx);
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_tare(uint32_t) {}

hipError_t launch_tare(pbf_ctx* ctx) {
    track_begin(ctx, "eventTare");
    PBF_KERNEL(k_tare, dim3(1), dim3(64), 0, ctx->stream, 0u);
    track_end(ctx);
    return hipGetLastError();
}

#undef PBF_LAUNCH_TILED
#undef COMMA

}  // namespace pbf
