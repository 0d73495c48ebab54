// pbf_sort.hip -- onesweep LSD radix sort of the cell keys (gfx950, wave64).
//
// Replaces the reference's 6-pass x 5-kernel radix sort (radixsort.cl:61-308
// driven by src/Simulation.cpp:466-520, with a host finish() per pass).  The
// result is the same unique stable order: keys_sorted ascending, perm[i] =
// pre-sort index of the particle that lands at i (radixsort.cl:19,204-209).
//
// Design (one kernel launch per 8-bit digit pass, no host syncs):
//  * the digit histograms of all passes are accumulated by the predict/hash
//    kernel (pbf_forward.hip), so each pass knows every digit's global base;
//  * each workgroup claims a tile of 4096 keys in launch order through an
//    atomic counter, ranks its keys stably with wave64 ballot matching
//    (8 ballots per key give the lanes holding the same digit), and learns the
//    keys of the same digit in earlier tiles by decoupled look-back over a
//    per-tile status word (aggregate / inclusive-prefix flag + 30-bit count);
//  * every wait in the look-back is bounded: on timeout the kernel sets
//    PBF_FLAG_SORT_TIMEOUT and continues, so no launch can hang.
#include "pbf_internal.h"

namespace pbf {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPre = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__global__ __launch_bounds__(kSortThreads) void k_onesweep(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, uint32_t n,
                                                           uint32_t shift, const uint32_t* __restrict__ hist,
                                                           uint32_t* status, uint32_t* tile_ctr,
                                                           uint32_t* err) {
    constexpr int kWaves = kSortThreads / 64;
    static_assert(kSortThreads >= kRadix && kRadix % 64 == 0, "one thread per digit in the first waves");
    __shared__ uint32_t s_cnt[kWaves][kRadix];
    __shared__ uint32_t s_off[kRadix];
    __shared__ uint32_t s_wsum[kRadix / 64];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const int lane = tid & 63;

    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (int q = tid; q < kWaves * kRadix; q += kSortThreads) (&s_cnt[0][0])[q] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;

    // Wave w owns keys [base, base + 64*kSortItems) of the tile; item k of
    // lane l is key base + 64k + l, so (k, lane) order is input order.
    const uint32_t base = tile * (uint32_t)kSortTile + (uint32_t)w * (64u * kSortItems);
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = base + 64u * k + lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0u;
        val[k] = vin ? (valid ? vin[idx] : 0u) : idx;
    }

    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = base + 64u * k + lane;
        const bool valid = idx < n;
        const uint32_t d = (key[k] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : lane;
        const uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t old = 0;
        if (valid && lane == leader) old = atomicAdd(&s_cnt[w][d], cnt);
        old = __shfl(old, leader, 64);
        rank[k] = old + (uint32_t)__popcll(peers & lt_mask);
    }
    __syncthreads();

    // The first kRadix threads, one per digit: wave-exclusive offsets, the
    // tile's count, the global base and the look-back.
    const uint32_t d = (uint32_t)tid;
    const bool digit = tid < kRadix;
    uint32_t tot = 0, hv = 0, incl = 0;
    uint32_t* my = status + (size_t)tile * kRadix + (digit ? d : 0u);
    if (digit) {
#pragma unroll
        for (int q = 0; q < kWaves; ++q) {
            const uint32_t cq = s_cnt[q][d];
            s_cnt[q][d] = tot;
            tot += cq;
        }
        __hip_atomic_store(my, (tile == 0 ? kFlagPre : kFlagAgg) | tot, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // global base of digit d: exclusive scan of this pass's histogram
        hv = hist[d];
        incl = wave_inclusive_scan(hv, lane);
        if (lane == 63) s_wsum[w] = incl;
    }
    __syncthreads();
    if (digit) {
        uint32_t gbase = incl - hv;
        for (int q = 0; q < w; ++q) gbase += s_wsum[q];

        uint32_t excl = 0;
        if (tile > 0) {
            int t = (int)tile - 1;
            uint32_t spins = 0;
            while (t >= 0) {
                const uint32_t s = __hip_atomic_load(status + (size_t)t * kRadix + d, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t flag = s & ~kValMask;
                if (flag == 0u) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err, (uint32_t)PBF_FLAG_SORT_TIMEOUT);
                        log_event(err, PBF_LOG_SORT_TIMEOUT, tile, d, shift);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += s & kValMask;
                if (flag == kFlagPre) break;
                --t;
            }
            __hip_atomic_store(my, kFlagPre | (excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_off[d] = gbase + excl;
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = base + 64u * k + lane;
        if (idx < n) {
            const uint32_t dg = (key[k] >> shift) & (kRadix - 1);
            const uint32_t dst = s_off[dg] + s_cnt[w][dg] + rank[k];
            if (dst < n) {
                kout[dst] = key[k];
                vout[dst] = val[k];
            } else {
                atomicOr(err, (uint32_t)PBF_FLAG_INTERNAL);
                log_event(err, PBF_LOG_INTERNAL, 3u, idx, dst);
            }
        }
    }
}

__global__ void k_clear_words(uint32_t* __restrict__ p, size_t words) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0u;
}

}  // namespace

// Zeroes the sort workspace; run once at create (k_gather_cells zeroes it
// after every sort).
hipError_t launch_clear(pbf_ctx* ctx) {
    const size_t words = ctx->sort.ws_words;
    const unsigned blocks = (unsigned)((words + 255) / 256);
    hipLaunchKernelGGL(k_clear_words, dim3(blocks), dim3(256), 0, ctx->stream, ctx->sort.ws, words);
    return hipGetLastError();
}

hipError_t launch_sort(pbf_ctx* ctx) {
    const uint32_t P = ctx->sort.passes;
    const uint32_t n = ctx->n;
    uint32_t* B = ctx->keys_in;
    uint32_t* C = ctx->keys_tmp;
    uint32_t* S = ctx->keys_sorted;
    for (uint32_t p = 0; p < P; ++p) {
        const bool last = (p == P - 1);
        const uint32_t* kin = (p == 0) ? B : ((p & 1) ? C : B);
        uint32_t* kout = last ? S : ((p & 1) ? B : C);
        const uint32_t* vin = (p == 0) ? nullptr : ((p & 1) ? ctx->vals_a : ctx->vals_b);
        uint32_t* vout = last ? ctx->perm : ((p & 1) ? ctx->vals_b : ctx->vals_a);
        track_begin(ctx, "onesweep", (int)p);
        PBF_KERNEL(k_onesweep, dim3(ctx->sort.tiles), dim3(kSortThreads), 0, ctx->stream, kin, vin, kout,
                           vout, n, (uint32_t)(kRadixBits * p), ctx->sort.hist + (size_t)p * kRadix,
                           ctx->sort.status + (size_t)p * ctx->sort.tiles * kRadix, ctx->sort.ctr + p, ctx->err);
        track_end(ctx);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pbf
