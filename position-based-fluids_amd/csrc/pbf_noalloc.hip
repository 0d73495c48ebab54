// pbf_noalloc.hip -- pbf_run_noalloc: steps of one scene with no context and
// no device allocation.
//
// Every buffer the substep needs beyond the caller's inputs and outputs
// (working positions and velocities, predicted positions, keys, the sort
// workspace, cells, friend lists, densities, lambda, vorticity, the block
// tables, the scene constants and the error words) is carved out of one
// file-scope device array, `scratch`, sized for kMaxParticles particles and
// kMaxCells hash buckets.  Each device holds its own copy of the array for
// the life of the process; its address on a device is looked up once
// (hipGetSymbolAddress) and kept.  The step itself is the context path's
// enqueue_step over a view whose pointers lie in the array, so the kernels
// and their results are those of pbf_run.
//
// Calls on one device share the array: two calls whose work overlaps on that
// device (different streams, or different host threads) race.  Everything a
// call needs to start from a known value is set by kernels on its stream
// before its first step.
#include <mutex>

#include "pbf_internal.h"

namespace pbf {
namespace {

constexpr uint32_t kMaxParticles = 16384;
constexpr uint32_t kMaxCells = 1u << 16;
constexpr size_t kAlign = 256;

constexpr size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

constexpr uint32_t passes_for(uint32_t cells) {
    uint32_t bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)cells) ++bits;
    if (bits == 0) bits = 1;
    return (bits + kRadixBits - 1) / kRadixBits;
}

constexpr size_t sort_words(uint32_t n, uint32_t cells) {
    const size_t tiles = n ? (n + kSortTile - 1) / kSortTile : 1;
    return (size_t)kMaxPasses * kRadix + 64 + (size_t)passes_for(cells) * tiles * kRadix;
}

// Bytes of the layout below for n particles and `cells` buckets.
constexpr size_t scratch_bytes(uint32_t n, uint32_t cells) {
    const size_t v4 = align_up(16 * (size_t)n), s1 = align_up(4 * (size_t)n);
    size_t b = 7 * v4 + 9 * s1;   // x, x_alt, xs[2], v, v_alt, omega; rho, lambda, 3 keys, 2 vals, perm, counts
    b += align_up(4 * friend_slots(n));
    b += align_up(16 * ((size_t)(n + 63) / 64) * 64 * kQuads);
    b += align_up(8 * (size_t)cells);
    b += align_up(4 * sort_words(n, cells));
    b += align_up(4 * (size_t)kErrAllocWords);
    b += align_up(sizeof(SceneConsts));
    for (int k = 0; k < kNumBlockSizes; ++k) b += align_up(16 * (((size_t)n + (64u << k) - 1) / (64u << k)));
    b += align_up(16 * (((size_t)n + kPredictTile - 1) / kPredictTile));
    return b;
}

constexpr size_t kScratchBytes = scratch_bytes(kMaxParticles, kMaxCells);

}  // namespace
}  // namespace pbf

// The scratch of pbf_run_noalloc (about 36 MB on each device it runs on).  In
// host code its name is only a stand-in for the symbol: kernels and copies
// take the device address that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[pbf::kScratchBytes];

namespace pbf {
namespace {

// Per-device address of `scratch`, looked up once per device with that
// device current.
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= 64) return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, PBF_E_HIP, "pbf_run_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return PBF_OK;
}

__global__ void k_put_consts(SceneConsts* __restrict__ dst, SceneConsts c) { *dst = c; }

// Block table of one scene: {0, b per, min(n, (b + 1) per), 0}.
__global__ void k_fill_blocks(uint4* __restrict__ t, uint32_t count, uint32_t n, uint32_t per) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < count) t[b] = make_uint4(0u, b * per, min(n, (b + 1) * per), 0u);
}

__global__ void k_fill_words(uint32_t* __restrict__ p, size_t words, uint32_t value) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) p[w] = value;
}

unsigned grid_for(size_t items) { return (unsigned)std::min<size_t>((items + 255) / 256, 4096); }

}  // namespace
}  // namespace pbf

extern "C" int pbf_run_noalloc(const pbf_params* params, const pbf_options* opt, const float* x4, const float* v4,
                               uint32_t n, uint32_t steps, float* x4_out, float* v4_out, void* stream) {
    using namespace pbf;
    int rc = validate(params, nullptr, 0);
    if (rc) return rc;
    if (!x4 || !x4_out || !v4_out) return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: x4, x4_out and v4_out are required");
    if (n != params->particleCount)
        return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: n = %u but particleCount = %u", n, params->particleCount);
    if (n == 0 || n > kMaxParticles)
        return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: %u particles; its scratch holds 1..%u", n, kMaxParticles);
    if (params->gridBufSize > kMaxCells)
        return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: gridBufSize %u; its scratch holds at most %u buckets",
                    params->gridBufSize, kMaxCells);
    pbf_options o;
    if (opt)
        o = *opt;
    else
        pbf_default_options(&o);

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, PBF_E_HIP, "pbf_run_noalloc: hipGetDevice failed");
    unsigned char* base = nullptr;
    if ((rc = scratch_address(dev, &base)) != PBF_OK) return rc;

    // a view over the scratch: the context path's buffers, carved in the
    // order scratch_bytes counts them (its host-side vectors stay empty)
    pbf_ctx c;
    c.S = 1;
    c.n = n;
    c.iters = params->simIterations;
    c.total_cells = params->gridBufSize;
    c.device = dev;
    c.stream = static_cast<hipStream_t>(stream);
    // the CU count (a device attribute, no allocation) sizes launch choices
    if (hipDeviceGetAttribute(&c.cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(nullptr, PBF_E_HIP, "pbf_run_noalloc: hipDeviceGetAttribute failed");
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += align_up(bytes);
        return p;
    };
    const size_t v4b = 16 * (size_t)n, s1 = 4 * (size_t)n;
    c.x = reinterpret_cast<float4*>(take(v4b));
    c.x_alt = reinterpret_cast<float4*>(take(v4b));
    c.xs[0] = reinterpret_cast<float4*>(take(v4b));
    c.xs[1] = reinterpret_cast<float4*>(take(v4b));
    c.v = reinterpret_cast<float4*>(take(v4b));
    c.v_alt = reinterpret_cast<float4*>(take(v4b));
    c.omega = reinterpret_cast<float4*>(take(v4b));
    c.rho = reinterpret_cast<float*>(take(s1));
    c.lambda = reinterpret_cast<float*>(take(s1));
    c.keys_in = reinterpret_cast<uint32_t*>(take(s1));
    c.keys_tmp = reinterpret_cast<uint32_t*>(take(s1));
    c.keys_sorted = reinterpret_cast<uint32_t*>(take(s1));
    c.vals_a = reinterpret_cast<uint32_t*>(take(s1));
    c.vals_b = reinterpret_cast<uint32_t*>(take(s1));
    c.perm = reinterpret_cast<uint32_t*>(take(s1));
    c.nb_cnt = reinterpret_cast<uint32_t*>(take(s1));
    c.nb_list = reinterpret_cast<uint32_t*>(take(4 * friend_slots(n)));
    c.nb_quads = reinterpret_cast<uint4*>(take(16 * ((size_t)(n + 63) / 64) * 64 * kQuads));
    c.cells = reinterpret_cast<uint32_t*>(take(8 * (size_t)c.total_cells));
    c.sort.passes = passes_for(c.total_cells);
    c.sort.tiles = (n + kSortTile - 1) / kSortTile;
    c.sort.ws_words = sort_words(n, c.total_cells);
    c.sort.ws = reinterpret_cast<uint32_t*>(take(4 * c.sort.ws_words));
    c.sort.hist = c.sort.ws;
    c.sort.ctr = c.sort.ws + (size_t)kMaxPasses * kRadix;
    c.sort.status = c.sort.ctr + 64;
    c.err = reinterpret_cast<uint32_t*>(take(4 * (size_t)kErrAllocWords));
    c.sc_d = reinterpret_cast<SceneConsts*>(take(sizeof(SceneConsts)));
    for (int k = 0; k < kNumBlockSizes; ++k) {
        const uint32_t per = 64u << k;
        c.bt[k].count = (n + per - 1) / per;
        c.bt[k].d = reinterpret_cast<uint4*>(take(16 * (size_t)c.bt[k].count));
    }
    c.bt_predict.count = (n + kPredictTile - 1) / kPredictTile;
    c.bt_predict.d = reinterpret_cast<uint4*>(take(16 * (size_t)c.bt_predict.count));
    if (off > kScratchBytes)   // the layout above and scratch_bytes disagree
        return fail(nullptr, PBF_E_INVALID, "pbf_run_noalloc: layout of %zu bytes exceeds the scratch (%zu)", off,
                    kScratchBytes);

    // initial values, all set on the stream by kernels and copies
    hipLaunchKernelGGL(k_put_consts, dim3(1), dim3(1), 0, c.stream, c.sc_d, make_consts(*params, o, 0u, 0u));
    for (int k = 0; k < kNumBlockSizes; ++k)
        hipLaunchKernelGGL(k_fill_blocks, dim3((c.bt[k].count + 255) / 256), dim3(256), 0, c.stream, c.bt[k].d,
                           c.bt[k].count, n, 64u << k);
    hipLaunchKernelGGL(k_fill_blocks, dim3((c.bt_predict.count + 255) / 256), dim3(256), 0, c.stream, c.bt_predict.d,
                       c.bt_predict.count, n, (uint32_t)kPredictTile);
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for(2 * (size_t)c.total_cells)), dim3(256), 0, c.stream, c.cells,
                       2 * (size_t)c.total_cells, 0xFFFFFFFFu);
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for(n)), dim3(256), 0, c.stream, c.nb_cnt, (size_t)n, 0u);
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for(kErrAllocWords)), dim3(256), 0, c.stream, c.err,
                       (size_t)kErrAllocWords, 0u);
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for(c.sort.ws_words)), dim3(256), 0, c.stream, c.sort.ws,
                       c.sort.ws_words, 0u);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(c.x, x4, v4b, hipMemcpyDeviceToDevice, c.stream);
    if (e == hipSuccess) {
        if (v4)
            e = hipMemcpyAsync(c.v, v4, v4b, hipMemcpyDeviceToDevice, c.stream);
        else
            hipLaunchKernelGGL(k_fill_words, dim3(grid_for(4 * (size_t)n)), dim3(256), 0, c.stream,
                               reinterpret_cast<uint32_t*>(c.v), 4 * (size_t)n, 0u);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, PBF_E_HIP, "pbf_run_noalloc: set-up: %s", hipGetErrorString(e));
    c.have_state = true;
    c.x_base = c.x;
    c.v_base = c.v;
    for (uint32_t s = 0; s < steps; ++s) {
        if ((rc = enqueue_step(&c, 0u)) != PBF_OK)
            return fail(nullptr, rc, "pbf_run_noalloc: step %u: %s", s, c.last_error.c_str());
    }
    e = hipMemcpyAsync(x4_out, c.x, v4b, hipMemcpyDeviceToDevice, c.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(v4_out, c.v, v4b, hipMemcpyDeviceToDevice, c.stream);
    if (e != hipSuccess) return fail(nullptr, PBF_E_HIP, "pbf_run_noalloc: result copy: %s", hipGetErrorString(e));
    return PBF_OK;
}
