// pbf_internal.h -- context layout and launcher declarations (not exported).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pbf/pbf.h"
#include "pbf_device.h"

namespace pbf {

// Onesweep radix sort geometry: 1024 threads, 4 keys per thread, 8-bit digits.
constexpr int kSortThreads = 1024;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 4;

// Device error words: [0] PBF_FLAG_* bits, [2..3] dropped-friend count (u64),
// [4] tiles whose neighbourhood did not fit in LDS (diagnostic).  The same
// allocation holds the event log after them (kLogIndex, kLogRing in
// pbf_device.h).
constexpr int kErrWords = 8;

// Per-particle kernels run over a block table: one uint4 {scene, first,
// end, 0} per workgroup, so a workgroup never straddles two scenes.  Tables
// exist for these workgroup sizes (index = log2(size) - 6).
constexpr int kNumBlockSizes = 5;  // 64, 128, 256, 512, 1024
constexpr int kPredictThreads = 1024;  // threads of a k_predict_hash workgroup
constexpr int kPredictItems = 4;      // particles per thread in k_predict_hash
constexpr int kPredictTile = kPredictThreads * kPredictItems;   // particles per block

struct BlockTable {
    uint4* d = nullptr;
    uint32_t count = 0;
};

// Kernel families whose workgroup size the autotuner chooses.
enum Family { F_FRIENDS = 0, F_LAMBDA, F_DELTA, F_VISC, F_VORT, F_COUNT };

struct SortBuffers {
    uint32_t* ws = nullptr;       // [hist P*256][ctr 64][status P*tiles*256]
    size_t ws_words = 0;
    uint32_t* hist = nullptr;
    uint32_t* ctr = nullptr;
    uint32_t* status = nullptr;
    uint32_t tiles = 0;
    uint32_t passes = 0;
    uint32_t bits = 0;
};

struct Tracker {
    std::string name;
    pbf_timing t{};
};

struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
    int tracker = -1;
};

// One instantiated substep graph per variant of the buffer rotation; the
// host applies the rotation the captured step performs when it replays it.
struct GraphEntry {
    uint32_t key = 0;
    uint32_t steps = 1;   // consecutive substeps the graph holds
    hipGraphExec_t exec = nullptr;
    float4 *post_x = nullptr, *post_x_alt = nullptr, *post_v = nullptr, *post_v_alt = nullptr;
    int post_xs_cur = 0;
    bool post_hold = false, post_ord = false;   // cells_hold_keys, order_sorted after them
};

// pbf_run replays its substeps in graphs of this many: the substeps of a
// group follow each other inside one graph, with no graph launch between.
constexpr uint32_t kRunGraphSteps = 4;

}  // namespace pbf

// hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, args...) on a
// context, with the current tracker's events when one is open.
#define PBF_KERNEL(KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                                                  \
    do {                                                                                                     \
        if (ctx->launch_start)                                                                               \
            hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, ctx->launch_start, ctx->launch_stop, 0u, \
                                  __VA_ARGS__);                                                              \
        else                                                                                                 \
            hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, __VA_ARGS__);                             \
    } while (0)

#define HIP_TRY(ctx, call)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return pbf::fail((ctx), PBF_E_HIP, "%s: %s", #call, hipGetErrorString(e_));    \
    } while (0)

struct pbf_ctx {
    std::vector<pbf_params> P;        // per scene
    std::vector<pbf_options> opt;     // per scene
    std::vector<pbf::SceneConsts> sc; // per scene (host copy)
    pbf::SceneConsts* sc_d = nullptr;
    uint32_t S = 0;                   // scenes
    uint32_t n = 0;                   // particles, all scenes
    uint32_t iters = 0;
    uint32_t total_cells = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    pbf::BlockTable bt[pbf::kNumBlockSizes];
    pbf::BlockTable bt_predict;       // kPredictTile particles per block
    int bs_log2[pbf::F_COUNT] = {8, 8, 8, 8, 8};  // chosen workgroup sizes
    bool friends_staged = false;      // friends build: LDS-staged candidates or global reads
    // lanes per particle of the global-read friends build: 1, 2, 4, 8 or 16,
    // or 0 to choose from the particle count (friends_lanes_for)
    uint32_t friends_lanes = 0;
    // tiles per workgroup of each family: 1 = one workgroup per tile, k > 1 =
    // workgroups striding over k tiles each, 0 = a persistent grid of
    // resident[f][size] workgroups (the device's capacity for that kernel)
    uint32_t tiles_per_wg[pbf::F_COUNT] = {1, 1, 1, 1, 1};
    uint32_t resident[pbf::F_COUNT][4] = {};
    int cu_count = 0;

    // particle arrays (current cell-sorted order)
    float4* x = nullptr;        // positions, w = |v|
    float4* x_alt = nullptr;    // gather target
    float4* xs[2] = {nullptr, nullptr};  // predicted ping/pong
    int xs_cur = 0;             // index of the buffer holding the latest x*
    float4* v = nullptr;
    float4* v_alt = nullptr;
    float4* omega = nullptr;
    float* rho = nullptr;
    float* lambda = nullptr;

    // hashing / sorting
    uint32_t* keys_in = nullptr;     // unsorted keys of this step
    uint32_t* keys_tmp = nullptr;    // sort ping-pong
    uint32_t* keys_sorted = nullptr; // persistent sorted keys
    uint32_t* vals_a = nullptr;
    uint32_t* vals_b = nullptr;
    uint32_t* perm = nullptr;
    pbf::SortBuffers sort;

    // grid and friends
    uint32_t* cells = nullptr;       // uint2[total_cells]
    uint32_t* nb_cnt = nullptr;      // packed 5x6-bit circle counts
    uint32_t* nb_list = nullptr;     // friends, [n/64][5 circles][kSeg][64]
    uint4* nb_quads = nullptr;       // processed friends, sweep quads

    uint32_t* err = nullptr;     // error words, then the event log (kErrAllocWords)
    uint32_t log_read = 0;       // log write index up to which the host has read

    // collision quads (pbf_set_collision_quads), per scene on the host; one
    // device buffer holds every scene's QuadConsts and mask bits, and the
    // SceneConsts point into it
    std::vector<std::vector<pbf::QuadConsts>> quads;
    std::vector<std::vector<uint32_t>> mask_bits;
    void* geom_d = nullptr;

    uint32_t pause_now = 0;     // pause flag of the step being enqueued
    bool have_state = false;
    bool cells_hold_keys = false;  // cells hold the entries of keys_sorted (reset them first)
    bool order_sorted = false;     // particle order == keys_sorted order (pause may skip the sort)
    bool cells_fresh = false;      // this step's cells were built from its own predicted positions
    float4* x_base = nullptr;      // buffer identities for the graph keys
    float4* v_base = nullptr;
    bool profiling = false;
    bool use_graph = false;
    std::vector<pbf::GraphEntry> graphs;
    std::vector<pbf::Tracker> trackers;
    std::vector<pbf::EventPair> events;  // recorded (start, stop) pairs
    size_t events_used = 0;
    // the pair of the tracked launch being enqueued (null outside one)
    hipEvent_t launch_start = nullptr, launch_stop = nullptr;
    std::string last_error;
};

namespace pbf {
// Launchers (pbf_forward.hip, pbf_sort.hip).  All enqueue on ctx->stream.
hipError_t launch_clear(pbf_ctx* ctx);
hipError_t launch_predict_hash(pbf_ctx* ctx, bool reset_old);
hipError_t launch_predict_only(pbf_ctx* ctx, uint32_t pause);
hipError_t launch_sort(pbf_ctx* ctx);
hipError_t launch_gather(pbf_ctx* ctx);
// lambda0: the friends build also runs the substep's first computeScaling
hipError_t launch_friends(pbf_ctx* ctx, bool lambda0 = false);
// Lanes per particle the friends build runs with (ctx->friends_lanes, or
// when 0 a choice from the particle count and the CU count).
uint32_t friends_lanes_for(const pbf_ctx* ctx);
// fills ctx->resident from the kernels' occupancy on ctx->cu_count CUs
hipError_t query_residency(pbf_ctx* ctx);
// aux / fut: per-particle records for the adjoint (nullptr: plain step)
hipError_t launch_lambda(pbf_ctx* ctx, int iter, float4* aux = nullptr);
hipError_t launch_delta(pbf_ctx* ctx, int iter, bool last, float4* fut = nullptr);
hipError_t launch_viscosity(pbf_ctx* ctx);
hipError_t launch_vorticity(pbf_ctx* ctx);
// profiling only: an empty kernel under its own tracker ("eventTare"), the
// fixed part of a tracked launch's duration
hipError_t launch_tare(pbf_ctx* ctx);

// Records a formatted error on ctx (or for the next pbf_create when ctx is
// NULL) and returns code.
int fail(pbf_ctx* ctx, int code, const char* fmt, ...);

// Checks one scene's parameters (PBF_E_INVALID with a message on ctx).
int validate(const pbf_params* p, pbf_ctx* ctx, uint32_t scene);
// The device constants of a scene whose cells start at global cell id
// cell_base and whose particles start at index first (no collision quads).
SceneConsts make_consts(const pbf_params& P, const pbf_options& o, uint32_t cell_base, uint32_t first);
// Enqueues one substep (Simulation::Step) on ctx->stream.
int enqueue_step(pbf_ctx* ctx, uint32_t pause);

// Profiling brackets (no-ops unless ctx->profiling) around one kernel
// launch, which PBF_KERNEL makes: the tracker's events are stamped by the
// kernel itself at its start and end (hipExtLaunchKernelGGL), so a duration
// excludes the dispatch of the launch, as a profiler's kernel trace does.
void track_begin(pbf_ctx* ctx, const char* name, int index = -1);
void track_end(pbf_ctx* ctx);
void collect_timings(pbf_ctx* ctx);

inline const BlockTable& table_for(const pbf_ctx* ctx, Family f) { return ctx->bt[ctx->bs_log2[f] - 6]; }
}  // namespace pbf

#define HIP_TRY(ctx, call)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return pbf::fail((ctx), PBF_E_HIP, "%s: %s", #call, hipGetErrorString(e_));    \
    } while (0)
