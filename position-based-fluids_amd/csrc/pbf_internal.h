// pbf_internal.h -- context layout and launcher declarations (not exported).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pbf/pbf.h"
#include "pbf_device.h"

namespace pbf {

// Onesweep radix sort geometry: 256 threads, 8 keys per thread, 8-bit digits.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 4;

// Device error words: [0] PBF_FLAG_* bits, [1..2] dropped-friend count (u64).
constexpr int kErrWords = 4;

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

}  // namespace pbf

struct pbf_ctx {
    pbf_params P{};
    pbf_options opt{};
    pbf::StepConsts c{};
    uint32_t n = 0;
    int device = 0;
    hipStream_t stream = nullptr;

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
    uint32_t* cells = nullptr;       // uint2[gbs]
    uint32_t* nb_cnt = nullptr;      // packed 5x6-bit circle counts
    uint32_t* nb_idx = nullptr;      // [5][50][n]

    uint32_t* err = nullptr;

    uint32_t pause_now = 0;     // pause flag of the step being enqueued
    bool have_state = false;
    bool have_sorted = false;
    bool profiling = false;
    std::vector<pbf::Tracker> trackers;
    std::vector<pbf::EventPair> events;  // recorded (start, stop) pairs
    size_t events_used = 0;
    std::string last_error;
};

namespace pbf {
// Launchers (pbf_forward.hip, pbf_sort.hip).  All enqueue on ctx->stream.
hipError_t launch_clear(pbf_ctx* ctx);
hipError_t launch_predict_hash(pbf_ctx* ctx, bool reset_old_cells);
hipError_t launch_predict_only(pbf_ctx* ctx, uint32_t pause);
hipError_t launch_sort(pbf_ctx* ctx);
hipError_t launch_gather(pbf_ctx* ctx);
hipError_t launch_cells(pbf_ctx* ctx);
hipError_t launch_friends(pbf_ctx* ctx);
hipError_t launch_lambda(pbf_ctx* ctx, int iter);
hipError_t launch_delta(pbf_ctx* ctx, int iter, bool last);
hipError_t launch_viscosity(pbf_ctx* ctx);
hipError_t launch_vorticity(pbf_ctx* ctx);

// Profiling brackets (no-ops unless ctx->profiling).
void track_begin(pbf_ctx* ctx, const char* name, int index = -1);
void track_end(pbf_ctx* ctx);
void collect_timings(pbf_ctx* ctx);
}  // namespace pbf
