// pbf_capi.hip -- the exported C ABI (include/pbf/pbf.h): context set-up,
// the substep driver and the host-side getters.
//
// pbf_step enqueues the whole substep on the context's stream with no host
// synchronisation (the reference blocks six times per step inside its sort,
// src/Simulation.cpp:488, and once at GL unlock, :551).  A context holds one
// scene (pbf_create) or a batch of independent scenes (pbf_create_batch)
// that every kernel covers in one launch through per-scene offset tables.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdarg>
#include <new>

#include "pbf_internal.h"

namespace {
// message of the last failure with no context (pbf_create, pbf_run_noalloc);
// a fixed buffer, so that recording it allocates nothing
char g_create_error[512];
}  // namespace

namespace pbf {
int fail(pbf_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->last_error = buf;
    else
        snprintf(g_create_error, sizeof g_create_error, "%s", buf);
    return code;
}
}  // namespace pbf

namespace {

using pbf::fail;

// Simulation.cpp:167,179-180: the reference streams the double expressions
// with std::showpoint at the default precision (6 significant digits) and a
// trailing 'f', so its kernels see the float nearest that decimal.
void kernel_factors(float h, float* poly6, float* spiky) {
    char buf[64];
    const double hd = (double)h;
    const double p6 = 315.0f / (64.0f * M_PI * pow(hd, 9));
    const double sp = 45.0f / (M_PI * pow(hd, 6));
    snprintf(buf, sizeof buf, "%#g", p6);
    *poly6 = strtof(buf, nullptr);
    snprintf(buf, sizeof buf, "%#g", sp);
    *spiky = strtof(buf, nullptr);
}

// The reference's friend circle of an in-range candidate
// (build_friends_list.cl:75-76), evaluated in binary32 as the kernels did.
int circle_of(float r2, float min_r, float span) {
    float a = (sqrtf(r2) - min_r) / span;
    a = a > 0.0f ? a : 0.0f;
    const int c = (int)(a * a * a * 5.0f);
    return c < 4 ? c : 4;
}

// Smallest non-negative binary32 r^2 whose circle is >= k (k = 1..4), by
// bisection over the bit patterns (ordered like the values for r^2 >= 0); a
// circle no in-range r^2 reaches gets +inf.
void circle_thresholds(float h2, float min_r, float span, float* t) {
    uint32_t hb;
    memcpy(&hb, &h2, 4);
    for (int k = 1; k <= 4; ++k) {
        uint32_t lo = 0, hi = hb;   // circle(lo) < k is not assumed; search [0, hb]
        float hv;
        memcpy(&hv, &hi, 4);
        if (circle_of(hv, min_r, span) < k) {
            t[k - 1] = INFINITY;
            continue;
        }
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float mv;
            memcpy(&mv, &mid, 4);
            if (circle_of(mv, min_r, span) >= k) hi = mid; else lo = mid + 1;
        }
        memcpy(&t[k - 1], &lo, 4);
    }
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
}

void drop_graphs(pbf_ctx* ctx) {
    for (pbf::GraphEntry& g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
}

void free_all(pbf_ctx* ctx) {
    drop_graphs(ctx);
    void* ptrs[] = {ctx->x, ctx->x_alt, ctx->xs[0], ctx->xs[1], ctx->v, ctx->v_alt, ctx->omega, ctx->rho,
                    ctx->lambda, ctx->keys_in, ctx->keys_tmp, ctx->keys_sorted, ctx->vals_a, ctx->vals_b,
                    ctx->perm, ctx->sort.ws, ctx->cells, ctx->nb_cnt, ctx->nb_list, ctx->nb_quads, ctx->err, ctx->sc_d,
                    ctx->bt_predict.d, ctx->geom_d};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& t : ctx->bt)
        if (t.d) (void)hipFree(t.d);
    for (pbf::EventPair& e : ctx->events) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    ctx->events.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

}  // namespace

namespace pbf {

int validate(const pbf_params* p, pbf_ctx* ctx, uint32_t scene) {
    if (!p) return fail(ctx, PBF_E_INVALID, "params is NULL");
    if (!(p->h >= 0x1p-20f) || !isfinite(p->h))
        return fail(ctx, PBF_E_INVALID, "scene %u: h must be finite and at least 2^-20", scene);
    if (!(p->timeStep > 0.0f)) return fail(ctx, PBF_E_INVALID, "scene %u: timeStep must be positive", scene);
    if (p->simIterations < 1 || p->simIterations > 64)
        return fail(ctx, PBF_E_INVALID, "scene %u: simIterations must be in [1, 64] (got %u)", scene, p->simIterations);
    if (p->gridBufSize < 1 || p->gridBufSize > (1u << 30))
        return fail(ctx, PBF_E_INVALID, "scene %u: gridBufSize must be in [1, 2^30] (got %u)", scene, p->gridBufSize);
    if (p->friendsCircles != (uint32_t)pbf::kCircles || p->particlesPerCircle != (uint32_t)pbf::kPerCircle)
        return fail(ctx, PBF_E_INVALID,
                    "scene %u: friendsCircles/particlesPerCircle must be 5/50 (call pbf_params_finalize)", scene);
    if (!(p->restDensity != 0.0f)) return fail(ctx, PBF_E_INVALID, "scene %u: restDensity must be nonzero", scene);
    return PBF_OK;
}

pbf::SceneConsts make_consts(const pbf_params& P, const pbf_options& o, uint32_t cell_base, uint32_t first) {
    pbf::SceneConsts c{};
    c.h = P.h;
    c.h2 = P.h_2;
    c.min_r = 0.3f * P.h;            // build_friends_list.cl:32
    c.span = P.h - c.min_r;          // build_friends_list.cl:75
    c.dt = P.timeStep;
    if (o.use_gravity_vector) {
        c.g0 = o.gravity[0];
        c.g1 = o.gravity[1];
        c.g2 = o.gravity[2];
    } else {
        c.g0 = 0.0f;                 // predict_positions.cl:13
        c.g1 = -P.garvity;
        c.g2 = 0.0f;
    }
    c.rest_density = P.restDensity;
    c.epsilon = P.epsilon;
    kernel_factors(P.h, &c.poly6, &c.spiky);
    c.st_k = P.surfaceTenstionK;
    const float dqh = P.surfaceTenstionDist * P.h;   // compute_delta.cl:81-82
    const float q2 = dqh * dqh;
    const float tq = P.h_2 - q2;
    c.inv_poly6_q = 1.0f / (tq * tq * tq);   // the ratio's divisor as one reciprocal per step
    c.viscosity = P.viscosityFactor;
    c.vorticity = P.vorticityFactor;
    c.lo0 = P.xMin; c.lo1 = P.yMin; c.lo2 = P.zMin;
    c.hi0 = P.xMax; c.hi1 = P.yMax; c.hi2 = P.zMax;
    c.walls = o.walls ? 1u : 0u;
    c.gbs = P.gridBufSize;
    c.gbs_mask = (P.gridBufSize > 1 && (P.gridBufSize & (P.gridBufSize - 1)) == 0) ? P.gridBufSize - 1 : 0u;
    c.cell_base = cell_base;
    c.first = first;
    c.count = P.particleCount;
    circle_thresholds(c.h2, c.min_r, c.span, c.circ_t);
    return c;
}

}  // namespace pbf

namespace {

using pbf::make_consts;
using pbf::validate;

// Block table: one {scene, first, end, 0} per workgroup of `per` particles.
std::vector<uint4> block_table(const pbf_ctx* ctx, uint32_t per) {
    std::vector<uint4> t;
    for (uint32_t s = 0; s < ctx->S; ++s) {
        const uint32_t first = ctx->sc[s].first, end = first + ctx->sc[s].count;
        for (uint32_t b = first; b < end; b += per) t.push_back(make_uint4(s, b, std::min(end, b + per), 0u));
    }
    return t;
}

int create_impl(const pbf_params* params, const pbf_options* opts, uint32_t S, pbf_ctx** out) {
    if (!out) return fail(nullptr, PBF_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!params || S == 0) return fail(nullptr, PBF_E_INVALID, "no scenes");
    uint64_t n_total = 0, cells_total = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const int rc = validate(&params[s], nullptr, s);
        if (rc) return rc;
        if (params[s].simIterations != params[0].simIterations)
            return fail(nullptr, PBF_E_INVALID, "scene %u: all scenes of a batch share simIterations", s);
        n_total += params[s].particleCount;
        cells_total += params[s].gridBufSize;
    }
    if (n_total >= (1u << 29)) return fail(nullptr, PBF_E_INVALID, "%llu particles >= 2^29", (unsigned long long)n_total);
    if (cells_total > (1u << 30)) return fail(nullptr, PBF_E_INVALID, "total gridBufSize > 2^30");

    pbf_ctx* ctx = new (std::nothrow) pbf_ctx();
    if (!ctx) return fail(nullptr, PBF_E_NOMEM, "host allocation failed");
    ctx->S = S;
    ctx->n = (uint32_t)n_total;
    ctx->total_cells = (uint32_t)cells_total;
    ctx->iters = params[0].simIterations;
    uint32_t first = 0, cell_base = 0;
    for (uint32_t s = 0; s < S; ++s) {
        ctx->P.push_back(params[s]);
        pbf_options o;
        if (opts)
            o = opts[s];
        else
            pbf_default_options(&o);
        ctx->opt.push_back(o);
        ctx->sc.push_back(make_consts(params[s], o, cell_base, first));
        first += params[s].particleCount;
        cell_base += params[s].gridBufSize;
    }

    auto bail = [&](int code) {
        snprintf(g_create_error, sizeof g_create_error, "%s", ctx->last_error.c_str());
        free_all(ctx);
        delete ctx;
        return code;
    };
#define CREATE_TRY(call)                                                                                    \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            const int code_ = e_ == hipErrorOutOfMemory ? PBF_E_NOMEM : PBF_E_HIP;                          \
            fail(ctx, code_, "%s: %s", #call, hipGetErrorString(e_));                                       \
            return bail(code_);                                                                             \
        }                                                                                                   \
    } while (0)

    CREATE_TRY(hipGetDevice(&ctx->device));
    CREATE_TRY(hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, ctx->device));
    CREATE_TRY(pbf::query_residency(ctx));
    CREATE_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const size_t n = ctx->n;
    CREATE_TRY(dalloc(&ctx->x, n));
    CREATE_TRY(dalloc(&ctx->x_alt, n));
    CREATE_TRY(dalloc(&ctx->xs[0], n));
    CREATE_TRY(dalloc(&ctx->xs[1], n));
    CREATE_TRY(dalloc(&ctx->v, n));
    CREATE_TRY(dalloc(&ctx->v_alt, n));
    CREATE_TRY(dalloc(&ctx->omega, n));
    CREATE_TRY(dalloc(&ctx->rho, n));
    CREATE_TRY(dalloc(&ctx->lambda, n));
    CREATE_TRY(dalloc(&ctx->keys_in, n));
    CREATE_TRY(dalloc(&ctx->keys_tmp, n));
    CREATE_TRY(dalloc(&ctx->keys_sorted, n));
    CREATE_TRY(dalloc(&ctx->vals_a, n));
    CREATE_TRY(dalloc(&ctx->vals_b, n));
    CREATE_TRY(dalloc(&ctx->perm, n));
    CREATE_TRY(dalloc(&ctx->cells, 2 * (size_t)ctx->total_cells));
    CREATE_TRY(dalloc(&ctx->nb_cnt, n));
    CREATE_TRY(dalloc(&ctx->nb_list, pbf::friend_slots(n)));
    CREATE_TRY(dalloc(&ctx->nb_quads, ((n + 63) / 64) * 64 * (size_t)pbf::kQuads));
    CREATE_TRY(dalloc(&ctx->err, pbf::kErrAllocWords));
    CREATE_TRY(dalloc(&ctx->sc_d, S));
    CREATE_TRY(hipMemcpy(ctx->sc_d, ctx->sc.data(), sizeof(pbf::SceneConsts) * S, hipMemcpyHostToDevice));

    for (int k = 0; k < pbf::kNumBlockSizes; ++k) {
        const std::vector<uint4> t = block_table(ctx, 64u << k);
        CREATE_TRY(dalloc(&ctx->bt[k].d, t.size()));
        if (!t.empty()) CREATE_TRY(hipMemcpy(ctx->bt[k].d, t.data(), sizeof(uint4) * t.size(), hipMemcpyHostToDevice));
        ctx->bt[k].count = (uint32_t)t.size();
    }
    {
        const std::vector<uint4> t = block_table(ctx, (uint32_t)pbf::kPredictTile);
        CREATE_TRY(dalloc(&ctx->bt_predict.d, t.size()));
        if (!t.empty())
            CREATE_TRY(hipMemcpy(ctx->bt_predict.d, t.data(), sizeof(uint4) * t.size(), hipMemcpyHostToDevice));
        ctx->bt_predict.count = (uint32_t)t.size();
    }

    // sort geometry: enough 8-bit passes for keys < total_cells
    uint32_t bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)ctx->total_cells) ++bits;
    if (bits == 0) bits = 1;
    ctx->sort.bits = bits;
    ctx->sort.passes = (bits + pbf::kRadixBits - 1) / pbf::kRadixBits;
    ctx->sort.tiles = (uint32_t)std::max<size_t>((n + pbf::kSortTile - 1) / pbf::kSortTile, 1);
    const size_t hist_words = (size_t)pbf::kMaxPasses * pbf::kRadix;
    const size_t ctr_words = 64;
    const size_t status_words = (size_t)ctx->sort.passes * ctx->sort.tiles * pbf::kRadix;
    ctx->sort.ws_words = hist_words + ctr_words + status_words;
    CREATE_TRY(dalloc(&ctx->sort.ws, ctx->sort.ws_words));
    ctx->sort.hist = ctx->sort.ws;
    ctx->sort.ctr = ctx->sort.ws + hist_words;
    ctx->sort.status = ctx->sort.ws + hist_words + ctr_words;

    CREATE_TRY(hipMemsetAsync(ctx->cells, 0xFF, 8 * (size_t)ctx->total_cells, ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->nb_cnt, 0, 4 * std::max<size_t>(n, 1), ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->v, 0, 16 * std::max<size_t>(n, 1), ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->err, 0, 4 * pbf::kErrAllocWords, ctx->stream));
    CREATE_TRY(pbf::launch_clear(ctx));   // the sort workspace starts zero
    CREATE_TRY(hipStreamSynchronize(ctx->stream));
#undef CREATE_TRY
    *out = ctx;
    return PBF_OK;
}

// The state of the buffer rotation after a step: what a captured graph
// leaves behind, applied on the host when the graph is replayed.
struct Rotation {
    float4 *x, *x_alt, *v, *v_alt;
    int xs_cur;
};

Rotation rotation_of(const pbf_ctx* ctx) { return {ctx->x, ctx->x_alt, ctx->v, ctx->v_alt, ctx->xs_cur}; }

}  // namespace

int pbf::enqueue_step(pbf_ctx* ctx, uint32_t pause) {
    ctx->pause_now = pause ? 1u : 0u;
    // Simulation.cpp:563-567: predict always; radixsort unless paused.  With
    // the particle order not sorted (no step yet, or new state) the step
    // sorts anyway.
    ctx->cells_fresh = !(pause && ctx->order_sorted);
    if (pause && ctx->order_sorted) {
        HIP_TRY(ctx, pbf::launch_predict_only(ctx, 1u));
    } else {
        // the sort workspace is zero between steps: cleared at create, then
        // by k_gather_cells after every sort
        HIP_TRY(ctx, pbf::launch_predict_hash(ctx, ctx->cells_hold_keys));
        HIP_TRY(ctx, pbf::launch_sort(ctx));
        HIP_TRY(ctx, pbf::launch_gather(ctx));   // + updateCells
        ctx->cells_hold_keys = true;
        ctx->order_sorted = true;
    }
    // the friends build runs the first computeScaling as well, each thread
    // from the list it has just written
    HIP_TRY(ctx, pbf::launch_friends(ctx, true));
    const int K = (int)ctx->iters;
    for (int it = 0; it < K; ++it) {
        if (it > 0) HIP_TRY(ctx, pbf::launch_lambda(ctx, it));
        HIP_TRY(ctx, pbf::launch_delta(ctx, it, it == K - 1));
    }
    HIP_TRY(ctx, pbf::launch_viscosity(ctx));
    HIP_TRY(ctx, pbf::launch_vorticity(ctx));
    if (ctx->profiling) HIP_TRY(ctx, pbf::launch_tare(ctx));
    return PBF_OK;
}

namespace {

using pbf::enqueue_step;

// Graph key: everything that changes the launch sequence or its arguments.
uint32_t graph_key(const pbf_ctx* ctx, uint32_t pause) {
    const bool sorts = !(pause && ctx->order_sorted);
    return (ctx->x == ctx->x_base ? 1u : 0u) | (ctx->v == ctx->v_base ? 2u : 0u) | (sorts ? 4u : 0u) |
           (ctx->cells_hold_keys ? 8u : 0u) | (pause ? 16u : 0u) | ((uint32_t)ctx->xs_cur << 5);
}

// Replays `steps` consecutive substeps as one graph, captured on first use
// for each variant (the key of the first substep; the later ones follow from
// it).
int step_graph(pbf_ctx* ctx, uint32_t pause, uint32_t steps) {
    const uint32_t key = graph_key(ctx, pause);
    for (pbf::GraphEntry& g : ctx->graphs) {
        if (g.key != key || g.steps != steps) continue;
        HIP_TRY(ctx, hipGraphLaunch(g.exec, ctx->stream));
        ctx->x = g.post_x; ctx->x_alt = g.post_x_alt; ctx->v = g.post_v; ctx->v_alt = g.post_v_alt;
        ctx->xs_cur = g.post_xs_cur;
        ctx->cells_hold_keys = g.post_hold;
        ctx->order_sorted = g.post_ord;
        return PBF_OK;
    }
    // capture this variant once
    const bool hold = ctx->cells_hold_keys, ord = ctx->order_sorted;
    const Rotation before = rotation_of(ctx);
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = PBF_OK;
    for (uint32_t s = 0; s < steps && rc == PBF_OK; ++s) rc = enqueue_step(ctx, pause);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != PBF_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(ctx, PBF_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    pbf::GraphEntry g;
    g.key = key;
    g.steps = steps;
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(ctx, PBF_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    g.post_x = ctx->x; g.post_x_alt = ctx->x_alt; g.post_v = ctx->v; g.post_v_alt = ctx->v_alt;
    g.post_xs_cur = ctx->xs_cur;
    g.post_hold = ctx->cells_hold_keys;
    g.post_ord = ctx->order_sorted;
    ctx->graphs.push_back(g);
    // the capture recorded the work without running it: rewind and launch
    ctx->x = before.x; ctx->x_alt = before.x_alt; ctx->v = before.v; ctx->v_alt = before.v_alt;
    ctx->xs_cur = before.xs_cur;
    ctx->cells_hold_keys = hold;
    ctx->order_sorted = ord;
    return step_graph(ctx, pause, steps);
}

// What every step checks first: PBF_OK to go on, or the code to return
// (*skip with PBF_OK when there is nothing to step).
int step_ready(pbf_ctx* ctx, bool* skip) {
    *skip = false;
    if (!ctx) return PBF_E_INVALID;
    if (!ctx->have_state) return fail(ctx, PBF_E_STATE, "pbf_step before pbf_set_state");
    if (ctx->n == 0) {
        *skip = true;
        return PBF_OK;
    }
    if (ctx->x_base == nullptr) {
        ctx->x_base = ctx->x;
        ctx->v_base = ctx->v;
    }
    return PBF_OK;
}

}  // namespace

namespace pbf {

void track_begin(pbf_ctx* ctx, const char* name, int index) {
    if (!ctx->profiling) return;
    std::string nm(name);
    if (index >= 0) nm += "_" + std::to_string(index);
    int tid = -1;
    for (size_t q = 0; q < ctx->trackers.size(); ++q)
        if (ctx->trackers[q].name == nm) { tid = (int)q; break; }
    if (tid < 0) {
        Tracker t;
        t.name = nm;
        snprintf(t.t.name, sizeof t.t.name, "%s", nm.c_str());
        ctx->trackers.push_back(t);
        tid = (int)ctx->trackers.size() - 1;
    }
    if (ctx->events_used == ctx->events.size()) {
        EventPair e{};
        (void)hipEventCreate(&e.start);
        (void)hipEventCreate(&e.stop);
        ctx->events.push_back(e);
    }
    EventPair& e = ctx->events[ctx->events_used++];
    e.tracker = tid;
    ctx->launch_start = e.start;   // the next launch stamps them
    ctx->launch_stop = e.stop;
}

void track_end(pbf_ctx* ctx) {
    ctx->launch_start = nullptr;
    ctx->launch_stop = nullptr;
}

// Reads every recorded event pair (synchronising) into the trackers: the
// OCLPerfMon::UpdateTimings update (src/OCLPerfMon.cpp:36-50, weight 0.5).
void collect_timings(pbf_ctx* ctx) {
    for (size_t q = 0; q < ctx->events_used; ++q) {
        EventPair& e = ctx->events[q];
        float ms = 0.0f;
        if (hipEventSynchronize(e.stop) != hipSuccess) continue;
        if (hipEventElapsedTime(&ms, e.start, e.stop) != hipSuccess) continue;
        pbf_timing& t = ctx->trackers[e.tracker].t;
        t.last_ms = ms;
        t.ema_ms = ms * 0.5 + t.ema_ms * 0.5;
        t.sum_ms += ms;
        t.calls++;
    }
    ctx->events_used = 0;
}

// BouncePointQuad's per-quad constants (compute_delta.cl:3-9) in binary32,
// dot products summed left to right; normalize() as v * (1 / |v|).
QuadConsts quad_consts(const pbf_quad& q) {
    QuadConsts o{};
    o.b0 = q.origin[0]; o.b1 = q.origin[1]; o.b2 = q.origin[2];
    o.e00 = q.e0[0]; o.e01 = q.e0[1]; o.e02 = q.e0[2];
    o.e10 = q.e1[0]; o.e11 = q.e1[1]; o.e12 = q.e1[2];
    o.aa = o.e00 * o.e00 + o.e01 * o.e01 + o.e02 * o.e02;
    o.bb = o.e00 * o.e10 + o.e01 * o.e11 + o.e02 * o.e12;
    o.cc = o.e10 * o.e10 + o.e11 * o.e11 + o.e12 * o.e12;
    o.invdet = 1.0f / (o.aa * o.cc - o.bb * o.bb);
    const float cx = o.e01 * o.e12 - o.e02 * o.e11;
    const float cy = o.e02 * o.e10 - o.e00 * o.e12;
    const float cz = o.e00 * o.e11 - o.e01 * o.e10;
    const float il = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
    o.n0 = cx * il; o.n1 = cy * il; o.n2 = cz * il;
    o.mask_y = q.mask_y;
    o.mask_h = q.mask_h;
    return o;
}

// Replaces scene `scene`'s quads and mask with q, m (swapped in on
// success): one device buffer then holds every scene's quads and mask bits,
// and the scene constants point into it.  On failure the context is left as
// it was.  The caller has drained the stream.
int upload_geometry(pbf_ctx* ctx, uint32_t scene, std::vector<QuadConsts>& q, std::vector<uint32_t>& m,
                    uint32_t mask_w, uint32_t mask_rows, float edge_base) {
    ctx->quads.resize(ctx->S);
    ctx->mask_bits.resize(ctx->S);
    auto quads_of = [&](uint32_t s) -> const std::vector<QuadConsts>& { return s == scene ? q : ctx->quads[s]; };
    auto mask_of = [&](uint32_t s) -> const std::vector<uint32_t>& { return s == scene ? m : ctx->mask_bits[s]; };
    size_t bytes = 0;
    for (uint32_t s = 0; s < ctx->S; ++s) {
        bytes = (bytes + quads_of(s).size() * sizeof(QuadConsts) + 255) & ~size_t(255);
        bytes = (bytes + mask_of(s).size() * 4 + 255) & ~size_t(255);
    }
    char* d = nullptr;
    if (bytes && hipMalloc(&d, bytes) != hipSuccess)
        return fail(ctx, PBF_E_NOMEM, "collision quads: hipMalloc(%zu) failed", bytes);
    std::vector<SceneConsts> sc = ctx->sc;
    sc[scene].quad_count = (uint32_t)q.size();
    sc[scene].mask_w = q.empty() ? 0u : mask_w;
    sc[scene].mask_rows = q.empty() ? 0u : mask_rows;
    sc[scene].edge_base = q.empty() ? 0.0f : edge_base;
    size_t off = 0;
    hipError_t e = hipSuccess;
    for (uint32_t s = 0; s < ctx->S && e == hipSuccess; ++s) {
        const auto& qs = quads_of(s);
        const auto& ms = mask_of(s);
        sc[s].quads = qs.empty() ? nullptr : reinterpret_cast<const QuadConsts*>(d + off);
        if (!qs.empty()) e = hipMemcpy(d + off, qs.data(), qs.size() * sizeof(QuadConsts), hipMemcpyHostToDevice);
        off = (off + qs.size() * sizeof(QuadConsts) + 255) & ~size_t(255);
        sc[s].mask = ms.empty() ? nullptr : reinterpret_cast<const uint32_t*>(d + off);
        if (!ms.empty() && e == hipSuccess) e = hipMemcpy(d + off, ms.data(), ms.size() * 4, hipMemcpyHostToDevice);
        off = (off + ms.size() * 4 + 255) & ~size_t(255);
    }
    if (e == hipSuccess) e = hipMemcpy(ctx->sc_d, sc.data(), sizeof(SceneConsts) * ctx->S, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        // the old table is still valid: restore the device copy of the constants
        (void)hipMemcpy(ctx->sc_d, ctx->sc.data(), sizeof(SceneConsts) * ctx->S, hipMemcpyHostToDevice);
        return fail(ctx, PBF_E_HIP, "collision quads: %s", hipGetErrorString(e));
    }
    if (ctx->geom_d) (void)hipFree(ctx->geom_d);
    ctx->geom_d = d;
    ctx->sc.swap(sc);
    ctx->quads[scene].swap(q);
    ctx->mask_bits[scene].swap(m);
    return PBF_OK;
}

}  // namespace pbf

extern "C" {

const char* pbf_version(void) { return "pbf_amd 0.2 (gfx950)"; }

void pbf_default_options(pbf_options* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->walls = 1;
}

void pbf_params_finalize(pbf_params* p) {
    if (!p) return;
    p->h_2 = p->h * p->h;   // src/ParamUtils.cpp:97-99
    p->friendsCircles = 5;
    p->particlesPerCircle = 50;
}

int pbf_create(const pbf_params* params, const pbf_options* opt, pbf_ctx** out) {
    return create_impl(params, opt, 1, out);
}

int pbf_create_batch(const pbf_params* params, const pbf_options* opts, uint32_t scenes, pbf_ctx** out) {
    return create_impl(params, opts, scenes, out);
}

int pbf_scene_offsets(pbf_ctx* ctx, uint32_t* offsets) {
    if (!ctx || !offsets) return PBF_E_INVALID;
    for (uint32_t s = 0; s < ctx->S; ++s) offsets[s] = ctx->sc[s].first;
    offsets[ctx->S] = ctx->n;
    return PBF_OK;
}

int pbf_scene_count(pbf_ctx* ctx) { return ctx ? (int)ctx->S : PBF_E_INVALID; }

int pbf_friend_circle_thresholds(float h, float* t4) {
    if (!t4 || !(h > 0.0f)) return PBF_E_INVALID;
    const float h2 = h * h, min_r = 0.3f * h, span = h - min_r;
    circle_thresholds(h2, min_r, span, t4);
    return PBF_OK;
}

void pbf_destroy(pbf_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
}

const char* pbf_last_error(const pbf_ctx* ctx) { return ctx ? ctx->last_error.c_str() : g_create_error; }

int pbf_set_state(pbf_ctx* ctx, const float* x4, const float* v4, uint32_t n) {
    if (!ctx) return PBF_E_INVALID;
    if (!x4) return fail(ctx, PBF_E_INVALID, "x is NULL");
    if (n != ctx->n) return fail(ctx, PBF_E_INVALID, "n = %u but the context holds %u particles", n, ctx->n);
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->x, x4, 16 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
        if (v4)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->v, v4, 16 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
        else
            HIP_TRY(ctx, hipMemsetAsync(ctx->v, 0, 16 * (size_t)n, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    ctx->have_state = true;
    ctx->order_sorted = false;   // cells still hold the old keys until reset
    return PBF_OK;
}

int pbf_update_params(pbf_ctx* ctx, uint32_t scene, const pbf_params* params) {
    if (!ctx) return PBF_E_INVALID;
    if (scene >= ctx->S) return fail(ctx, PBF_E_INVALID, "scene %u out of range (%u scenes)", scene, ctx->S);
    const int rc = validate(params, ctx, scene);
    if (rc) return rc;
    const pbf_params& old = ctx->P[scene];
    if (params->particleCount != old.particleCount || params->gridBufSize != old.gridBufSize ||
        params->simIterations != old.simIterations)
        return fail(ctx, PBF_E_INVALID,
                    "scene %u: particleCount, gridBufSize and simIterations size the context; create a new one", scene);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const pbf::SceneConsts& prev = ctx->sc[scene];
    pbf::SceneConsts c = make_consts(*params, ctx->opt[scene], prev.cell_base, prev.first);
    c.quads = prev.quads;
    c.mask = prev.mask;
    c.quad_count = prev.quad_count;
    c.mask_w = prev.mask_w;
    c.mask_rows = prev.mask_rows;
    c.edge_base = prev.edge_base;
    HIP_TRY(ctx, hipMemcpy(ctx->sc_d + scene, &c, sizeof c, hipMemcpyHostToDevice));
    ctx->sc[scene] = c;
    ctx->P[scene] = *params;
    ctx->order_sorted = false;   // h may have moved every particle's cell: sort before the next friends build
    return PBF_OK;
}

int pbf_set_collision_quads(pbf_ctx* ctx, uint32_t scene, const pbf_quad* quads, uint32_t count, const uint32_t* mask,
                            uint32_t mask_w, uint32_t mask_rows, float edge_base) {
    if (!ctx) return PBF_E_INVALID;
    if (scene >= ctx->S) return fail(ctx, PBF_E_INVALID, "scene %u out of range (%u scenes)", scene, ctx->S);
    if (count > PBF_MAX_QUADS) return fail(ctx, PBF_E_INVALID, "%u quads > %d", count, PBF_MAX_QUADS);
    std::vector<pbf::QuadConsts> qc;
    std::vector<uint32_t> bits;
    if (count) {
        if (!quads || !mask) return fail(ctx, PBF_E_INVALID, "quads or mask is NULL");
        if (mask_w == 0 || mask_rows == 0 || (uint64_t)mask_w * mask_rows > (1ull << 31))
            return fail(ctx, PBF_E_INVALID, "mask must have 1..2^31 texels (got %u x %u)", mask_w, mask_rows);
        if (!isfinite(edge_base)) return fail(ctx, PBF_E_INVALID, "edge_base must be finite");
        for (uint32_t k = 0; k < count; ++k) {
            const pbf_quad& q = quads[k];
            for (int a = 0; a < 3; ++a)
                if (!isfinite(q.origin[a]) || !isfinite(q.e0[a]) || !isfinite(q.e1[a]))
                    return fail(ctx, PBF_E_INVALID, "quad %u: non-finite corner or edge", k);
            if (q.mask_y < -(1 << 24) || q.mask_y > (1 << 24) || q.mask_h < -(1 << 24) || q.mask_h > (1 << 24))
                return fail(ctx, PBF_E_INVALID, "quad %u: mask_y and mask_h must be within +-2^24", k);
            const pbf::QuadConsts c = pbf::quad_consts(q);
            if (!(c.aa * c.cc - c.bb * c.bb > 0.0f) || !isfinite(c.invdet) || !isfinite(c.n0) || !isfinite(c.n1) ||
                !isfinite(c.n2))
                return fail(ctx, PBF_E_INVALID, "quad %u: its edges are parallel or degenerate", k);
            qc.push_back(c);
        }
        const uint32_t stride = (mask_w + 31) / 32;
        bits.assign((size_t)stride * mask_rows, 0u);
        for (uint32_t y = 0; y < mask_rows; ++y)
            for (uint32_t x = 0; x < mask_w; ++x)
                if (mask[(size_t)y * mask_w + x]) bits[(size_t)y * stride + (x >> 5)] |= 1u << (x & 31);
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return pbf::upload_geometry(ctx, scene, qc, bits, mask_w, mask_rows, edge_base);
}

int pbf_step(pbf_ctx* ctx, uint32_t pause) {
    bool skip;
    const int rc = step_ready(ctx, &skip);
    if (rc != PBF_OK || skip) return rc;
    if (ctx->use_graph && !ctx->profiling) return step_graph(ctx, pause, 1);
    return enqueue_step(ctx, pause);
}

int pbf_run(pbf_ctx* ctx, uint32_t steps, uint32_t pause) {
    if (steps == 0) return ctx ? PBF_OK : PBF_E_INVALID;
    bool skip;
    const int rc0 = step_ready(ctx, &skip);
    if (rc0 != PBF_OK || skip) return rc0;
    uint32_t s = 0;
    if (ctx->use_graph && !ctx->profiling)
        for (; steps - s >= pbf::kRunGraphSteps; s += pbf::kRunGraphSteps) {
            const int rc = step_graph(ctx, pause, pbf::kRunGraphSteps);
            if (rc) return rc;
        }
    for (; s < steps; ++s) {
        const int rc = pbf_step(ctx, pause);
        if (rc) return rc;
    }
    return PBF_OK;
}

int pbf_set_graph(pbf_ctx* ctx, int enable) {
    if (!ctx) return PBF_E_INVALID;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (!enable) drop_graphs(ctx);
    ctx->use_graph = enable != 0;
    return PBF_OK;
}

int pbf_sync(pbf_ctx* ctx) {
    if (!ctx) return PBF_E_INVALID;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PBF_OK;
}

int pbf_get_state(pbf_ctx* ctx, float* x4, float* v4, float* xs4, float* rho, float* lambda, float* omega4) {
    if (!ctx) return PBF_E_INVALID;
    const size_t n = ctx->n;
    if (n == 0) return PBF_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (x4) HIP_TRY(ctx, hipMemcpy(x4, ctx->x, 16 * n, hipMemcpyDeviceToHost));
    if (v4) HIP_TRY(ctx, hipMemcpy(v4, ctx->v, 16 * n, hipMemcpyDeviceToHost));
    if (xs4) HIP_TRY(ctx, hipMemcpy(xs4, ctx->xs[ctx->xs_cur], 16 * n, hipMemcpyDeviceToHost));
    if (rho) HIP_TRY(ctx, hipMemcpy(rho, ctx->rho, 4 * n, hipMemcpyDeviceToHost));
    if (lambda) HIP_TRY(ctx, hipMemcpy(lambda, ctx->lambda, 4 * n, hipMemcpyDeviceToHost));
    if (omega4) HIP_TRY(ctx, hipMemcpy(omega4, ctx->omega, 16 * n, hipMemcpyDeviceToHost));
    return PBF_OK;
}

int pbf_get_structure(pbf_ctx* ctx, uint32_t* keys_sorted, uint32_t* perm, uint32_t* cells, uint32_t* friends) {
    if (!ctx) return PBF_E_INVALID;
    const size_t n = ctx->n;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (keys_sorted && n) HIP_TRY(ctx, hipMemcpy(keys_sorted, ctx->keys_sorted, 4 * n, hipMemcpyDeviceToHost));
    if (perm && n) HIP_TRY(ctx, hipMemcpy(perm, ctx->perm, 4 * n, hipMemcpyDeviceToHost));
    if (cells) HIP_TRY(ctx, hipMemcpy(cells, ctx->cells, 8 * (size_t)ctx->total_cells, hipMemcpyDeviceToHost));
    if (friends && n) {
        // counts[5][n] unpacked from the 6-bit fields, then indices[5][50][n]
        // regrouped from the blocked circle slots (zero past each count)
        std::vector<uint32_t> packed(n);
        HIP_TRY(ctx, hipMemcpy(packed.data(), ctx->nb_cnt, 4 * n, hipMemcpyDeviceToHost));
        std::vector<uint32_t> slots(pbf::friend_slots(n));
        HIP_TRY(ctx, hipMemcpy(slots.data(), ctx->nb_list, 4 * slots.size(), hipMemcpyDeviceToHost));
        std::vector<uint32_t> sweep(((n + 63) / 64) * 64 * (size_t)pbf::kQuads * 4);
        HIP_TRY(ctx, hipMemcpy(sweep.data(), ctx->nb_quads, 4 * sweep.size(), hipMemcpyDeviceToHost));
        uint32_t* idx = friends + (size_t)pbf::kCircles * n;
        memset(idx, 0, 4 * n * pbf::kCircles * pbf::kPerCircle);
        for (size_t i = 0; i < n; ++i)
            for (int q = 0; q < pbf::kCircles; ++q) {
                const uint32_t cnt = (packed[i] >> (6 * q)) & 63u;
                friends[(size_t)q * n + i] = cnt;
                for (uint32_t f = 0; f < cnt; ++f)
                    idx[((size_t)q * pbf::kPerCircle + f) * n + i] =
                        pbf::friend_at(slots.data(), sweep.data(), (uint32_t)i, q, (int)f);
            }
    }
    return PBF_OK;
}

int pbf_device_buffers(pbf_ctx* ctx, pbf_dev_bufs* out) {
    if (!ctx || !out) return PBF_E_INVALID;
    out->positions = ctx->x;
    out->predicted = ctx->xs[ctx->xs_cur];
    out->velocities = ctx->v;
    out->density = ctx->rho;
    out->lambda = ctx->lambda;
    out->omega = ctx->omega;
    out->stream = (void*)ctx->stream;
    out->n = ctx->n;
    return PBF_OK;
}

int pbf_set_profiling(pbf_ctx* ctx, int enable) {
    if (!ctx) return PBF_E_INVALID;
    if (!enable && ctx->profiling) pbf::collect_timings(ctx);
    ctx->profiling = enable != 0;
    return PBF_OK;
}

int pbf_timings(pbf_ctx* ctx, pbf_timing* out, int cap, int* count) {
    if (!ctx) return PBF_E_INVALID;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    pbf::collect_timings(ctx);
    const int total = (int)ctx->trackers.size();
    if (count) *count = total;
    for (int q = 0; q < total && q < cap && out; ++q) out[q] = ctx->trackers[q].t;
    return PBF_OK;
}

int pbf_reset_timings(pbf_ctx* ctx) {
    if (!ctx) return PBF_E_INVALID;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    pbf::collect_timings(ctx);
    ctx->trackers.clear();
    return PBF_OK;
}

int pbf_error_flags(pbf_ctx* ctx, uint32_t* flags, uint64_t* dropped, int reset) {
    if (!ctx) return PBF_E_INVALID;
    uint32_t w[pbf::kErrWords];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(w, ctx->err, sizeof w, hipMemcpyDeviceToHost));
    if (flags) *flags = w[0];
    if (dropped) {
        uint64_t d;
        memcpy(&d, w + 2, 8);
        *dropped = d;
    }
    if (reset) {
        HIP_TRY(ctx, hipMemsetAsync(ctx->err, 0, sizeof w, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return PBF_OK;
}

int pbf_read_log(pbf_ctx* ctx, pbf_log_msg* out, uint32_t cap, uint32_t* count, uint64_t* lost) {
    if (!ctx || !count || (cap && !out)) return PBF_E_INVALID;
    static_assert(pbf::kLogRing == 4 * PBF_LOG_CAPACITY, "four words per message");
    std::vector<uint32_t> w(pbf::kLogRing);
    uint32_t at = 0;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(&at, ctx->err + pbf::kLogIndex, 4, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(w.data(), ctx->err + pbf::kLogRingStart, 4 * w.size(), hipMemcpyDeviceToHost));
    // indices are in words and advance by 4; unsigned differences survive the wrap
    uint32_t from = ctx->log_read;
    uint64_t missed = 0;
    if (at - from > pbf::kLogRing) {
        missed = (at - from - pbf::kLogRing) / 4;
        from = at - pbf::kLogRing;
    }
    uint32_t k = 0;
    for (; k < cap && from != at; ++k, from += 4) {
        const uint32_t* m = w.data() + (from % pbf::kLogRing);
        out[k] = {m[0], m[1], m[2], m[3]};
    }
    ctx->log_read = from;
    *count = k;
    if (lost) *lost = missed;
    return PBF_OK;
}

int pbf_stats(pbf_ctx* ctx, pbf_stats_t* out) {
    if (!ctx || !out) return PBF_E_INVALID;
    uint32_t w[pbf::kErrWords];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(w, ctx->err, sizeof w, hipMemcpyDeviceToHost));
    memset(out, 0, sizeof *out);
    out->unstaged_tiles = w[4];
    out->friends_staged = ctx->friends_staged ? 1u : 0u;
    out->friends_lanes = ctx->friends_staged ? 1u : pbf::friends_lanes_for(ctx);
    out->sort_passes = ctx->sort.passes;
    out->sort_tiles = ctx->sort.tiles;
    out->total_cells = ctx->total_cells;
    out->particles = ctx->n;
    out->scenes = ctx->S;
    for (int f = 0; f < pbf::F_COUNT; ++f) {
        out->workgroup_size[f] = 1u << ctx->bs_log2[f];
        out->tiles_per_workgroup[f] = ctx->tiles_per_wg[f];
    }
    return PBF_OK;
}

int pbf_set_friends_mode(pbf_ctx* ctx, int staged) {
    if (!ctx) return PBF_E_INVALID;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->friends_staged = staged != 0;
    return PBF_OK;
}

int pbf_set_friends_lanes(pbf_ctx* ctx, uint32_t lanes) {
    if (!ctx) return PBF_E_INVALID;
    if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16))
        return fail(ctx, PBF_E_INVALID, "friends lanes %u is not one of 0, 1, 2, 4, 8, 16", lanes);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    ctx->friends_lanes = lanes;
    return PBF_OK;
}

int pbf_set_workgroup_sizes(pbf_ctx* ctx, const uint32_t* sizes) {
    if (!ctx || !sizes) return PBF_E_INVALID;
    int lg[pbf::F_COUNT];
    for (int f = 0; f < pbf::F_COUNT; ++f) {
        int l = 6;
        while (l <= 9 && (1u << l) != sizes[f]) ++l;
        if (l > 9) return fail(ctx, PBF_E_INVALID, "workgroup size %u is not one of 64, 128, 256, 512", sizes[f]);
        lg[f] = l;
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    for (int f = 0; f < pbf::F_COUNT; ++f) ctx->bs_log2[f] = lg[f];
    return PBF_OK;
}

int pbf_set_tiles_per_workgroup(pbf_ctx* ctx, const uint32_t* tiles) {
    if (!ctx || !tiles) return PBF_E_INVALID;
    for (int f = 0; f < pbf::F_COUNT; ++f)
        if (tiles[f] > 64) return fail(ctx, PBF_E_INVALID, "tiles per workgroup %u > 64", tiles[f]);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    for (int f = 0; f < pbf::F_COUNT; ++f) ctx->tiles_per_wg[f] = tiles[f];
    return PBF_OK;
}

// Times every workgroup size for every family on the current state and keeps
// the fastest per family.  The particle state is restored afterwards; the
// next step re-sorts (the order is no longer known to be sorted).
int pbf_autotune(pbf_ctx* ctx, uint32_t steps, uint32_t* chosen) {
    if (!ctx) return PBF_E_INVALID;
    if (!ctx->have_state) return fail(ctx, PBF_E_STATE, "pbf_autotune before pbf_set_state");
    if (ctx->n == 0) return PBF_OK;
    if (steps == 0) steps = 1;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const size_t bytes = 16 * (size_t)ctx->n;
    float4 *sx = nullptr, *sv = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&sx, bytes));
    if (hipMalloc((void**)&sv, bytes) != hipSuccess) {
        (void)hipFree(sx);
        return fail(ctx, PBF_E_NOMEM, "pbf_autotune: no memory for the state copy");
    }
    int rc = PBF_OK;
    const bool prev_graph = ctx->use_graph, prev_prof = ctx->profiling;
    ctx->use_graph = false;
    pbf::collect_timings(ctx);
    const std::vector<pbf::Tracker> saved_trackers = ctx->trackers;
    double best[pbf::F_COUNT];
    int best_l[pbf::F_COUNT];
    for (int f = 0; f < pbf::F_COUNT; ++f) {
        best[f] = 1e300;
        best_l[f] = ctx->bs_log2[f];
    }
    auto restore = [&]() -> int {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->x, sx, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->v, sv, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        ctx->order_sorted = false;
        return PBF_OK;
    };
    static const char* kFamilyTrackers[pbf::F_COUNT] = {"buildFriendsList", "computeScaling", "computeDelta",
                                                        "applyViscosity", "applyVorticity"};
    if (hipMemcpyAsync(sx, ctx->x, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(sv, ctx->v, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        rc = fail(ctx, PBF_E_HIP, "pbf_autotune: state copy failed");
    }
    // one timed trial from the saved state: per-family milliseconds in ms
    auto trial_ms = [&](double* ms) -> int {
        int r;
        if ((r = restore()) != PBF_OK) return r;
        ctx->profiling = false;
        if ((r = enqueue_step(ctx, 0)) != PBF_OK) return r;   // warm-up
        ctx->trackers.clear();
        ctx->profiling = true;
        for (uint32_t s = 0; s < steps && r == PBF_OK; ++s) r = enqueue_step(ctx, 0);
        if (r != PBF_OK) return r;
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, PBF_E_HIP, "pbf_autotune: sync");
        pbf::collect_timings(ctx);
        ctx->profiling = false;
        for (int f = 0; f < pbf::F_COUNT; ++f) {
            ms[f] = 0.0;
            const size_t len = strlen(kFamilyTrackers[f]);
            for (const pbf::Tracker& t : ctx->trackers)
                if (t.name.compare(0, len, kFamilyTrackers[f]) == 0) ms[f] += t.t.sum_ms;
        }
        return PBF_OK;
    };
    // workgroup sizes (and the friends build variant), one tile per workgroup
    // A candidate replaces the default (256, global-read friends build, one
    // tile per workgroup) only when it is more than kMargin faster: the
    // trials are a few substeps long, and choices within their noise would
    // otherwise change from run to run.
    constexpr double kMargin = 0.97;
    double ms_at[2][4][pbf::F_COUNT];   // [staged trial][size][family]
    for (int f = 0; f < pbf::F_COUNT; ++f) ctx->tiles_per_wg[f] = 1;
    for (int trial = 0; trial < 8 && rc == PBF_OK; ++trial) {
        const int l = 6 + (trial & 3);
        const bool staged_trial = trial >= 4;
        ctx->friends_staged = staged_trial;
        for (int f = 0; f < pbf::F_COUNT; ++f) ctx->bs_log2[f] = l;
        rc = trial_ms(ms_at[staged_trial][l - 6]);
    }
    if (rc == PBF_OK) {
        for (int f = 0; f < pbf::F_COUNT; ++f) {   // the fastest size that beats 256 by the margin
            const double dflt = ms_at[0][2][f];
            best_l[f] = 8;
            best[f] = dflt;
            for (int l = 6; l <= 9; ++l)
                if (ms_at[0][l - 6][f] < kMargin * dflt && ms_at[0][l - 6][f] < best[f]) {
                    best[f] = ms_at[0][l - 6][f];
                    best_l[f] = l;
                }
        }
        // the staged friends build against the chosen global-read one
        ctx->friends_staged = false;
        const double direct = best[pbf::F_FRIENDS];
        for (int l = 6; l <= 9; ++l)
            if (ms_at[1][l - 6][pbf::F_FRIENDS] < kMargin * direct &&
                ms_at[1][l - 6][pbf::F_FRIENDS] < best[pbf::F_FRIENDS]) {
                best[pbf::F_FRIENDS] = ms_at[1][l - 6][pbf::F_FRIENDS];
                best_l[pbf::F_FRIENDS] = l;
                ctx->friends_staged = true;
            }
    }
    for (int f = 0; f < pbf::F_COUNT; ++f) ctx->bs_log2[f] = best_l[f];
    // tiles per workgroup at those sizes: 4, then a persistent grid, against
    // one tile per workgroup (the staged friends build runs one tile each)
    uint32_t best_t[pbf::F_COUNT] = {1, 1, 1, 1, 1};
    double one_tile[pbf::F_COUNT];
    for (int f = 0; f < pbf::F_COUNT; ++f) one_tile[f] = best[f];
    static const uint32_t kTileTrials[2] = {4, 0};
    for (int k = 0; k < 2 && rc == PBF_OK; ++k) {
        for (int f = 0; f < pbf::F_COUNT; ++f) ctx->tiles_per_wg[f] = kTileTrials[k];
        double ms[pbf::F_COUNT];
        if ((rc = trial_ms(ms)) != PBF_OK) break;
        for (int f = 0; f < pbf::F_COUNT; ++f) {
            if (f == pbf::F_FRIENDS && ctx->friends_staged) continue;
            if (ms[f] < kMargin * one_tile[f] && ms[f] < best[f]) {
                best[f] = ms[f];
                best_t[f] = kTileTrials[k];
            }
        }
    }
    for (int f = 0; f < pbf::F_COUNT; ++f) {
        ctx->tiles_per_wg[f] = best_t[f];
        if (chosen) chosen[f] = 1u << best_l[f];
    }
    // lanes per particle of the global-read friends build, on scenes of at
    // most four 64-particle waves per CU (above, one lane is faster): 1, 4
    // and 16 against the default choice
    if (rc == PBF_OK && !ctx->friends_staged && ctx->friends_lanes == 0 &&
        ((uint64_t)ctx->n + 63) / 64 <= 4ull * (uint64_t)std::max(ctx->cu_count, 1)) {
        const uint32_t dflt = pbf::friends_lanes_for(ctx);
        double ms[pbf::F_COUNT];
        if ((rc = trial_ms(ms)) == PBF_OK) {
            const double at_default = ms[pbf::F_FRIENDS];
            double best_ms = at_default;
            uint32_t best_lanes = 0;
            static const uint32_t kLaneTrials[3] = {1, 4, 16};
            for (int k = 0; k < 3 && rc == PBF_OK; ++k) {
                if (kLaneTrials[k] == dflt) continue;
                ctx->friends_lanes = kLaneTrials[k];
                if ((rc = trial_ms(ms)) != PBF_OK) break;
                if (ms[pbf::F_FRIENDS] < kMargin * at_default && ms[pbf::F_FRIENDS] < best_ms) {
                    best_ms = ms[pbf::F_FRIENDS];
                    best_lanes = kLaneTrials[k];
                }
            }
            ctx->friends_lanes = best_lanes;
        }
    }
    if (rc == PBF_OK) rc = restore();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(sx);
    (void)hipFree(sv);
    ctx->trackers = saved_trackers;
    ctx->profiling = prev_prof;
    drop_graphs(ctx);
    ctx->use_graph = prev_graph;
    return rc;
}

}  // extern "C"
