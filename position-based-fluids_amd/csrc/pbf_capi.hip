// pbf_capi.hip -- the exported C ABI (include/pbf/pbf.h): context set-up,
// the substep driver and the host-side getters.
//
// pbf_step enqueues the whole substep on the context's stream with no host
// synchronisation (the reference blocks six times per step inside its sort,
// src/Simulation.cpp:488, and once at GL unlock, :551).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdarg>
#include <new>

#include "pbf_internal.h"

namespace {

std::string g_create_error;

int fail(pbf_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->last_error = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIP_TRY(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail((ctx), PBF_E_HIP, "%s: %s", #call, hipGetErrorString(e_));    \
    } while (0)

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

template <class T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
}

void free_all(pbf_ctx* ctx) {
    void* ptrs[] = {ctx->x, ctx->x_alt, ctx->xs[0], ctx->xs[1], ctx->v, ctx->v_alt, ctx->omega, ctx->rho,
                    ctx->lambda, ctx->keys_in, ctx->keys_tmp, ctx->keys_sorted, ctx->vals_a, ctx->vals_b,
                    ctx->perm, ctx->sort.ws, ctx->cells, ctx->nb_cnt, ctx->nb_idx, ctx->err};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (pbf::EventPair& e : ctx->events) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    ctx->events.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

int validate(const pbf_params* p, pbf_ctx* ctx) {
    if (!p) return fail(ctx, PBF_E_INVALID, "params is NULL");
    if (p->particleCount >= (1u << 30)) return fail(ctx, PBF_E_INVALID, "particleCount %u >= 2^30", p->particleCount);
    if (!(p->h > 0.0f) || !isfinite(p->h)) return fail(ctx, PBF_E_INVALID, "h must be positive and finite");
    if (!(p->timeStep > 0.0f)) return fail(ctx, PBF_E_INVALID, "timeStep must be positive");
    if (p->simIterations < 1 || p->simIterations > 64)
        return fail(ctx, PBF_E_INVALID, "simIterations must be in [1, 64] (got %u)", p->simIterations);
    if (p->gridBufSize < 1 || p->gridBufSize > (1u << 30))
        return fail(ctx, PBF_E_INVALID, "gridBufSize must be in [1, 2^30] (got %u)", p->gridBufSize);
    if (p->friendsCircles != (uint32_t)pbf::kCircles || p->particlesPerCircle != (uint32_t)pbf::kPerCircle)
        return fail(ctx, PBF_E_INVALID, "friendsCircles/particlesPerCircle must be 5/50 (call pbf_params_finalize)");
    if (!(p->restDensity != 0.0f)) return fail(ctx, PBF_E_INVALID, "restDensity must be nonzero");
    return PBF_OK;
}

pbf::StepConsts make_consts(const pbf_params& P, const pbf_options& o) {
    pbf::StepConsts c{};
    c.h = P.h;
    c.h2 = P.h_2;
    c.min_r = 0.3f * P.h;            // build_friends_list.cl:32
    c.span = P.h - c.min_r;          // build_friends_list.cl:75
    c.dt = P.timeStep;
    if (o.use_gravity_vector) {
        c.g[0] = o.gravity[0];
        c.g[1] = o.gravity[1];
        c.g[2] = o.gravity[2];
    } else {
        c.g[0] = 0.0f;               // predict_positions.cl:13
        c.g[1] = -P.garvity;
        c.g[2] = 0.0f;
    }
    c.rest_density = P.restDensity;
    c.epsilon = P.epsilon;
    kernel_factors(P.h, &c.poly6, &c.spiky);
    c.st_k = P.surfaceTenstionK;
    const float dqh = P.surfaceTenstionDist * P.h;   // compute_delta.cl:81-82
    const float q2 = dqh * dqh;
    const float tq = P.h_2 - q2;
    c.poly6_q = tq * tq * tq;
    c.viscosity = P.viscosityFactor;
    c.vorticity = P.vorticityFactor;
    c.lo[0] = P.xMin; c.lo[1] = P.yMin; c.lo[2] = P.zMin;
    c.hi[0] = P.xMax; c.hi[1] = P.yMax; c.hi[2] = P.zMax;
    c.walls = o.walls ? 1u : 0u;
    c.gbs = P.gridBufSize;
    c.gbs_mask = ((P.gridBufSize & (P.gridBufSize - 1)) == 0) ? P.gridBufSize - 1 : 0u;
    if (P.gridBufSize == 1) c.gbs_mask = 0;  // m % 1 == 0 handled by the modulo path
    c.n = P.particleCount;
    return c;
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
    (void)hipEventRecord(e.start, ctx->stream);
}

void track_end(pbf_ctx* ctx) {
    if (!ctx->profiling || ctx->events_used == 0) return;
    (void)hipEventRecord(ctx->events[ctx->events_used - 1].stop, ctx->stream);
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

}  // namespace pbf

extern "C" {

const char* pbf_version(void) { return "pbf_amd 0.1 (gfx950)"; }

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
    if (!out) return fail(nullptr, PBF_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(params, nullptr);
    if (rc) return rc;
    pbf_ctx* ctx = new (std::nothrow) pbf_ctx();
    if (!ctx) return fail(nullptr, PBF_E_NOMEM, "host allocation failed");
    ctx->P = *params;
    if (opt)
        ctx->opt = *opt;
    else
        pbf_default_options(&ctx->opt);
    ctx->c = make_consts(ctx->P, ctx->opt);
    ctx->n = params->particleCount;

    auto bail = [&](int code) {
        g_create_error = ctx->last_error;
        free_all(ctx);
        delete ctx;
        return code;
    };
#define CREATE_TRY(call)                                                                                    \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            fail(ctx, e_ == hipErrorOutOfMemory ? PBF_E_NOMEM : PBF_E_HIP, "%s: %s", #call,                 \
                 hipGetErrorString(e_));                                                                    \
            return bail(e_ == hipErrorOutOfMemory ? PBF_E_NOMEM : PBF_E_HIP);                               \
        }                                                                                                   \
    } while (0)

    CREATE_TRY(hipGetDevice(&ctx->device));
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
    CREATE_TRY(dalloc(&ctx->cells, 2 * (size_t)params->gridBufSize));
    CREATE_TRY(dalloc(&ctx->nb_cnt, n));
    CREATE_TRY(dalloc(&ctx->nb_idx, n * pbf::kCircles * pbf::kPerCircle));
    CREATE_TRY(dalloc(&ctx->err, pbf::kErrWords));

    // sort geometry: enough 8-bit passes for keys < gridBufSize
    uint32_t bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)params->gridBufSize) ++bits;
    if (bits == 0) bits = 1;
    ctx->sort.bits = bits;
    ctx->sort.passes = (bits + pbf::kRadixBits - 1) / pbf::kRadixBits;
    ctx->sort.tiles = (uint32_t)((n + pbf::kSortTile - 1) / pbf::kSortTile);
    if (ctx->sort.tiles == 0) ctx->sort.tiles = 1;
    const size_t hist_words = (size_t)pbf::kMaxPasses * pbf::kRadix;
    const size_t ctr_words = 64;
    const size_t status_words = (size_t)ctx->sort.passes * ctx->sort.tiles * pbf::kRadix;
    ctx->sort.ws_words = hist_words + ctr_words + status_words;
    CREATE_TRY(dalloc(&ctx->sort.ws, ctx->sort.ws_words));
    ctx->sort.hist = ctx->sort.ws;
    ctx->sort.ctr = ctx->sort.ws + hist_words;
    ctx->sort.status = ctx->sort.ws + hist_words + ctr_words;

    CREATE_TRY(hipMemsetAsync(ctx->cells, 0xFF, 8 * (size_t)params->gridBufSize, ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->nb_cnt, 0, 4 * std::max<size_t>(n, 1), ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->v, 0, 16 * std::max<size_t>(n, 1), ctx->stream));
    CREATE_TRY(hipMemsetAsync(ctx->err, 0, 4 * pbf::kErrWords, ctx->stream));
    CREATE_TRY(hipStreamSynchronize(ctx->stream));
#undef CREATE_TRY
    *out = ctx;
    return PBF_OK;
}

void pbf_destroy(pbf_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
}

const char* pbf_last_error(const pbf_ctx* ctx) { return ctx ? ctx->last_error.c_str() : g_create_error.c_str(); }

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
    ctx->have_sorted = false;
    return PBF_OK;
}

int pbf_step(pbf_ctx* ctx, uint32_t pause) {
    if (!ctx) return PBF_E_INVALID;
    if (!ctx->have_state) return fail(ctx, PBF_E_STATE, "pbf_step before pbf_set_state");
    if (ctx->n == 0) return PBF_OK;
    ctx->pause_now = pause ? 1u : 0u;
    // Simulation.cpp:563-567: predict always; radixsort unless paused.  With
    // no sort yet there are no sorted keys, so the first step always sorts.
    if (pause && ctx->have_sorted) {
        HIP_TRY(ctx, pbf::launch_predict_only(ctx, 1u));
    } else {
        HIP_TRY(ctx, pbf::launch_clear(ctx));
        HIP_TRY(ctx, pbf::launch_predict_hash(ctx, ctx->have_sorted));
        HIP_TRY(ctx, pbf::launch_sort(ctx));
        HIP_TRY(ctx, pbf::launch_gather(ctx));
        HIP_TRY(ctx, pbf::launch_cells(ctx));
        ctx->have_sorted = true;
    }
    HIP_TRY(ctx, pbf::launch_friends(ctx));
    const int K = (int)ctx->P.simIterations;
    for (int it = 0; it < K; ++it) {
        HIP_TRY(ctx, pbf::launch_lambda(ctx, it));
        HIP_TRY(ctx, pbf::launch_delta(ctx, it, it == K - 1));
    }
    HIP_TRY(ctx, pbf::launch_viscosity(ctx));
    HIP_TRY(ctx, pbf::launch_vorticity(ctx));
    return PBF_OK;
}

int pbf_run(pbf_ctx* ctx, uint32_t steps, uint32_t pause) {
    for (uint32_t s = 0; s < steps; ++s) {
        const int rc = pbf_step(ctx, pause);
        if (rc) return rc;
    }
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
    if (cells)
        HIP_TRY(ctx, hipMemcpy(cells, ctx->cells, 8 * (size_t)ctx->P.gridBufSize, hipMemcpyDeviceToHost));
    if (friends && n) {
        // counts[5][n] unpacked from the 6-bit fields, then indices[5][50][n]
        std::vector<uint32_t> packed(n);
        HIP_TRY(ctx, hipMemcpy(packed.data(), ctx->nb_cnt, 4 * n, hipMemcpyDeviceToHost));
        for (int q = 0; q < pbf::kCircles; ++q)
            for (size_t i = 0; i < n; ++i) friends[(size_t)q * n + i] = (packed[i] >> (6 * q)) & 63u;
        uint32_t* idx = friends + (size_t)pbf::kCircles * n;
        HIP_TRY(ctx, hipMemcpy(idx, ctx->nb_idx, 4 * n * pbf::kCircles * pbf::kPerCircle, hipMemcpyDeviceToHost));
        // slots past each circle's count hold stale data; zero them
        for (int q = 0; q < pbf::kCircles; ++q)
            for (int f = 0; f < pbf::kPerCircle; ++f) {
                uint32_t* row = idx + ((size_t)q * pbf::kPerCircle + f) * n;
                for (size_t i = 0; i < n; ++i)
                    if ((uint32_t)f >= friends[(size_t)q * n + i]) row[i] = 0;
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

}  // extern "C"
