/*
 * pbf.h -- C ABI of the MI355X PBF substep library (libpbf_amd.so).
 *
 * The library replaces the per-substep kernel chain that the reference's
 * Simulation class drives (src/Simulation.hpp:105-150,
 * src/Simulation.cpp:554-621).  Each entry point names the reference
 * interface it stands in for.  Conventions:
 *   - every function returns PBF_OK (0) or a negative PBF_E* code; no C++
 *     exception crosses this boundary; pbf_last_error() gives the message;
 *   - the library owns device memory, the caller owns host pointers;
 *   - one HIP stream per context; a context is not thread-safe;
 *   - pbf_step is asynchronous; the getters synchronise the stream.
 * Arrays of float4 are float[4*n], particle-major (x, y, z, w).
 */
#ifndef PBF_PBF_H
#define PBF_PBF_H

#include <stdint.h>
#include "params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_OK 0
#define PBF_E_INVALID (-1)   /* bad argument or parameter                  */
#define PBF_E_HIP (-2)       /* HIP runtime error (message has the detail) */
#define PBF_E_NOMEM (-3)     /* device allocation failed                   */
#define PBF_E_STATE (-4)     /* call out of order (e.g. step before state) */

/* Error/diagnostic flags accumulated on the device (pbf_error_flags). */
#define PBF_FLAG_SORT_TIMEOUT 1u   /* radix-sort look-back wait hit its bound    */
#define PBF_FLAG_FRIENDS_OVERFLOW 2u /* a circle held 50 friends; extras dropped  */
#define PBF_FLAG_NAN 4u            /* non-finite predicted position seen        */
#define PBF_FLAG_INTERNAL 8u       /* an internal bounds check failed (a result
                                      of this step is invalid)                 */

typedef struct pbf_ctx pbf_ctx;

/* Settings with no field in the reference Parameters block. */
typedef struct pbf_options {
    float gravity[3];          /* used when use_gravity_vector != 0          */
    uint32_t use_gravity_vector; /* 0: gravity = (0, -garvity, 0) as in
                                  predict_positions.cl:13                    */
    uint32_t walls;            /* 1: six box planes at x/y/zMin/Max in the
                                  delta stage; 0: no walls                   */
    uint32_t reserved[8];
} pbf_options;

/* Fills *o with the defaults: reference gravity, walls on. */
void pbf_default_options(pbf_options *o);

/* Sets h_2 = h*h, friendsCircles = 5, particlesPerCircle = 50
 * (src/ParamUtils.cpp:97-99). */
void pbf_params_finalize(pbf_params *p);

/* Parses the reference's .par text (src/ParamUtils.cpp:24-100: '#'
 * comments, case-insensitive keys, unknown keys skipped) into *p and
 * finalizes it.  Returns the number of unknown keys, or a PBF_E* code. */
int pbf_load_parameters(const char *text, pbf_params *p);

/* Simulation::Simulation + InitBuffers + InitCells + InitKernels
 * (src/Simulation.cpp:56-63,206-255,129-204): allocates every device
 * buffer for params->particleCount particles. opt may be NULL.
 * PBF_E_INVALID unless h is finite and >= 2^-20, timeStep > 0,
 * simIterations in [1, 64], gridBufSize in [1, 2^30], restDensity != 0 and
 * friendsCircles / particlesPerCircle are 5 / 50 (pbf_params_finalize). */
int pbf_create(const pbf_params *params, const pbf_options *opt, pbf_ctx **out);

/* A batch of independent scenes stepped together: params[s] and opts[s]
 * (opts may be NULL) describe scene s; all scenes share simIterations.
 * Particles of scene s occupy [offset[s], offset[s+1]) of every array
 * (pbf_scene_offsets), friend and permutation indices are global, cell keys
 * are offset by the preceding scenes' gridBufSize.  Each stage of a step is
 * one launch over all scenes.  No counterpart in the reference, which runs
 * one scene per Simulation. */
int pbf_create_batch(const pbf_params *params, const pbf_options *opts, uint32_t scenes, pbf_ctx **out);
int pbf_scene_count(pbf_ctx *ctx);
int pbf_scene_offsets(pbf_ctx *ctx, uint32_t *offsets /* scenes + 1 */);

/* Initial particle state, the role of Simulation::CreateParticles
 * (src/Simulation.cpp:71-101): n must equal particleCount; v may be NULL
 * (zero velocities). */
int pbf_set_state(pbf_ctx *ctx, const float *x4, const float *v4, uint32_t n);

/* Live parameter change, the hot-reload path of Runner::run
 * (src/Runner.cpp:43-77: a changed scenario is reloaded and the kernels are
 * rebuilt with its constants, src/Simulation.cpp:129-204): scene `scene`
 * takes params for its next step (h, timeStep, restDensity, epsilon,
 * garvity, viscosity/vorticity factors, surface tension, box bounds) and the
 * next step re-sorts even when paused.  particleCount, gridBufSize and
 * simIterations size the context and must not change (PBF_E_INVALID: create
 * a new context, as the reference re-runs InitBuffers).  The particle state,
 * options and collision quads are kept.  Waits for queued work. */
int pbf_update_params(pbf_ctx *ctx, uint32_t scene, const pbf_params *params);

/* One collision quad: corner B, edges E0 and E1 (BouncePointQuad's
 * arguments, compute_delta.cl:1) and the band of mask rows it samples:
 * texel row mask_y + (int)(s * mask_h) at quad parameter s along E0. */
typedef struct pbf_quad {
    float origin[3];
    float e0[3];
    float e1[3];
    int32_t mask_y, mask_h;
} pbf_quad;

#define PBF_MAX_QUADS 26

/* The reference's force plates (compute_delta.cl:1-42, applied after the
 * walls as at :169-182, with the surface mask that Simulation::LoadForceMasks
 * loads, src/Simulation.cpp:257-281): in the delta stage each quad in turn
 * moves a point that is inside its parallelogram, behind its plane offset by
 * edge_base + edge and within the force zone, onto that offset plane, where
 * the mask texel ((int)(t mask_w), mask_y + (int)(s mask_h)), clamped to the
 * image, is nonzero.  edge is the walls' noise edge; edge_base is the
 * reference's "1.0 +" (compute_delta.cl:166).  mask: mask_w x mask_rows
 * texels, row-major (the reference's CL_R / CL_UNSIGNED_INT32 image of the
 * PNG's red channel).  count = 0 removes the scene's quads (quads and mask may
 * then be NULL).  count <= PBF_MAX_QUADS; every quad needs finite entries
 * and non-parallel edges.  Replaces the scene's previous quads; waits for the
 * context's queued work. */
int pbf_set_collision_quads(pbf_ctx *ctx, uint32_t scene, const pbf_quad *quads, uint32_t count,
                            const uint32_t *mask, uint32_t mask_w, uint32_t mask_rows, float edge_base);

/* Simulation::Step (src/Simulation.cpp:554-621).  pause mirrors bPauseSim:
 * no gravity and no re-sort. */
int pbf_step(pbf_ctx *ctx, uint32_t pause);

/* steps x Simulation::Step, as Runner::run does per frame
 * (src/Runner.cpp:124-132).  With graphs on (pbf_set_graph), groups of four
 * substeps are replayed as one graph each, the rest one graph per substep;
 * the results equal pbf_step called `steps` times. */
int pbf_run(pbf_ctx *ctx, uint32_t steps, uint32_t pause);

/* pbf_run for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Runs `steps` substeps of one scene from
 * positions x4 and velocities v4 (v4 may be NULL for zero) and writes the
 * final state to x4_out and v4_out in the cell-sorted order pbf_get_state
 * gives; all four are DEVICE arrays of n float4.  No context is created and
 * nothing is allocated: every other buffer of the substep lives in one
 * static device array of the library (about 36 MB, held on each device the
 * library runs on until the process ends), which limits a call to
 * n <= 16384 particles and gridBufSize <= 65536 (PBF_E_INVALID beyond);
 * there are no collision quads.  The work is enqueued on `stream` (a
 * hipStream_t; NULL for the null stream) and the call returns without
 * waiting.  Calls share the array of their device: two calls whose work
 * overlaps on one device race.  Results equal pbf_create + pbf_set_state +
 * pbf_run + pbf_get_state bit for bit.  Errors: pbf_last_error(NULL). */
int pbf_run_noalloc(const pbf_params *params, const pbf_options *opt, const float *x4, const float *v4, uint32_t n,
                    uint32_t steps, float *x4_out, float *v4_out, void *stream);

/* Replays each substep (pbf_step) or group of four substeps (pbf_run) as an
 * instantiated HIP graph, captured on first use for every variant of the
 * buffer rotation.  Ignored while profiling. */
int pbf_set_graph(pbf_ctx *ctx, int enable);

/* Blocks until all queued work of the context is done. */
int pbf_sync(pbf_ctx *ctx);

/* State after the last step, in the current (cell-sorted) particle order.
 * x4: positions, w = |v| (the render VBO, update_velocities.cl:14);
 * v4: velocities; xs4: predicted positions, w = density (the render
 * texture, update_velocities.cl:16-17); rho, lambda: last solver
 * iteration; omega4: vorticity.  Any pointer may be NULL. */
int pbf_get_state(pbf_ctx *ctx, float *x4, float *v4, float *xs4, float *rho, float *lambda, float *omega4);

/* Structure of the last step: sorted cell keys, the permutation applied by
 * the sort (radixsort.cl:22-37), cells (uint32[2*gridBufSize],
 * update_cells.cl) and the friends list in the reference layout
 * (build_friends_list.cl:1-21: counts[5][n] then indices[5][50][n]).
 * Any pointer may be NULL. */
int pbf_get_structure(pbf_ctx *ctx, uint32_t *keys_sorted, uint32_t *perm, uint32_t *cells, uint32_t *friends);

/* Device pointers for zero-copy interop (the shared GL buffers of
 * src/Simulation.cpp:209-211 on the reference side). */
typedef struct pbf_dev_bufs {
    void *positions;   /* float4[n], w = |v|       */
    void *predicted;   /* float4[n], w = density   */
    void *velocities;  /* float4[n]                */
    void *density;     /* float[n]                 */
    void *lambda;      /* float[n]                 */
    void *omega;       /* float4[n]                */
    void *stream;      /* hipStream_t of the context */
    uint32_t n;
} pbf_dev_bufs;
int pbf_device_buffers(pbf_ctx *ctx, pbf_dev_bufs *out);

/* Per-stage timing, the OCLPerfMon trackers (src/OCLPerfMon.cpp:3-50):
 * same names, "_<iteration>" suffix, EMA weight 0.5.  Enabled with
 * pbf_set_profiling(ctx, 1); pbf_timings synchronises.  Each duration is
 * that of a pair of HIP events the kernel stamps at its start and end
 * (hipExtLaunchKernelGGL).  The extra tracker "eventTare" times an empty
 * kernel launched the same way once per profiled substep: the fixed part of
 * every launch's duration. */
typedef struct pbf_timing {
    char name[48];
    double last_ms;    /* last measured duration                    */
    double ema_ms;     /* total_time of OCLPerfMon (weight 0.5)      */
    double sum_ms;     /* accumulated since profiling was enabled    */
    uint64_t calls;
} pbf_timing;
int pbf_set_profiling(pbf_ctx *ctx, int enable);
int pbf_timings(pbf_ctx *ctx, pbf_timing *out, int cap, int *count);
int pbf_reset_timings(pbf_ctx *ctx);

/* Launch geometry.  Families: 0 friends build, 1 lambda, 2 delta,
 * 3 viscosity, 4 vorticity; sizes from {64, 128, 256, 512}.  The friends
 * build stages one workgroup's neighbourhood (about size/8 cells) in LDS, so
 * its size is also the number of cells per tile. */
#define PBF_FAMILIES 5
int pbf_set_workgroup_sizes(pbf_ctx *ctx, const uint32_t *sizes /* PBF_FAMILIES */);
/* Friends build variant: 1 stages each tile's 27-cell neighbourhood in LDS,
 * 0 reads candidates from global memory (L2).  Both give the same lists. */
int pbf_set_friends_mode(pbf_ctx *ctx, int staged);
/* Lanes per particle of the global-read friends build: 1 (one thread scans
 * a particle's 27 cells), 2, 4, 8 or 16 (each lane scans its share of the 27
 * cells; the lanes' lists are joined in the reference's scan order, so the
 * lists are the same), or 0 (the default): 16 lanes when the scene has at
 * most two 64-particle waves per CU (32,768 particles on 256 CUs), else 1.
 * More lanes shorten the build of scenes too small to fill the device; the
 * lanes build scans twice, so it is slower on larger scenes.  No
 * counterpart in the reference
 * (its buildFriendsList runs one work-item per particle,
 * assets/kernels/build_friends_list.cl:28).  PBF_E_INVALID for other values. */
int pbf_set_friends_lanes(pbf_ctx *ctx, uint32_t lanes);
/* Tiles per workgroup (the particles each thread handles in turn) of every
 * family: 1 launches one workgroup per tile of the block table; k > 1
 * launches about tiles / k workgroups; 0 launches a persistent grid, as many
 * workgroups as the device holds at once for the family's kernel and size.
 * Workgroups of a strided launch walk their XCD's run of tiles.  The
 * LDS-staged friends build always runs one tile per workgroup.  Results do
 * not depend on it.  Values up to 64. */
int pbf_set_tiles_per_workgroup(pbf_ctx *ctx, const uint32_t *tiles /* PBF_FAMILIES */);
/* Autotuner: times every candidate size for each family (and both friends
 * build variants), then 4 tiles per workgroup and the persistent grid at the
 * chosen sizes, and on scenes of at most four 64-particle waves per CU the
 * friends build with 1, 4 and 16 lanes per particle (unless lanes were set
 * with pbf_set_friends_lanes), over `steps` substeps from the current state
 * (which is restored afterwards).  A choice other than the default (256, global-read
 * friends build, one tile per workgroup) is kept only when it is more than
 * 3% faster.  Returns the chosen sizes in `chosen` (may be NULL; pbf_stats
 * reports the variant and the tiles per workgroup). */
int pbf_autotune(pbf_ctx *ctx, uint32_t steps, uint32_t *chosen /* PBF_FAMILIES */);

typedef struct pbf_stats_t {
    uint32_t particles, scenes, total_cells;
    uint32_t sort_passes, sort_tiles;
    uint32_t unstaged_tiles;      /* friends tiles that read global memory */
    uint32_t workgroup_size[PBF_FAMILIES];
    uint32_t friends_staged;      /* friends build variant (pbf_set_friends_mode) */
    uint32_t tiles_per_workgroup[PBF_FAMILIES]; /* pbf_set_tiles_per_workgroup */
    uint32_t friends_lanes;       /* lanes per particle the friends build runs with */
    uint32_t reserved[1];
} pbf_stats_t;
int pbf_stats(pbf_ctx *ctx, pbf_stats_t *out);

/* Accumulated PBF_FLAG_* bits and the number of dropped friends; reset
 * clears them.  Synchronises. */
int pbf_error_flags(pbf_ctx *ctx, uint32_t *flags, uint64_t *dropped_friends, int reset);

/* Device event log: the role of the reference's device logger
 * (assets/kernels/logging.cl, src/OCL_Logger.cpp:80-116), a ring in device
 * memory that kernels append to and the host reads after a step.  Each
 * message is {code, a, b, c}; the kernels log: */
#define PBF_LOG_FRIENDS_OVERFLOW 1u /* a = particle (current order), b = friends
                                       dropped, c = its packed circle counts
                                       (6 bits per circle)                      */
#define PBF_LOG_NAN 2u              /* a = first particle of a workgroup with a
                                       non-finite predicted position, b = how
                                       many in that workgroup, c = scene        */
#define PBF_LOG_SORT_TIMEOUT 3u     /* a = tile, b = digit, c = key shift        */
#define PBF_LOG_INTERNAL 4u         /* a = where (1 gather, 2 cells, 3 sort
                                       scatter), b = index, c = bad value       */
#define PBF_LOG_CAPACITY 1024u      /* messages the ring holds                   */

typedef struct pbf_log_msg {
    uint32_t code, a, b, c;
} pbf_log_msg;

/* Messages logged since the previous call, oldest first: up to cap into out
 * (*count receives how many; the rest stay for the next call).  *lost (may
 * be NULL) receives how many were overwritten before they were read (more
 * than PBF_LOG_CAPACITY between reads).  Synchronises. */
int pbf_read_log(pbf_ctx *ctx, pbf_log_msg *out, uint32_t cap, uint32_t *count, uint64_t *lost);

/* The r^2 at which the friend circle of build_friends_list.cl:75-76 reaches
 * 1, 2, 3, 4 for smoothing length h (+inf if never in range): the kernels
 * classify a friend by comparing its binary32 r^2 against these. */
int pbf_friend_circle_thresholds(float h, float *t4);

/* Plain HBM copy of `bytes` (float4 loads/stores, 2048 workgroups), `reps`
 * times; *gbps receives read+write bytes per second.  No counterpart in the
 * reference: it calibrates the roofline (sustained copy bandwidth) and the
 * rocprofv3 FETCH_SIZE / WRITE_SIZE counters (kernel k_copy_probe). */
int pbf_copy_probe(uint64_t bytes, uint32_t reps, double *gbps);


/* Compares, on the device, the square root and reciprocal the sweeps use
 * (hardware estimate plus one FMA correction) with IEEE sqrtf and 1.0f/x:
 * the square root for every binary32 input from 2^-96 (below it a pair
 * counts as coincident), the reciprocal for every normal input whose
 * reciprocal is normal.  Counts of mismatches and the first mismatching bit
 * pattern of each (0 if none).  No counterpart in the reference; it backs
 * the bit-for-bit parity of the sweeps with the IEEE oracle. */
int pbf_check_fast_math(uint64_t *sqrt_bad, uint64_t *rcp_bad, uint32_t *first_sqrt, uint32_t *first_rcp);

void pbf_destroy(pbf_ctx *ctx);

/* Message of the last failing call on ctx (or of the last pbf_create
 * failure when ctx is NULL). */
const char *pbf_last_error(const pbf_ctx *ctx);

const char *pbf_version(void);

#ifdef __cplusplus
}
#endif
#endif
