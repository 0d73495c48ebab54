/*
 * pbf_oracle.h -- CPU restatement of the reference PBF substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker; the product library never links,
 * loads or falls back to it.
 *
 * Every stage restates one kernel of the reference (WorkingXXX/
 * position-based-fluids, assets/kernels/<stage>.cl) driven in the order of
 * src/Simulation.cpp:554-621, with IEEE single precision, no implicit FMA
 * contraction (built with -ffp-contract=off; explicit fmaf only where the GPU
 * kernels fuse the same operations) and the operation order of the
 * reference's expressions.  Where the reference relies on -cl-fast-relaxed-math or has
 * undefined behaviour, the choice made here is stated at the function; one
 * applies to every friend sweep: the vector r / |r| (compute_scaling.cl:83,
 * compute_delta.cl:144, apply_viscosity.cl:65, apply_vorticity.cl) is
 * evaluated as r * (1/|r|), one correctly rounded reciprocal per pair, which
 * the relaxed-math build of the reference also permits; at r = 0 (coincident
 * particles) the reference's 0/0 gives NaN, which then spreads through every
 * later step, and here the direction is 0 (the pair still counts in the
 * density; its gradient terms vanish).
 *
 * Data layout (all little-endian, particle-major):
 *   float4 arrays   : float[4*n]   (x, y, z, w)
 *   friends list    : uint32[n*5*(1+50)] in the reference layout
 *                     (build_friends_list.cl:1-21): counts[5][n] followed by
 *                     indices[5][50][n].
 *   cells           : uint32[2*grid_buf_size], start/end (inclusive) per hash,
 *                     0xFFFFFFFF = empty (update_cells.cl, reset_grid.cl).
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBFO_CIRCLES 5          /* Params.friendsCircles, ParamUtils.cpp:98     */
#define PBFO_PER_CIRCLE 50      /* Params.particlesPerCircle, ParamUtils.cpp:99 */
#define PBFO_EMPTY 0xFFFFFFFFu  /* END_OF_CELL_LIST, Simulation.cpp:170         */

typedef struct pbfo_cfg {
    uint32_t n;              /* Params.particleCount                               */
    float h;                 /* Params.h (smoothing length = cell size)            */
    float dt;                /* Params.timeStep                                    */
    uint32_t iters;          /* Params.simIterations                               */
    float rest_density;      /* Params.restDensity (unit-mass units)               */
    float epsilon;           /* Params.epsilon (relaxation)                        */
    float gravity[3];        /* acceleration; the reference uses (0,-garvity,0)    */
    float vorticity;         /* Params.vorticityFactor                             */
    float viscosity;         /* Params.viscosityFactor                             */
    float st_k;              /* Params.surfaceTenstionK                            */
    float st_dist;           /* Params.surfaceTenstionDist                         */
    uint32_t grid_buf_size;  /* Params.gridBufSize                                 */
    float box_lo[3];         /* Params.xMin,yMin,zMin: box walls                   */
    float box_hi[3];         /* Params.xMax,yMax,zMax                              */
    uint32_t walls;          /* 1: six box planes in the delta stage, 0: none      */
    /* Collision quads (BouncePointQuad, compute_delta.cl:1-42, chained as
     * :169-182), applied after the walls; quad_count <= 26. */
    const struct pbfo_quad *quads;
    uint32_t quad_count;
    const uint32_t *mask;    /* mask_w x mask_rows texels; a quad moves a point  */
    uint32_t mask_w;         /* only where its texel is nonzero (the red channel */
    uint32_t mask_rows;      /* of Scene_fp_mask.png, Simulation.cpp:257-281)    */
    float edge_base;         /* the "1.0 +" of compute_delta.cl:166 (the
                                reference scene's scale); walls use none      */
} pbfo_cfg;

/* One collision quad: corner B, edges E0, E1 and its band of mask rows. */
typedef struct pbfo_quad {
    float b[3], e0[3], e1[3];
    int32_t mask_y, mask_h;
} pbfo_quad;

/* Kernel normalisation constants as the reference bakes them into the
 * program: computed in double, printed with 6 significant digits, parsed as a
 * float literal (Simulation.cpp:167,179-180). */
void pbfo_kernel_factors(float h, float *poly6, float *spiky);

/* Morton hash (utilities.cl:69-87) and cell coordinate (radixsort.cl:12). */
uint32_t pbfo_expand_bits(int32_t x);
uint32_t pbfo_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t grid_buf_size);
int32_t pbfo_cell_coord(float q);
float pbfo_sinf(float x);

/* Stage restatements. */
void pbfo_predict(const pbfo_cfg *c, uint32_t pause, const float *x, float *v, float *xs);
void pbfo_keys(const pbfo_cfg *c, const float *xs, uint32_t *keys);
void pbfo_stable_sort(const uint32_t *keys, uint32_t n, uint32_t *keys_sorted, uint32_t *perm);
void pbfo_gather(const uint32_t *perm, uint32_t n, const float *src, float *dst);
void pbfo_cells(const uint32_t *keys_sorted, uint32_t n, uint32_t *cells);
void pbfo_friends(const pbfo_cfg *c, const float *xs, const uint32_t *cells, uint32_t *friends);
void pbfo_scaling(const pbfo_cfg *c, const float *xs, const uint32_t *friends, float *rho, float *lambda);
void pbfo_pack(uint32_t n, const float *src4, const float *w, float *dst4);
void pbfo_delta(const pbfo_cfg *c, const float *x, const float *xs, const uint32_t *friends, float *delta);
void pbfo_delta_hits(const pbfo_cfg *c, const float *x, const float *xs, const uint32_t *friends, float *delta,
                     uint32_t *hits);
void pbfo_update_predicted(uint32_t n, const float *xs, const float *delta, float *xs_out);
void pbfo_update_velocities(const pbfo_cfg *c, float *x, const float *xs, float *v);
void pbfo_viscosity(const pbfo_cfg *c, const float *xs, const float *v_in, const uint32_t *friends,
                    float *v_out, float *omega);
void pbfo_vorticity(const pbfo_cfg *c, const float *xs, float *v, const float *omega, const uint32_t *friends);

/* Truncation rule of the friend sweeps, the literal fp32 test of
 * compute_scaling.cl:50 / compute_delta.cl:100: nonzero when circle is skipped. */
int pbfo_skip_circle(int processed, int total, float threshold);

/* Whole-substep driver following Simulation::Step (Simulation.cpp:554-621). */
typedef struct pbfo_sim pbfo_sim;
pbfo_sim *pbfo_create(const pbfo_cfg *c);
void pbfo_destroy(pbfo_sim *s);
void pbfo_set_state(pbfo_sim *s, const float *x, const float *v);
void pbfo_step(pbfo_sim *s, uint32_t pause);
/* which: 0 x, 1 v, 2 xs (w = rho), 3 rho, 4 lambda, 5 omega, 6 keys_sorted,
 * 7 perm, 8 cells, 9 friends, 10 keys (unsorted, this step) */
void *pbfo_buffer(pbfo_sim *s, int which);
/* Each later step writes, per solver iteration, which planes moved each
 * particle (iters x n words, bits as pbfo_delta_hits); NULL stops it. */
void pbfo_record_walls(pbfo_sim *s, uint32_t *hits);
int pbfo_threads(void);

#ifdef __cplusplus
}
#endif
#endif
