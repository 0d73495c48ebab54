/* pbf_adjoint.h -- reverse-mode adjoint of the PBF substep (C ABI).
 *
 * The reference has no differentiation; these entry points have no
 * counterpart there.  They differentiate a rollout of `substeps` calls of
 * the substep that pbf_step runs (src/Simulation.cpp:554-621 with the kernels
 * of assets/kernels) with respect to the initial positions and velocities and
 * to each scene's XSPH factor, vorticity factor and gravity vector.
 *
 * Frozen structure.  The sort order, the friend lists and the branches the
 * walls took are integer/branch structure: the adjoint differentiates the
 * substep with them held at the values the forward produced (they are
 * recomputed bit for bit from checkpoints, never approximated).  Everything
 * else -- the lambda/delta-p iterations with s_corr, the walls' edge offset,
 * the velocity update, XSPH and vorticity confinement -- is differentiated
 * exactly, in binary32.
 *
 * Checkpointing.  The forward keeps the particle state every `interval`
 * substeps; the backward replays each interval from its checkpoint,
 * recording the solver iterates, and walks it in reverse.  The last interval
 * is recorded by the forward itself.
 *
 * All array arguments of pbf_adjoint_forward/backward are DEVICE pointers
 * (float4 per particle, w ignored on input and written as 0 for gradients),
 * ordered by particle identity: entry i is the particle that entry i of x0
 * was.  Both calls return after their work has finished on the device.
 */
#ifndef PBF_ADJOINT_H
#define PBF_ADJOINT_H

#include <stdint.h>

#include "pbf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbf_adjoint pbf_adjoint;

/* Coefficients the adjoint differentiates, per scene, in the units of the
 * reference Parameters: viscosityFactor (apply_viscosity.cl:81),
 * vorticityFactor (apply_vorticity.cl:82) and the gravity vector
 * (predict_positions.cl:13 applies (0, -garvity, 0)).  Setting them switches
 * the scene to the explicit gravity vector. */
int pbf_set_scene_coefficients(pbf_ctx* ctx, uint32_t scene, float viscosity, float vorticity,
                               const float gravity[3]);
/* out5 = {viscosity, vorticity, g0, g1, g2} */
int pbf_get_scene_coefficients(pbf_ctx* ctx, uint32_t scene, float* out5);

/* Device bytes of one recorded substep and of one checkpoint. */
int pbf_adjoint_tape_bytes(pbf_ctx* ctx, uint64_t* per_substep, uint64_t* per_checkpoint);

/* Checkpoint interval for `substeps` substeps under a device-memory budget.
 * The backward replays every interval but the last (which the forward
 * records), so the plan takes the fewest intervals whose recorded substeps
 * plus checkpoints fit, and among those the one with the longest last
 * interval; interval = substeps replays nothing. */
int pbf_adjoint_plan(pbf_ctx* ctx, uint32_t substeps, uint64_t budget_bytes, uint32_t* interval);

/* A differentiable rollout of `substeps` substeps of ctx (which it then
 * drives; one adjoint per context at a time). */
int pbf_adjoint_create(pbf_ctx* ctx, uint32_t substeps, uint32_t interval, pbf_adjoint** out);

/* Runs the rollout from (x0, v0): ctx ends in the final state; xT, vT (may be
 * NULL) receive it. */
int pbf_adjoint_forward(pbf_adjoint* a, const float* x0, const float* v0, float* xT, float* vT);

/* Given dL/dxT and dL/dvT (either may be NULL for zero), writes dL/dx0,
 * dL/dv0 (either may be NULL) and, if gcoef is not NULL, per scene
 * {dL/dviscosity, dL/dvorticity, dL/dg0, dL/dg1, dL/dg2} (5 floats per scene,
 * device memory).  Requires a preceding pbf_adjoint_forward; may be called
 * again for another loss.  Leaves ctx's particles at an intermediate
 * substep.  Substeps that dropped friends at the 50-per-circle cap or visited
 * a hash bucket twice are differentiated too (their friend lists are
 * transposed by a search of the neighbour cells).  Fails with PBF_E_STATE
 * only when a particle is listed by more than 252 others (its transposed
 * list does not fit its slot: a dense cluster at the cap), or when a listed
 * friend lies outside the particle's 27 neighbour cells (possible only at
 * the rounding of a cell boundary), rather than differentiate a different
 * function. */
int pbf_adjoint_backward(pbf_adjoint* a, const float* gxT, const float* gvT, float* gx0, float* gv0,
                         float* gcoef);

/* Scenes whose gradients from the last pbf_adjoint_backward exceed binary32.
 * The sensitivities of a long rollout can grow geometrically with the
 * horizon; the backward keeps each scene's adjoint state as true * 2^e (exact
 * power-of-two rescaling), so it stays finite, but a gradient whose true
 * value is beyond binary32 is returned as +-inf.  *scenes receives how many
 * scenes had one; per_scene (may be NULL; one word per scene) receives bit 0
 * when a coefficient gradient overflowed and bit 1 when a dL/dx0 or dL/dv0
 * component did.  No counterpart in the reference.  PBF_E_STATE before a
 * backward. */
int pbf_adjoint_overflow(pbf_adjoint* a, uint32_t* scenes, uint32_t* per_scene);

/* The gradients of the last pbf_adjoint_backward as mantissas with one binary
 * exponent per scene, finite where their true values exceed binary32: for
 * the particles and coefficients of scene s, dL/dx0 = gx0[i] * 2^expo[s],
 * dL/dv0 = gv0[i] * 2^expo[s], dL/dcoef = gcoef[5s + k] * 2^expo[s].  gx0,
 * gv0, gcoef are device arrays laid out as in pbf_adjoint_backward, expo a
 * host array of one int32 per scene; any may be NULL.  The largest x0 or v0
 * mantissa component of a scene is in [2^-17, 2^16) unless all are zero.  No counterpart in the reference.  PBF_E_STATE before a
 * backward. */
int pbf_adjoint_scaled_gradients(pbf_adjoint* a, float* gx0, float* gv0, float* gcoef, int32_t* expo);

/* Milliseconds of the last forward and backward (host clock around the
 * device work), and of the replays inside the last backward. */
int pbf_adjoint_times(pbf_adjoint* a, double* forward_ms, double* backward_ms, double* replay_ms);

void pbf_adjoint_destroy(pbf_adjoint* a);

#ifdef __cplusplus
}
#endif
#endif
