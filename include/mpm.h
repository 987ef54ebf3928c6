/*
 * mpm.h -- C ABI of libmpm355, an MI355X (gfx950) implementation of the
 * MLS-MPM fluid substep of WebGPU-Ocean (mls-mpm/mls-mpm.ts).
 *
 * The ABI is what a Node N-API addon (or any FFI) binds to replace the
 * reference's MLSMPMSimulator class.  Plain pointers and sizes only; every
 * function returns an mpm_status (0 = OK) and never throws or aborts across
 * the boundary.  mpm_last_error() returns a message for the calling thread.
 *
 * Threading: one simulator handle per host thread; no internal locking.  All
 * device work of a handle is ordered on that handle's HIP stream.
 *
 * Layouts (identical to the reference's GPU buffers):
 *   particle  80 B  position f32[3] @0, v f32[3] @16, C f32 3x3 column-major @32,
 *                   each column xyz + 4 B pad     (mls-mpm/p2g_1.wgsl:1-5)
 *   cell      16 B  vx, vy, vz, mass as i32 fixed point (mls-mpm/clearGrid.wgsl:1-6)
 *   posvel    32 B  position f32[3] @0, v f32[3] @16    (mls-mpm/copyPosition.wgsl:7-10)
 */
#ifndef MPM355_H
#define MPM355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 1

#define MPM_PARTICLE_STRUCT_SIZE 80 /* mlsmpmParticleStructSize, mls-mpm/mls-mpm.ts:10 */
#define MPM_CELL_STRUCT_SIZE 16     /* cellStructSize, mls-mpm/mls-mpm.ts:16 */
#define MPM_POSVEL_STRUCT_SIZE 32   /* main.ts:117 */

typedef enum {
    MPM_OK = 0,
    MPM_ERR_INVALID_ARGUMENT = 1,
    MPM_ERR_GRID_TOO_LARGE = 2, /* counterpart of the throw at mls-mpm/mls-mpm.ts:230-232 */
    MPM_ERR_OUT_OF_MEMORY = 3,
    MPM_ERR_DEVICE = 4,         /* a HIP runtime call failed */
    MPM_ERR_NO_DEVICE = 5,
    MPM_ERR_NOT_READY = 6,      /* the handle holds no scene: its last mpm_reset failed
                                   (everything was released; mpm_reset again) */
} mpm_status;

/* The reference's pipeline-override constants (mls-mpm/mls-mpm.ts:53-59). */
typedef struct {
    float stiffness;              /* 3.0  */
    float rest_density;           /* 4.0  */
    float dynamic_viscosity;      /* 0.1  */
    float dt;                     /* 0.2  */
    float fixed_point_multiplier; /* 1e7  */
} mpm_params;

/* The six reference kernels, in dispatch order (mls-mpm/mls-mpm.ts:248-265). */
typedef enum {
    MPM_STAGE_CLEAR_GRID = 0,   /* mls-mpm/clearGrid.wgsl   */
    MPM_STAGE_P2G_1 = 1,        /* mls-mpm/p2g_1.wgsl       */
    MPM_STAGE_P2G_2 = 2,        /* mls-mpm/p2g_2.wgsl       */
    MPM_STAGE_UPDATE_GRID = 3,  /* mls-mpm/updateGrid.wgsl  */
    MPM_STAGE_G2P = 4,          /* mls-mpm/g2p.wgsl         */
    MPM_STAGE_COPY_POSITION = 5,/* mls-mpm/copyPosition.wgsl */
    MPM_STAGE_COUNT = 6
} mpm_stage;

typedef struct mpm_sim mpm_sim;

/* Library version (MPM_ABI_VERSION). */
int mpm_abi_version(void);

/* Message describing the calling thread's most recent error ("" if none). */
const char *mpm_last_error(void);

/* Fills *out with the reference constants of mls-mpm/mls-mpm.ts:53-59. */
int mpm_default_params(mpm_params *out);

/* Construction from a particle buffer, box and parameters.
 * Replaces `new MLSMPMSimulator(particleBuffer, posvelBuffer, renderDiameter,
 * device)` (mls-mpm/mls-mpm.ts:42-193) followed by `reset` (:225-243):
 * particles = n records of 80 B on the host (may be NULL when n == 0);
 * init_box = the reference's initBoxSize; the real box starts equal to it.
 * params == NULL selects mpm_default_params.  Grid cells =
 * ceil(bx)*ceil(by)*ceil(bz) (:229); MPM_ERR_GRID_TOO_LARGE replaces the
 * 64^3 cap of :230-232 with the limit of 32-bit cell indices. */
int mpm_create(const mpm_params *params, const void *particles, uint32_t n,
               const float init_box[3], mpm_sim **out);

/* Releases the handle and its device memory. NULL is accepted. */
int mpm_destroy(mpm_sim *sim);

/* reset(numParticles, initBoxSize) (mls-mpm/mls-mpm.ts:225-243) with the
 * particle data supplied by the caller (the host side builds it, as
 * initDambreak does at :195-223).  Sets real box = init box. */
int mpm_reset(mpm_sim *sim, const void *particles, uint32_t n, const float init_box[3]);

/* changeBoxSize(realBoxSize) (mls-mpm/mls-mpm.ts:270-275): ordered on the
 * handle's stream after previously issued steps.  Components must be finite;
 * the grid strides stay those of the init box, as in the reference. */
int mpm_set_box(mpm_sim *sim, const float real_box[3]);

/* Advances `substeps` substeps; execute(commandEncoder) (mls-mpm.ts:245-268)
 * is substeps == 2.  Asynchronous on the handle's stream. */
int mpm_step(mpm_sim *sim, uint32_t substeps);

/* Runs one reference kernel exactly as dispatched at mls-mpm.ts:248-265
 * (test hook for per-stage bit-exact comparison).  Asynchronous. */
int mpm_run_stage(mpm_sim *sim, int stage);

/* Waits for all work issued on the handle. */
int mpm_synchronize(mpm_sim *sim);

/* Position readout: n records of 32-B PosVel (the output of copyPosition,
 * copyPosition.wgsl:15-21) into host memory.  Synchronizes. */
int mpm_read_posvel(mpm_sim *sim, void *dst, uint32_t n);

/* Device address of the N x 32-B PosVel buffer, for consumers that read it
 * on the GPU as the reference's renderer does (render/sphere.wgsl:28-33).
 * The address stays the same across mpm_step and mpm_set_box; mpm_reset and
 * mpm_destroy invalidate it.  Steps are asynchronous: a reader on another
 * stream waits for mpm_synchronize (or an event after the handle's work). */
int mpm_posvel_device_ptr(mpm_sim *sim, void **out);

/* Particle state in the reference's 80-B layout (read synchronizes). */
int mpm_read_particles(mpm_sim *sim, void *dst, uint32_t n);
int mpm_write_particles(mpm_sim *sim, const void *src, uint32_t n);

/* Grid accumulators, ncells x (vx, vy, vz, mass) i32 (read synchronizes).
 * After mpm_step on the default path the grid holds the last substep's raw
 * P2G accumulators (after p2g_1 and p2g_2, before updateGrid: the production
 * path keeps the updated velocities in a separate buffer, see
 * mpm_read_grid_velocity); after mpm_run_stage it holds that stage's output
 * exactly as the reference's cell buffer (mls-mpm/mls-mpm.ts:122-130). */
int mpm_read_grid(mpm_sim *sim, int32_t *dst, uint64_t ncells);
int mpm_write_grid(mpm_sim *sim, const int32_t *src, uint64_t ncells);

/* Grid velocities after updateGrid, decoded: ncells x (vx, vy, vz, 0) f32 =
 * f32(cell.v) / multiplier of the reference's cells after updateGrid.wgsl
 * (what g2p.wgsl:51-53 reads), as left by the last mpm_step on the default
 * path (read synchronizes).  Test hook for the production path's updateGrid;
 * undefined after mpm_run_stage. */
int mpm_read_grid_velocity(mpm_sim *sim, float *dst, uint64_t ncells);

/* numParticles and gridCount (mls-mpm/mls-mpm.ts:19-20). */
int mpm_get_counts(mpm_sim *sim, uint32_t *num_particles, uint64_t *grid_count);

/* Timing hooks (HIP events on the handle's stream, after a synchronize).
 * mpm_time_steps: runs `frames` x `substeps_per_frame` substeps; *ms = total.
 * mpm_time_stages: runs `substeps` substeps eagerly, timing every kernel of
 * the production path with events bound to the kernel's own start and end
 * (hipExtLaunchKernel), and every stage of the reference-shaped path with
 * events around it; stage_ms[i] = summed duration of slot i, calls[i] (may be
 * NULL) = its launches, names[i] (may be NULL) = a static name; *count = slots
 * used.  No reference counterpart (the reference only
 * times JS encoding, main.ts:209,278-279). */
int mpm_time_steps(mpm_sim *sim, uint32_t frames, uint32_t substeps_per_frame, float *ms);
int mpm_time_stages(mpm_sim *sim, uint32_t substeps, float *stage_ms, int *calls,
                    const char **names, int capacity, int *count);

/* Launch configuration of the production substep (no reference counterpart:
 * the reference fixes @workgroup_size(64) and one thread per item).
 *   path                 0 = binned tile-partial kernels (default),
 *                        1 = reference-shaped kernels (one launch per WGSL kernel;
 *                            single simulators only)
 *   particles_per_thread 1, 2 or 4: a P2G/G2P block takes a chunk of at most
 *                        256 x this particles of one tile
 *   grid_blocks          persistent blocks of the chunk kernels, 0 = 4096,
 *                        at most 2^20
 *   xcd_remap            1 = each XCD walks a contiguous range of chunks,
 *                        0 = runs of 16 consecutive chunks dealt round robin
 *                            over the XCDs (default)
 *   tile_edge            4 (default) or 8: P2G tiles of tile_edge^3 cells, whose
 *                        LDS windows are (tile_edge + 2)^3 cells
 * Results do not depend on these (integer adds commute; per-particle float
 * math is unchanged); only speed does.  mpm_backward runs its reverse
 * kernels on 4^3 tiles and restores the tuning afterwards. */
typedef struct {
    int path;
    int particles_per_thread;
    int grid_blocks;
    int xcd_remap;
    int tile_edge;
} mpm_tuning;

int mpm_get_tuning(mpm_sim *sim, mpm_tuning *out);
int mpm_set_tuning(mpm_sim *sim, const mpm_tuning *in);

/* Autotuner: times candidate tunings (tile_edge x particles_per_thread x
 * grid_blocks x xcd_remap) for `substeps_per_trial` substeps each on a snapshot of the
 * current state, restores the state and keeps the fastest.  No reference
 * counterpart. */
int mpm_autotune(mpm_sim *sim, uint32_t substeps_per_trial, mpm_tuning *chosen);

/* ---- Differentiable rollout (no reference counterpart: the reference has
 * no gradients).  Runs `substeps` substeps from the current state, keeping a
 * checkpoint every `checkpoint_every` substeps, evaluates
 *     loss = |centre of mass - target|^2 + ke_weight * mean |v|^2
 * on the final state, and back-propagates through a hand-written adjoint of
 * every substep (recomputing each segment from its checkpoint).  Fixed-point
 * encode/decode are treated as the identity.  Outputs (any may be NULL):
 * d loss / d initial velocity and position (n x 3 floats, caller order),
 * d loss / d stiffness and / d dynamic_viscosity.  Afterwards the simulator
 * holds the rollout's final state, as after mpm_step(sim, substeps). */
typedef struct {
    uint32_t substeps;
    uint32_t checkpoint_every;
    float target[3];
    float ke_weight;
} mpm_adjoint_config;

int mpm_backward(mpm_sim *sim, const mpm_adjoint_config *cfg, float *grad_v0, float *grad_x0,
                 double *loss, float *grad_stiffness, float *grad_viscosity);

/* ---- Allocation-free step.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Runs `substeps`
 * substeps of the reference-shaped kernels (clearGrid, p2g_1, p2g_2,
 * updateGrid, g2p, copyPosition; mls-mpm/mls-mpm.ts:248-265) on the caller's
 * device buffers of the current device: particles = n 80-B records, read and
 * written in place; posvel = n 32-B PosVel records, written (record order).
 * Results equal mpm_step's on the same state and boxes.  The grid and the
 * uniforms live in a library-owned static device array holding at most 64^3
 * cells (the reference's cell-buffer cap, mls-mpm/mls-mpm.ts:13-15): a larger
 * ceil(bx)*ceil(by)*ceil(bz) is rejected with MPM_ERR_GRID_TOO_LARGE.  No
 * device memory is allocated or freed.  stream: a hipStream_t (NULL = the
 * default stream); asynchronous.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  params NULL = mpm_default_params. */
#define MPM_NOALLOC_MAX_CELLS (64u * 64u * 64u)
int mpm_step_noalloc(const mpm_params *params, void *particles, void *posvel, uint32_t n,
                     const float init_box[3], const float real_box[3], uint32_t substeps,
                     void *stream);

/* ---- Batched scenes (no reference counterpart: the reference runs one scene).
 * Many independent scenes with their own particle counts, boxes and
 * parameters, packed into shared buffers and advanced by ONE launch per stage
 * (hipGraph-replayed substep).  Each scene's results equal a single
 * simulator's on the same inputs. */
typedef struct {
    const void *particles;  /* n x 80-B records (host) */
    uint32_t n;
    float init_box[3];
    mpm_params params;
} mpm_scene;

typedef struct mpm_batch mpm_batch;

int mpm_batch_create(const mpm_scene *scenes, uint32_t count, mpm_batch **out);
int mpm_batch_destroy(mpm_batch *batch);
int mpm_batch_step(mpm_batch *batch, uint32_t substeps);
int mpm_batch_set_box(mpm_batch *batch, uint32_t scene, const float real_box[3]);
int mpm_batch_synchronize(mpm_batch *batch);
int mpm_batch_get_counts(mpm_batch *batch, uint32_t scene, uint32_t *num_particles,
                         uint64_t *grid_count);
int mpm_batch_read_posvel(mpm_batch *batch, uint32_t scene, void *dst, uint32_t n);
int mpm_batch_read_particles(mpm_batch *batch, uint32_t scene, void *dst, uint32_t n);
int mpm_batch_get_tuning(mpm_batch *batch, mpm_tuning *out);
int mpm_batch_set_tuning(mpm_batch *batch, const mpm_tuning *in);
int mpm_batch_autotune(mpm_batch *batch, uint32_t substeps_per_trial, mpm_tuning *chosen);
int mpm_batch_time_steps(mpm_batch *batch, uint32_t substeps, float *ms);
int mpm_batch_time_stages(mpm_batch *batch, uint32_t substeps, float *stage_ms, int *calls,
                          const char **names, int capacity, int *count);

#ifdef __cplusplus
}
#endif
#endif
