/*
 * sph.h -- C ABI of the SPH fluid step in libmpm355: an MI355X (gfx950)
 * implementation of WebGPU-Ocean's SPHSimulator (sph/sph.ts).
 *
 * What a Node N-API addon (or any FFI) binds to replace the reference's
 * SPHSimulator class.  Plain pointers and sizes only; every function returns
 * an sph_status (0 = OK) and never throws or aborts across the boundary;
 * sph_last_error() returns a message for the calling thread.
 *
 * Threading: one handle per host thread; all device work of a handle is
 * ordered on that handle's HIP stream.
 *
 * Layouts (identical to the reference's GPU buffers):
 *   particle  64 B  position f32[3] @0, v f32[3] @16, force f32[3] @32,
 *                   density f32 @44, nearDensity f32 @48, 12 B pad
 *                   (sph/density.wgsl:1-7, sph/sph.ts:13,348-354)
 *   posvel    32 B  position f32[3] @0, v f32[3] @16 (sph/copyPosition.wgsl:9-12)
 *
 * Neighbour order.  The reference ranks particles within a cell by atomicAdd
 * (sph/grid/gridBuild.wgsl:58) and places rank r at the cell's end - r - 1
 * (sph/grid/reorderParticles.wgsl:57), so the order of its float sums over
 * neighbours depends on how its GPU scheduled the atomics.  This library
 * fixes that order to the one an in-order execution gives -- within a cell,
 * particles by decreasing index -- so results are deterministic and
 * bit-reproducible against a CPU evaluation.
 */
#ifndef SPH355_H
#define SPH355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_PARTICLE_STRUCT_SIZE 64 /* sphParticleStructSize, sph/sph.ts:13 */
#define SPH_POSVEL_STRUCT_SIZE 32   /* main.ts:117 */
#define SPH_SUBSTEPS_PER_FRAME 2    /* sph/sph.ts:303 */

typedef enum {
    SPH_OK = 0,
    SPH_ERR_INVALID_ARGUMENT = 1,
    SPH_ERR_GRID_TOO_LARGE = 2, /* the neighbour grid of the kernel radius exceeds 2^28 cells */
    SPH_ERR_OUT_OF_MEMORY = 3,
    SPH_ERR_DEVICE = 4,         /* a HIP runtime call failed */
    SPH_ERR_NO_DEVICE = 5,
    SPH_ERR_NOT_READY = 6,      /* the handle holds no scene: its last sph_reset failed */
} sph_status;

/* Host constants of sph/sph.ts:41,72-77 (JavaScript numbers, hence double;
 * they reach the kernels rounded to f32 as the reference's uniform buffers
 * hold them, :149-174).  The neighbour grid (:58-70) covers the half box 2.0
 * per axis plus a 2-cell margin, with cell size = kernel_radius. */
typedef struct {
    double kernel_radius;  /* 0.07  */
    double mass;           /* 1     */
    double dt;             /* 0.006 */
    double stiffness;      /* 20    */
    double near_stiffness; /* 1     */
    double rest_density;   /* 15000 */
    double viscosity;      /* 100   */
} sph_params;

/* The kernels of one substep of this library, in launch order (test and
 * timing hook).  The reference's nine dispatches (sph/sph.ts:304-332) map to
 * them as: gridClear -> folded into SPH_STAGE_SCAN (which clears the counts
 * it consumed); gridBuild -> COUNT; prefix sum -> SCAN_REDUCE + SCAN;
 * reorderParticles -> SCATTER + SORT_GATHER; density -> DENSITY (which also
 * writes the sorted copy the second reorderParticles makes); force,
 * integrate and copyPosition -> FORCE. */
typedef enum {
    SPH_STAGE_COUNT = 0,
    SPH_STAGE_SCAN_REDUCE = 1,
    SPH_STAGE_SCAN = 2,
    SPH_STAGE_SCATTER = 3,
    SPH_STAGE_SORT_GATHER = 4,
    SPH_STAGE_DENSITY = 5,
    SPH_STAGE_FORCE = 6,
    SPH_STAGE_NUM = 7
} sph_stage;

typedef struct sph_sim sph_sim;

/* Message describing the calling thread's most recent SPH error ("" if none). */
const char *sph_last_error(void);

/* Fills *out with the reference constants of sph/sph.ts:41,72-77. */
int sph_default_params(sph_params *out);

/* Construction + reset: replaces `new SPHSimulator(particleBuffer,
 * posvelBuffer, renderDiameter, device)` (sph/sph.ts:47-279) followed by
 * `reset(numParticles, initHalfBoxSize)` (:281-299) with the particle data
 * supplied by the caller (the host side builds it as initDambreak does,
 * :338-363).  particles = n records of 64 B on the host (may be NULL when
 * n == 0); init_half_box = initHalfBoxSize, which is also the starting real
 * box (:290-292).  params == NULL selects sph_default_params. */
int sph_create(const sph_params *params, const void *particles, uint32_t n,
               const float init_half_box[3], sph_sim **out);

/* Releases the handle and its device memory.  NULL is accepted. */
int sph_destroy(sph_sim *sim);

/* reset(numParticles, initHalfBoxSize) (sph/sph.ts:281-299). */
int sph_reset(sph_sim *sim, const void *particles, uint32_t n, const float init_half_box[3]);

/* changeBoxSize(realBoxSize) (sph/sph.ts:365-370): the half extents the
 * wall forces of integrate.wgsl:41-61 push against; ordered after
 * previously issued steps.  Components must be finite. */
int sph_set_box(sph_sim *sim, const float real_half_box[3]);

/* Advances `substeps` substeps; execute(commandEncoder) (sph/sph.ts:301-336)
 * is substeps == 2.  Asynchronous on the handle's stream. */
int sph_step(sph_sim *sim, uint32_t substeps);

/* Runs one kernel of the substep (sph_stage), for per-stage testing. */
int sph_run_stage(sph_sim *sim, int stage);

/* Waits for all work issued on the handle. */
int sph_synchronize(sph_sim *sim);

/* Position readout: n records of 32-B PosVel (copyPosition.wgsl:34-38) into
 * host memory.  Synchronizes. */
int sph_read_posvel(sph_sim *sim, void *dst, uint32_t n);

/* Device address of the n x 32-B PosVel buffer (what the reference's
 * renderer reads, render/sphere.wgsl:28-33); valid until sph_reset or
 * sph_destroy. */
int sph_posvel_device_ptr(sph_sim *sim, void **out);

/* Particle state in the reference's 64-B layout (read synchronizes). */
int sph_read_particles(sph_sim *sim, void *dst, uint32_t n);
int sph_write_particles(sph_sim *sim, const void *src, uint32_t n);

/* Neighbour grid of the last binning (read synchronizes):
 * sph_read_cell_start: gridCount + 1 entries, the exclusive prefix sum of
 * the per-cell particle counts (the reference's cellParticleCountBuffer after
 * its prefix sum, sph/sph.ts:179-183,310-313); sph_read_sorted_order: the
 * particle index at each of the `count` sorted slots (the reference's
 * targetParticles buffer holds those particles' records, :184-188). */
int sph_read_cell_start(sph_sim *sim, uint32_t *dst, uint64_t count);
int sph_read_sorted_order(sph_sim *sim, uint32_t *dst, uint32_t count);

/* numParticles and gridCount (sph/sph.ts:42-43,69). */
int sph_get_counts(sph_sim *sim, uint32_t *num_particles, uint64_t *grid_count);

/* Timing hooks (HIP events on the handle's stream; no reference
 * counterpart).  sph_time_steps: `substeps` graph-replayed substeps, *ms =
 * total.  sph_time_stages: `substeps` eager substeps with events bound to
 * each kernel's own start and end; stage_ms[s] = summed duration of stage s
 * (SPH_STAGE_NUM entries). */
int sph_time_steps(sph_sim *sim, uint32_t substeps, float *ms);
int sph_time_stages(sph_sim *sim, uint32_t substeps, float *stage_ms);

#ifdef __cplusplus
}
#endif
#endif
