/*
 * mpm_oracle.h -- CPU restatement of WebGPU-Ocean's MLS-MPM substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP
 * library, never a product path: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * Layouts follow the reference exactly:
 *   Particle (80 B): position vec3f @0, v vec3f @16, C mat3x3f @32
 *                    (column-major, each column xyz + 4 B pad)
 *                    -- mls-mpm/p2g_1.wgsl:1-5, mls-mpm/mls-mpm.ts:10,205-209
 *   Cell     (16 B): vx, vy, vz, mass as i32 fixed point
 *                    -- mls-mpm/clearGrid.wgsl:1-6
 *   PosVel   (32 B): position vec3f @0, v vec3f @16  -- copyPosition.wgsl:7-10
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float position[3]; float _pad0;
    float v[3];        float _pad1;
    float C[3][4];     /* C[col][row], row 3 is padding */
} oracle_particle;

typedef struct { int32_t vx, vy, vz, mass; } oracle_cell;

typedef struct { float position[3]; float _pad0; float v[3]; float _pad1; } oracle_posvel;

typedef struct {
    float stiffness;
    float rest_density;
    float dynamic_viscosity;
    float dt;
    float fixed_point_multiplier;
} oracle_params;

/* Grid size rule: ceil(bx)*ceil(by)*ceil(bz) -- mls-mpm/mls-mpm.ts:229 */
uint64_t oracle_grid_count(const float init_box[3]);

/* The six reference kernels, each over exactly n particles / g cells. */
void oracle_clear_grid(oracle_cell *cells, uint64_t g);
void oracle_p2g_1(const oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                  const float init_box[3], const oracle_params *prm);
void oracle_p2g_2(const oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                  const float init_box[3], const oracle_params *prm);
void oracle_update_grid(oracle_cell *cells, uint64_t g, const float real_box[3],
                        const float init_box[3], const oracle_params *prm);
void oracle_g2p(oracle_particle *p, uint32_t n, const oracle_cell *cells, uint64_t g,
                const float real_box[3], const float init_box[3], const oracle_params *prm);
void oracle_copy_position(const oracle_particle *p, uint32_t n, oracle_posvel *out);

/* One substep: the six kernels in the order of mls-mpm/mls-mpm.ts:247-266. */
void oracle_substep(oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                    oracle_posvel *posvel, const float real_box[3], const float init_box[3],
                    const oracle_params *prm);

/* Same substep with OpenMP over particles/cells and relaxed atomic integer
 * adds in P2G.  Integer grid results equal oracle_substep's (adds commute);
 * used as the timed CPU baseline.  nthreads <= 0 means the OpenMP default. */
void oracle_substep_omp(oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                        oracle_posvel *posvel, const float real_box[3], const float init_box[3],
                        const oracle_params *prm, int nthreads);

/* FP64, unquantized rollout of `substeps` substeps (real box = init box);
 * returns |com - target|^2 + ke_weight * mean|v|^2 of the final state.
 * x0 and v0 (n x 3 each, may be NULL) override the records' positions and
 * velocities; final_state (may be NULL) receives the final particles as 80-B
 * f32 records. */
double oracle_rollout_loss_f64(const float *particles, const double *x0, const double *v0,
                               uint32_t n, const float init_box[3], double stiffness,
                               double viscosity,
                               double rest_density, double dt, uint32_t substeps,
                               const double target[3], double ke_weight, float *final_state);

/* Scalar helpers exported so tests can pin them directly. */
int32_t oracle_encode(float f, float mult);
float   oracle_decode(int32_t i, float mult);
int32_t oracle_f32_to_i32(float f);
float   oracle_pow5(float r);
void    oracle_weights(float x, float w[3]);

#ifdef __cplusplus
}
#endif
#endif
