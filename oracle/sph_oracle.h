/*
 * sph_oracle.h -- CPU restatement of WebGPU-Ocean's SPH step (sph/sph.ts,
 * the WGSL kernels under sph/).
 *
 * TEST INFRASTRUCTURE ONLY.  The parity checker of the HIP library's SPH
 * path, never a product path: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * Layouts follow the reference:
 *   Particle (64 B): position vec3f @0, v vec3f @16, force vec3f @32,
 *                    density f32 @44, nearDensity f32 @48, 12 B pad
 *                    -- sph/density.wgsl:1-7, sph/sph.ts:13,348-354
 *   PosVel   (32 B): position vec3f @0, v vec3f @16 -- sph/copyPosition.wgsl:9-12
 *   Environment / SPHParams -- sph/sph.ts:129-175
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float position[3]; float _pad0;
    float v[3];        float _pad1;
    float force[3];
    float density;
    float nearDensity;
    float _pad2[3];
} sph_oracle_particle;

typedef struct { float position[3]; float _pad0; float v[3]; float _pad1; } sph_oracle_posvel;

/* Host-side constants of sph/sph.ts:41,72-77 as JavaScript numbers. */
typedef struct {
    double kernel_radius;   /* 0.07  */
    double mass;            /* 1     */
    double dt;              /* 0.006 */
    double stiffness;       /* 20    */
    double near_stiffness;  /* 1     */
    double rest_density;    /* 15000 */
    double viscosity;       /* 100   */
} sph_oracle_params;

/* Environment uniform (sph/sph.ts:129-147). */
typedef struct {
    int32_t xGrids, yGrids, zGrids;
    float cellSize, xHalf, yHalf, zHalf, offset;
} sph_oracle_env;

/* SPHParams uniform without n (sph/sph.ts:149-174). */
typedef struct {
    float mass, kernelRadius, kernelRadiusPow2, kernelRadiusPow5, kernelRadiusPow6,
        kernelRadiusPow9, dt, stiffness, nearStiffness, restDensity, viscosity;
} sph_oracle_uniforms;

/* OpenMP threads of the per-particle loops (density, force, integrate);
 * results do not depend on it.  Default 1. */
void sph_oracle_set_threads(int n);

void sph_oracle_setup(const sph_oracle_params *p, sph_oracle_env *env, sph_oracle_uniforms *u);
uint64_t sph_oracle_grid_count(const sph_oracle_env *env);

void sph_oracle_cell_position(const sph_oracle_env *env, const float pos[3], int32_t c[3]);
int32_t sph_oracle_cell_id(const sph_oracle_env *env, const float pos[3]);

float sph_oracle_density_kernel(const sph_oracle_uniforms *u, float r);
float sph_oracle_near_density_kernel(const sph_oracle_uniforms *u, float r);
float sph_oracle_density_gradient(const sph_oracle_uniforms *u, float r);
float sph_oracle_near_density_gradient(const sph_oracle_uniforms *u, float r);
float sph_oracle_viscosity_laplacian(const sph_oracle_uniforms *u, float r);

/* Counting sort by cell: start = gridCount + 1 exclusive prefix entries,
 * order[k] / sorted[k] (sorted may be NULL) the k-th binned particle; within
 * a cell, particles appear by decreasing index.  Returns the number binned. */
uint32_t sph_oracle_bin(const sph_oracle_particle *p, uint32_t n, const sph_oracle_env *env,
                        uint32_t *start, uint32_t *order, sph_oracle_particle *sorted);

void sph_oracle_density(sph_oracle_particle *p, uint32_t n, const sph_oracle_particle *sorted,
                        const uint32_t *start, const sph_oracle_env *env,
                        const sph_oracle_uniforms *u);
void sph_oracle_force(sph_oracle_particle *p, uint32_t n, const sph_oracle_particle *sorted,
                      const uint32_t *start, const sph_oracle_env *env,
                      const sph_oracle_uniforms *u);
void sph_oracle_integrate(sph_oracle_particle *p, uint32_t n, const float real_half[3],
                          const sph_oracle_uniforms *u);

/* One substep of sph/sph.ts:304-332; posvel (may be NULL) receives copyPosition. */
void sph_oracle_substep(sph_oracle_particle *p, uint32_t n, const float real_half[3],
                        const sph_oracle_env *env, const sph_oracle_uniforms *u,
                        sph_oracle_posvel *posvel);

#ifdef __cplusplus
}
#endif
#endif
