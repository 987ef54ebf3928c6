/*
 * sph_oracle.c -- CPU restatement of WebGPU-Ocean's SPH step (sph/sph.ts,
 * the WGSL kernels under sph/).  TEST INFRASTRUCTURE ONLY: the parity checker of the HIP
 * library's SPH path; never a product path.
 *
 * Every float operation is one IEEE binary32 operation in the order the WGSL
 * writes it (compiled with -ffp-contract=off).  Where WGSL leaves the order
 * open, this file fixes it, and the HIP kernels follow the same choice:
 *   - dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z;
 *   - normalize(v) = v / sqrt(dot(v, v)) (IEEE sqrt and division);
 *   - abstract-float constant expressions (64. * 3.1415926535) fold in double
 *     and round to f32 once, as WGSL's AbstractFloat rules do;
 *   - the order of particles within a cell.  The reference takes each
 *     particle's rank in its cell from an atomicAdd (sph/grid/gridBuild.wgsl:
 *     58) and places rank r at end - r - 1 (reorderParticles.wgsl:57), so
 *     its order -- and every float sum over neighbours -- depends on the
 *     order the GPU happened to run the atomics.  Here the atomics are taken
 *     in particle-index order (what an in-order execution gives): within a
 *     cell, particles are sorted by decreasing index.
 * Particles step exactly n (the reference writes n into the u32 field as the
 * bits of a float, sph.ts:293-296, and so also steps stale "ghost" records
 * up to its buffer's size).  Linear cell ids outside
 * [0, gridCount) are not binned (the reference only tests the upper bound,
 * gridBuild.wgsl:57, leaving negative ids to WebGPU's robust access).
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_threads = 1; /* OpenMP threads of the per-particle loops */

void sph_oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* sph/sph.ts:58-77 constants, :129-174 uniform values */
void sph_oracle_setup(const sph_oracle_params *p, sph_oracle_env *env,
                      sph_oracle_uniforms *u)
{
    const double h = p->kernel_radius;
    const double cell = 1.0 * h;
    const double half = 2.0; /* xHalfMax = yHalfMax = zHalfMax (sph.ts:58-60) */
    const double len = 2.0 * half;
    const double sentinel = 4 * cell;
    const int grids = (int)ceil((len + sentinel) / cell);
    env->xGrids = env->yGrids = env->zGrids = grids;
    env->cellSize = (float)cell;
    env->xHalf = env->yHalf = env->zHalf = (float)half;
    env->offset = (float)(sentinel / 2);
    u->mass = p->mass;
    u->kernelRadius = (float)h;
    u->kernelRadiusPow2 = (float)pow(h, 2);
    u->kernelRadiusPow5 = (float)pow(h, 5);
    u->kernelRadiusPow6 = (float)pow(h, 6);
    u->kernelRadiusPow9 = (float)pow(h, 9);
    u->dt = p->dt;
    u->stiffness = p->stiffness;
    u->nearStiffness = p->near_stiffness;
    u->restDensity = p->rest_density;
    u->viscosity = p->viscosity;
}

uint64_t sph_oracle_grid_count(const sph_oracle_env *env)
{
    return (uint64_t)env->xGrids * (uint64_t)env->yGrids * (uint64_t)env->zGrids;
}

static int32_t f2i(float f)  /* i32(f32): truncation, saturating, NaN -> 0 */
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

/* cellPosition -- density.wgsl:53-58 (gridBuild.wgsl:41-48 for the id) */
void sph_oracle_cell_position(const sph_oracle_env *env, const float pos[3], int32_t c[3])
{
    c[0] = f2i(floorf((pos[0] + env->xHalf + env->offset) / env->cellSize));
    c[1] = f2i(floorf((pos[1] + env->yHalf + env->offset) / env->cellSize));
    c[2] = f2i(floorf((pos[2] + env->zHalf + env->offset) / env->cellSize));
}

/* cellId -- gridBuild.wgsl:41-48: xi + yi * xGrids + zi * xGrids * yGrids (i32) */
int32_t sph_oracle_cell_id(const sph_oracle_env *env, const float pos[3])
{
    int32_t c[3];
    sph_oracle_cell_position(env, pos, c);
    const uint32_t id = (uint32_t)c[0] + (uint32_t)c[1] * (uint32_t)env->xGrids +
                        (uint32_t)c[2] * (uint32_t)env->xGrids * (uint32_t)env->yGrids;
    return (int32_t)id;
}

static float dot3(const float a[3], const float b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* density.wgsl:41-51 */
float sph_oracle_density_kernel(const sph_oracle_uniforms *u, float r)
{
    const float scale = 315.0f / ((float)(64. * 3.1415926535) * u->kernelRadiusPow9);
    const float dd = u->kernelRadiusPow2 - r * r;
    return ((scale * dd) * dd) * dd;
}

float sph_oracle_near_density_kernel(const sph_oracle_uniforms *u, float r)
{
    const float scale = 15.0f / ((float)3.1415926535 * u->kernelRadiusPow6);
    const float d = u->kernelRadius - r;
    return ((scale * d) * d) * d;
}

/* force.wgsl:41-59 */
float sph_oracle_density_gradient(const sph_oracle_uniforms *u, float r)
{
    const float scale = 45.0f / ((float)3.1415926535 * u->kernelRadiusPow6);
    const float d = u->kernelRadius - r;
    return (scale * d) * d;
}

float sph_oracle_near_density_gradient(const sph_oracle_uniforms *u, float r)
{
    const float scale = 45.0f / ((float)3.1415926535 * u->kernelRadiusPow5);
    const float d = u->kernelRadius - r;
    return (scale * d) * d;
}

float sph_oracle_viscosity_laplacian(const sph_oracle_uniforms *u, float r)
{
    const float scale = 45.0f / ((float)3.1415926535 * u->kernelRadiusPow6);
    const float d = u->kernelRadius - r;
    return scale * d;
}

/* gridClear + gridBuild + prefix sum + reorderParticles (sph.ts:304-316):
 * start[c] = exclusive prefix of the cell counts (gridCount + 1 entries),
 * sorted[] = the binned particles, cell by cell, decreasing index within a
 * cell; order[k] = index of sorted[k].  Returns the number binned. */
uint32_t sph_oracle_bin(const sph_oracle_particle *p, uint32_t n, const sph_oracle_env *env,
                        uint32_t *start, uint32_t *order, sph_oracle_particle *sorted)
{
    const uint64_t G = sph_oracle_grid_count(env);
    memset(start, 0, sizeof(uint32_t) * (G + 1));
    int32_t *cell = malloc(sizeof(int32_t) * (n ? n : 1));
    uint32_t *rank = malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        cell[i] = sph_oracle_cell_id(env, p[i].position);
        if (cell[i] >= 0 && (uint64_t)cell[i] < G) rank[i] = start[cell[i]]++;
    }
    uint32_t run = 0;
    for (uint64_t c = 0; c <= G; ++c) {  /* exclusive prefix sum, in place */
        const uint32_t v = start[c];
        start[c] = run;
        run += v;
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (cell[i] < 0 || (uint64_t)cell[i] >= G) continue;
        const uint32_t t = start[cell[i] + 1] - rank[i] - 1; /* reorderParticles.wgsl:57 */
        order[t] = i;
        if (sorted) sorted[t] = p[i];
    }
    free(cell);
    free(rank);
    return run;
}

/* Neighbour rows of density.wgsl:72-85 / force.wgsl:82-96: for dz, dy in
 * [-1, 1] inside the grid, the sorted range from the first to the last cell
 * of the x row. */
typedef struct { uint32_t lo, hi; } range;

static int rows(const sph_oracle_env *env, const uint32_t *start, const float pos[3],
                range out[9])
{
    int32_t v[3];
    sph_oracle_cell_position(env, pos, v);
    if (!(v[0] < env->xGrids && 0 <= v[0] && v[1] < env->yGrids && 0 <= v[1] &&
          v[2] < env->zGrids && 0 <= v[2]))
        return -1;
    int k = 0;
    const int dz0 = v[2] > 0 ? -1 : -v[2], dz1 = env->zGrids - v[2] - 1 < 1 ? env->zGrids - v[2] - 1 : 1;
    const int dy0 = v[1] > 0 ? -1 : -v[1], dy1 = env->yGrids - v[1] - 1 < 1 ? env->yGrids - v[1] - 1 : 1;
    const int dx0 = v[0] > 0 ? -1 : -v[0], dx1 = env->xGrids - v[0] - 1 < 1 ? env->xGrids - v[0] - 1 : 1;
    for (int dz = dz0; dz <= dz1; ++dz)
        for (int dy = dy0; dy <= dy1; ++dy) {
            const int32_t s = (v[0] + dx0) + (v[1] + dy) * env->xGrids +
                              (v[2] + dz) * env->xGrids * env->yGrids;
            const int32_t e = (v[0] + dx1) + (v[1] + dy) * env->xGrids +
                              (v[2] + dz) * env->xGrids * env->yGrids;
            out[k].lo = start[s];
            out[k].hi = start[e + 1];
            ++k;
        }
    return k;
}

/* computeDensity -- density.wgsl:65-106 (pos_j from the sorted array). */
void sph_oracle_density(sph_oracle_particle *p, uint32_t n, const sph_oracle_particle *sorted,
                        const uint32_t *start, const sph_oracle_env *env,
                        const sph_oracle_uniforms *u)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (uint32_t i = 0; i < n; ++i) {
        float density = 0.0f, near_density = 0.0f;
        range r[9];
        const int nr = rows(env, start, p[i].position, r);
        for (int k = 0; k < nr; ++k)
            for (uint32_t j = r[k].lo; j < r[k].hi; ++j) {
                float d[3];
                for (int a = 0; a < 3; ++a) d[a] = p[i].position[a] - sorted[j].position[a];
                const float r2 = dot3(d, d);
                if (r2 < u->kernelRadiusPow2) {
                    density = density + u->mass * sph_oracle_density_kernel(u, sqrtf(r2));
                    near_density =
                        near_density + u->mass * sph_oracle_near_density_kernel(u, sqrtf(r2));
                }
            }
        p[i].density = density;
        p[i].nearDensity = near_density;
    }
}

/* computeForce -- force.wgsl:73-152 (neighbour data from the sorted array). */
void sph_oracle_force(sph_oracle_particle *p, uint32_t n, const sph_oracle_particle *sorted,
                      const uint32_t *start, const sph_oracle_env *env,
                      const sph_oracle_uniforms *u)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (uint32_t i = 0; i < n; ++i) {
        const float density_i = p[i].density, near_density_i = p[i].nearDensity;
        float fpress[3] = {0.0f, 0.0f, 0.0f}, fvisc[3] = {0.0f, 0.0f, 0.0f};
        range r[9];
        const int nr = rows(env, start, p[i].position, r);
        for (int k = 0; k < nr; ++k)
            for (uint32_t j = r[k].lo; j < r[k].hi; ++j) {
                const sph_oracle_particle *q = &sorted[j];
                float d[3];
                for (int a = 0; a < 3; ++a) d[a] = p[i].position[a] - q->position[a];
                const float r2 = dot3(d, d);
                if (q->density == 0.0f || q->nearDensity == 0.0f) continue;
                if (r2 < u->kernelRadiusPow2 && 0.0f < r2) { /* 1e-64 rounds to 0 in f32 */
                    const float rr = sqrtf(r2);
                    const float pressure_i = u->stiffness * (density_i - u->restDensity);
                    const float pressure_j = u->stiffness * (q->density - u->restDensity);
                    const float near_pressure_i = u->nearStiffness * near_density_i;
                    const float near_pressure_j = u->nearStiffness * q->nearDensity;
                    const float shared = (pressure_i + pressure_j) / 2.0f;
                    const float near_shared = (near_pressure_i + near_pressure_j) / 2.0f;
                    float dir[3];
                    for (int a = 0; a < 3; ++a) dir[a] = q->position[a] - p[i].position[a];
                    const float len = sqrtf(dot3(dir, dir));
                    for (int a = 0; a < 3; ++a) dir[a] = dir[a] / len;
                    const float grad = sph_oracle_density_gradient(u, rr);
                    const float near_grad = sph_oracle_near_density_gradient(u, rr);
                    const float lap = sph_oracle_viscosity_laplacian(u, rr);
                    const float sp = -u->mass * shared, nsp = -u->mass * near_shared;
                    for (int a = 0; a < 3; ++a)
                        fpress[a] = fpress[a] + ((sp * dir[a]) * grad) / q->density;
                    for (int a = 0; a < 3; ++a)
                        fpress[a] = fpress[a] + ((nsp * dir[a]) * near_grad) / q->nearDensity;
                    for (int a = 0; a < 3; ++a) {
                        const float rel = q->v[a] - p[i].v[a];
                        fvisc[a] = fvisc[a] + ((u->mass * rel) * lap) / q->density;
                    }
                }
            }
        for (int a = 0; a < 3; ++a) fvisc[a] = fvisc[a] * u->viscosity;
        const float grav[3] = {0.0f, -9.8f, 0.0f};
        for (int a = 0; a < 3; ++a)
            p[i].force[a] = (fpress[a] + fvisc[a]) + density_i * grav[a];
    }
}

/* integrate -- integrate.wgsl:35-65 */
void sph_oracle_integrate(sph_oracle_particle *p, uint32_t n, const float real_half[3],
                          const sph_oracle_uniforms *u)
{
    const float wall = 8000.0f;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (uint32_t i = 0; i < n; ++i) {
        if (p[i].density == 0.0f) continue;
        float a[3];
        for (int k = 0; k < 3; ++k) a[k] = p[i].force[k] / p[i].density;
        float plus[3], minus[3];
        for (int k = 0; k < 3; ++k) {
            plus[k] = real_half[k] - p[i].position[k];
            minus[k] = real_half[k] + p[i].position[k];
        }
        /* axis force vectors: unit vector * wallStiffness * min(dist, 0) */
        float f[3][3];  /* f[axis][component] = plus + minus force of that axis */
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < 3; ++k) {
                const float up = k == ax ? 1.0f : 0.0f, down = k == ax ? -1.0f : 0.0f;
                const float fp = (up * wall) * fminf(plus[ax], 0.0f);
                const float fm = (down * wall) * fminf(minus[ax], 0.0f);
                f[ax][k] = fp + fm;
            }
        for (int k = 0; k < 3; ++k) a[k] = a[k] + ((f[0][k] + f[1][k]) + f[2][k]);
        for (int k = 0; k < 3; ++k) p[i].v[k] = p[i].v[k] + u->dt * a[k];
        for (int k = 0; k < 3; ++k) p[i].position[k] = p[i].position[k] + u->dt * p[i].v[k];
    }
}

/* One substep of sph.ts:304-332 (the second reorder refreshes the sorted
 * copy with the new densities; the particle order does not change). */
void sph_oracle_substep(sph_oracle_particle *p, uint32_t n, const float real_half[3],
                        const sph_oracle_env *env, const sph_oracle_uniforms *u,
                        sph_oracle_posvel *posvel)
{
    const uint64_t G = sph_oracle_grid_count(env);
    uint32_t *start = malloc(sizeof(uint32_t) * (G + 1));
    uint32_t *order = malloc(sizeof(uint32_t) * (n ? n : 1));
    sph_oracle_particle *sorted = malloc(sizeof(sph_oracle_particle) * (n ? n : 1));
    const uint32_t m = sph_oracle_bin(p, n, env, start, order, sorted);
    sph_oracle_density(p, n, sorted, start, env, u);
    for (uint32_t k = 0; k < m; ++k) sorted[k] = p[order[k]];
    sph_oracle_force(p, n, sorted, start, env, u);
    sph_oracle_integrate(p, n, real_half, u);
    if (posvel)
        for (uint32_t i = 0; i < n; ++i) {  /* copyPosition.wgsl:34-38 */
            memset(&posvel[i], 0, sizeof(posvel[i]));
            memcpy(posvel[i].position, p[i].position, sizeof(float) * 3);
            memcpy(posvel[i].v, p[i].v, sizeof(float) * 3);
        }
    free(start);
    free(order);
    free(sorted);
}
