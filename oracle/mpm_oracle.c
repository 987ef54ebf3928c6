/*
 * mpm_oracle.c -- CPU restatement of WebGPU-Ocean's MLS-MPM substep
 * (mls-mpm/clearGrid.wgsl, p2g_1.wgsl, p2g_2.wgsl, updateGrid.wgsl, g2p.wgsl,
 * copyPosition.wgsl; dispatch order mls-mpm/mls-mpm.ts:245-268).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library; the product path never does.
 *
 * Build with -ffp-contract=off and no fast-math: every float operation below
 * is one IEEE-754 binary32 operation in the order the WGSL source writes it
 * (left-to-right for chains of * and +, mat*vec summed column by column).
 *
 * Semantics fixed here where WGSL leaves them to the implementation (the same
 * choices are made by the HIP kernels):
 *   - f32 -> i32 truncates toward zero and saturates; NaN converts to 0.
 *   - pow(r, 5.) is evaluated as ((r*r)*(r*r))*r.
 *   - f32 division is IEEE correctly rounded.
 *   - a cell index outside [0, G) is dropped on write and reads as 0.
 *   - mat3x3 * vec3 sums as (c0*v.x + c1*v.y) + c2*v.z.
 */
#include "mpm_oracle.h"

#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* i32(f32): truncate toward zero, saturating, NaN -> 0. */
int32_t oracle_f32_to_i32(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

/* encodeFixedPoint -- mls-mpm/p2g_1.wgsl:15-17 */
int32_t oracle_encode(float f, float mult) { return oracle_f32_to_i32(f * mult); }

/* decodeFixedPoint -- mls-mpm/p2g_2.wgsl:22-24 (a division, not a multiply) */
float oracle_decode(int32_t i, float mult) { return (float)i / mult; }

/* pow(r, 5.) -- mls-mpm/p2g_2.wgsl:63 */
float oracle_pow5(float r)
{
    float r2 = r * r;
    float r4 = r2 * r2;
    return r4 * r;
}

/* Quadratic B-spline weights of one axis -- mls-mpm/p2g_1.wgsl:30-34.
 * x is the particle coordinate; d = x - (floor(x) + 0.5). */
void oracle_weights(float x, float w[3])
{
    float ci = floorf(x);
    float d = x - (ci + 0.5f);
    w[0] = (0.5f * (0.5f - d)) * (0.5f - d);
    w[1] = 0.75f - d * d;
    w[2] = (0.5f * (0.5f + d)) * (0.5f + d);
}

/* ceil(bx)*ceil(by)*ceil(bz) -- mls-mpm/mls-mpm.ts:229 */
uint64_t oracle_grid_count(const float b[3])
{
    return (uint64_t)ceilf(b[0]) * (uint64_t)ceilf(b[1]) * (uint64_t)ceilf(b[2]);
}

/* Per-particle stencil setup shared by p2g_1, p2g_2 and g2p
 * (mls-mpm/p2g_1.wgsl:29-34, p2g_2.wgsl:35-40, g2p.wgsl:32-37). */
typedef struct {
    float cell_idx[3];
    float w[3][3]; /* w[k][axis] */
} stencil;

static void stencil_init(const float x[3], stencil *s)
{
    for (int a = 0; a < 3; ++a) {
        float ci = floorf(x[a]);
        float d = x[a] - (ci + 0.5f);
        s->cell_idx[a] = ci;
        s->w[0][a] = (0.5f * (0.5f - d)) * (0.5f - d);
        s->w[1][a] = 0.75f - d * d;
        s->w[2][a] = (0.5f * (0.5f + d)) * (0.5f + d);
    }
}

/* cell_x, cell_dist and cell_index of stencil node (gx,gy,gz)
 * (mls-mpm/p2g_1.wgsl:41-56).  Index math is i32 with wrap-around. */
static int64_t stencil_node(const stencil *s, const float x[3], int gx, int gy, int gz,
                            int32_t Y, int32_t Z, float dist[3])
{
    const int g[3] = {gx, gy, gz};
    int32_t ic[3];
    for (int a = 0; a < 3; ++a) {
        float cx = (s->cell_idx[a] + (float)g[a]) - 1.0f;
        dist[a] = (cx + 0.5f) - x[a];
        ic[a] = oracle_f32_to_i32(cx);
    }
    uint32_t idx = (uint32_t)ic[0] * (uint32_t)Y * (uint32_t)Z
                 + (uint32_t)ic[1] * (uint32_t)Z + (uint32_t)ic[2];
    return (int64_t)(int32_t)idx;
}

static inline float node_weight(const stencil *s, int gx, int gy, int gz)
{
    return (s->w[gx][0] * s->w[gy][1]) * s->w[gz][2];
}

static inline void grid_add(int32_t *dst, int32_t v, int atomic)
{
    if (atomic)
        __atomic_fetch_add((uint32_t *)dst, (uint32_t)v, __ATOMIC_RELAXED);
    else
        *dst = (int32_t)((uint32_t)*dst + (uint32_t)v);
}

/* clearGrid -- mls-mpm/clearGrid.wgsl:10-18 */
void oracle_clear_grid(oracle_cell *cells, uint64_t g)
{
    memset(cells, 0, (size_t)g * sizeof(oracle_cell));
}

/* p2g_1 body for one particle -- mls-mpm/p2g_1.wgsl:27-63 */
static void p2g1_particle(const oracle_particle *p, oracle_cell *cells, uint64_t g,
                          int32_t Y, int32_t Z, float mult, int atomic)
{
    stencil s;
    stencil_init(p->position, &s);
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy)
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(&s, gx, gy, gz);
                float dist[3];
                int64_t idx = stencil_node(&s, p->position, gx, gy, gz, Y, Z, dist);
                float Q[3];
                for (int r = 0; r < 3; ++r)
                    Q[r] = (p->C[0][r] * dist[0] + p->C[1][r] * dist[1]) + p->C[2][r] * dist[2];
                float mass_contrib = weight * 1.0f;
                float vel[3];
                for (int r = 0; r < 3; ++r) vel[r] = mass_contrib * (p->v[r] + Q[r]);
                if (idx < 0 || (uint64_t)idx >= g) continue;
                oracle_cell *c = &cells[idx];
                grid_add(&c->mass, oracle_encode(mass_contrib, mult), atomic);
                grid_add(&c->vx, oracle_encode(vel[0], mult), atomic);
                grid_add(&c->vy, oracle_encode(vel[1], mult), atomic);
                grid_add(&c->vz, oracle_encode(vel[2], mult), atomic);
            }
}

void oracle_p2g_1(const oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                  const float ib[3], const oracle_params *prm)
{
    const int32_t Y = oracle_f32_to_i32(ib[1]), Z = oracle_f32_to_i32(ib[2]);
    for (uint32_t i = 0; i < n; ++i)
        p2g1_particle(&p[i], cells, g, Y, Z, prm->fixed_point_multiplier, 0);
}

/* p2g_2 body for one particle -- mls-mpm/p2g_2.wgsl:33-92 */
static void p2g2_particle(const oracle_particle *p, oracle_cell *cells, uint64_t g,
                          int32_t Y, int32_t Z, const oracle_params *prm, int atomic)
{
    const float mult = prm->fixed_point_multiplier;
    stencil s;
    stencil_init(p->position, &s);
    float density = 0.0f;
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy)
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(&s, gx, gy, gz);
                float dist[3];
                int64_t idx = stencil_node(&s, p->position, gx, gy, gz, Y, Z, dist);
                int32_t m = (idx < 0 || (uint64_t)idx >= g) ? 0 : cells[idx].mass;
                density = density + oracle_decode(m, mult) * weight;
            }

    float volume = 1.0f / density;
    float pw = prm->stiffness * (oracle_pow5(density / prm->rest_density) - 1.0f);
    float pressure = (pw > -0.0f) ? pw : -0.0f; /* max(-0.0, .) */

    float M[3][3]; /* M[col][row] = eq_16_term0 */
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            float stress = (c == r) ? -pressure : 0.0f;
            float strain = p->C[c][r] + p->C[r][c];
            stress = stress + prm->dynamic_viscosity * strain;
            M[c][r] = ((-volume * 4.0f) * stress) * prm->dt;
        }

    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy)
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(&s, gx, gy, gz);
                float dist[3];
                int64_t idx = stencil_node(&s, p->position, gx, gy, gz, Y, Z, dist);
                float Mw[3][3];
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) Mw[c][r] = M[c][r] * weight;
                float mom[3];
                for (int r = 0; r < 3; ++r)
                    mom[r] = (Mw[0][r] * dist[0] + Mw[1][r] * dist[1]) + Mw[2][r] * dist[2];
                if (idx < 0 || (uint64_t)idx >= g) continue;
                oracle_cell *c = &cells[idx];
                grid_add(&c->vx, oracle_encode(mom[0], mult), atomic);
                grid_add(&c->vy, oracle_encode(mom[1], mult), atomic);
                grid_add(&c->vz, oracle_encode(mom[2], mult), atomic);
            }
}

void oracle_p2g_2(const oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                  const float ib[3], const oracle_params *prm)
{
    const int32_t Y = oracle_f32_to_i32(ib[1]), Z = oracle_f32_to_i32(ib[2]);
    for (uint32_t i = 0; i < n; ++i) p2g2_particle(&p[i], cells, g, Y, Z, prm, 0);
}

/* updateGrid body for one cell -- mls-mpm/updateGrid.wgsl:24-45 */
static void update_cell(oracle_cell *c, int32_t id, int32_t Y, int32_t Z,
                        const int32_t wall_hi[3], const oracle_params *prm)
{
    if (c->mass > 0) {
        const float mult = prm->fixed_point_multiplier;
        float fv[3] = {oracle_decode(c->vx, mult), oracle_decode(c->vy, mult),
                       oracle_decode(c->vz, mult)};
        float m = oracle_decode(c->mass, mult);
        for (int a = 0; a < 3; ++a) fv[a] = fv[a] / m;
        float gdt = -0.3f * prm->dt;
        c->vx = oracle_encode(fv[0], mult);
        c->vy = oracle_encode(fv[1] + gdt, mult);
        c->vz = oracle_encode(fv[2], mult);
        int32_t x = id / Z / Y;
        int32_t y = (id / Z) % Y;
        int32_t z = id % Z;
        if (x < 2 || x > wall_hi[0]) c->vx = 0;
        if (y < 2 || y > wall_hi[1]) c->vy = 0;
        if (z < 2 || z > wall_hi[2]) c->vz = 0;
    }
}

static void wall_limits(const float rb[3], int32_t hi[3])
{
    for (int a = 0; a < 3; ++a) hi[a] = oracle_f32_to_i32(ceilf(rb[a]) - 3.0f);
}

void oracle_update_grid(oracle_cell *cells, uint64_t g, const float rb[3], const float ib[3],
                        const oracle_params *prm)
{
    const int32_t Y = oracle_f32_to_i32(ib[1]), Z = oracle_f32_to_i32(ib[2]);
    int32_t hi[3];
    wall_limits(rb, hi);
    for (uint64_t id = 0; id < g; ++id) update_cell(&cells[id], (int32_t)id, Y, Z, hi, prm);
}

/* g2p body for one particle -- mls-mpm/g2p.wgsl:28-90 */
static void g2p_particle(oracle_particle *p, const oracle_cell *cells, uint64_t g, int32_t Y,
                         int32_t Z, const float rb[3], const oracle_params *prm)
{
    const float mult = prm->fixed_point_multiplier, dt = prm->dt;
    float v[3] = {0.0f, 0.0f, 0.0f};
    float x[3] = {p->position[0], p->position[1], p->position[2]};
    stencil s;
    stencil_init(x, &s);
    float B[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy)
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(&s, gx, gy, gz);
                float dist[3];
                int64_t idx = stencil_node(&s, x, gx, gy, gz, Y, Z, dist);
                int32_t cv[3] = {0, 0, 0};
                if (idx >= 0 && (uint64_t)idx < g) {
                    cv[0] = cells[idx].vx; cv[1] = cells[idx].vy; cv[2] = cells[idx].vz;
                }
                float wv[3];
                for (int r = 0; r < 3; ++r) wv[r] = oracle_decode(cv[r], mult) * weight;
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) B[c][r] = B[c][r] + wv[r] * dist[c];
                for (int r = 0; r < 3; ++r) v[r] = v[r] + wv[r];
            }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) p->C[c][r] = B[c][r] * 4.0f;
    for (int a = 0; a < 3; ++a) {
        float xa = x[a] + v[a] * dt;
        xa = fminf(fmaxf(xa, 1.0f), rb[a] - 2.0f); /* clamp(e, 1., real - 2.) */
        p->position[a] = xa;
    }
    const float k = 3.0f, wall_stiffness = 0.3f;
    for (int a = 0; a < 3; ++a) {
        float xn = p->position[a] + (v[a] * dt) * k;
        float wall_max = rb[a] - 4.0f;
        if (xn < 3.0f) v[a] = v[a] + wall_stiffness * (3.0f - xn);
        if (xn > wall_max) v[a] = v[a] + wall_stiffness * (wall_max - xn);
    }
    for (int a = 0; a < 3; ++a) p->v[a] = v[a];
}

void oracle_g2p(oracle_particle *p, uint32_t n, const oracle_cell *cells, uint64_t g,
                const float rb[3], const float ib[3], const oracle_params *prm)
{
    const int32_t Y = oracle_f32_to_i32(ib[1]), Z = oracle_f32_to_i32(ib[2]);
    for (uint32_t i = 0; i < n; ++i) g2p_particle(&p[i], cells, g, Y, Z, rb, prm);
}

/* copyPosition -- mls-mpm/copyPosition.wgsl:15-21 (padding written as 0) */
void oracle_copy_position(const oracle_particle *p, uint32_t n, oracle_posvel *out)
{
    for (uint32_t i = 0; i < n; ++i) {
        memset(&out[i], 0, sizeof(out[i]));
        memcpy(out[i].position, p[i].position, sizeof(out[i].position));
        memcpy(out[i].v, p[i].v, sizeof(out[i].v));
    }
}

void oracle_substep(oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                    oracle_posvel *posvel, const float rb[3], const float ib[3],
                    const oracle_params *prm)
{
    oracle_clear_grid(cells, g);
    oracle_p2g_1(p, n, cells, g, ib, prm);
    oracle_p2g_2(p, n, cells, g, ib, prm);
    oracle_update_grid(cells, g, rb, ib, prm);
    oracle_g2p(p, n, cells, g, rb, ib, prm);
    if (posvel) oracle_copy_position(p, n, posvel);
}

void oracle_substep_omp(oracle_particle *p, uint32_t n, oracle_cell *cells, uint64_t g,
                        oracle_posvel *posvel, const float rb[3], const float ib[3],
                        const oracle_params *prm, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int32_t Y = oracle_f32_to_i32(ib[1]), Z = oracle_f32_to_i32(ib[2]);
    const float mult = prm->fixed_point_multiplier;
    int32_t hi[3];
    wall_limits(rb, hi);
    const int64_t ng = (int64_t)g, np = (int64_t)n;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (int64_t i = 0; i < ng; ++i) memset(&cells[i], 0, sizeof(oracle_cell));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < np; ++i) p2g1_particle(&p[i], cells, g, Y, Z, mult, 1);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < np; ++i) p2g2_particle(&p[i], cells, g, Y, Z, prm, 1);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < ng; ++i) update_cell(&cells[i], (int32_t)i, Y, Z, hi, prm);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < np; ++i) g2p_particle(&p[i], cells, g, Y, Z, rb, prm);
    }
    if (posvel) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < np; ++i) oracle_copy_position(&p[i], 1, &posvel[i]);
    }
}
