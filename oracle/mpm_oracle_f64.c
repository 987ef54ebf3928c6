/*
 * mpm_oracle_f64.c -- FP64, unquantized MLS-MPM rollout and loss: the
 * function whose central finite differences check the HIP adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (see mpm_oracle.c).
 *
 * Same substep as mls-mpm/{p2g_1,p2g_2,updateGrid,g2p}.wgsl with every value
 * in double precision and the i32 fixed-point grid replaced by exact sums
 * (encode/decode are the identity), so the rollout is a piecewise-smooth
 * function of the initial velocities and of stiffness and viscosity.
 * Grid cells with mass > 0 are updated; others keep their raw momentum,
 * like updateGrid.wgsl:26.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

typedef struct {
    double x[3], v[3], C[3][3]; /* C[col][row] */
} p64;

static void weights64(double x, double *ci, double d_out[1], double w[3])
{
    double c = floor(x);
    double d = x - (c + 0.5);
    *ci = c;
    d_out[0] = d;
    w[0] = 0.5 * (0.5 - d) * (0.5 - d);
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * (0.5 + d) * (0.5 + d);
}

static void substep64(p64 *p, uint32_t n, double *grid /* G x 4: vx,vy,vz,m */, int64_t G,
                      const int32_t dims[3], const double rb[3], double k, double mu, double rho0,
                      double dt)
{
    const int64_t Y = dims[1], Z = dims[2];
    memset(grid, 0, sizeof(double) * 4 * (size_t)G);
    /* p2g_1 */
    for (uint32_t i = 0; i < n; ++i) {
        double ci[3], d[3], w[3][3];
        for (int a = 0; a < 3; ++a) weights64(p[i].x[a], &ci[a], &d[a], w[a]);
        for (int gx = 0; gx < 3; ++gx)
            for (int gy = 0; gy < 3; ++gy)
                for (int gz = 0; gz < 3; ++gz) {
                    const int g[3] = {gx, gy, gz};
                    double wt = w[0][gx] * w[1][gy] * w[2][gz], dist[3];
                    int64_t c[3];
                    for (int a = 0; a < 3; ++a) {
                        double cx = ci[a] + g[a] - 1.0;
                        dist[a] = cx + 0.5 - p[i].x[a];
                        c[a] = (int64_t)cx;
                    }
                    int64_t idx = c[0] * Y * Z + c[1] * Z + c[2];
                    if (idx < 0 || idx >= G) continue;
                    double *cell = grid + 4 * idx;
                    for (int r = 0; r < 3; ++r) {
                        double q = p[i].C[0][r] * dist[0] + p[i].C[1][r] * dist[1] +
                                   p[i].C[2][r] * dist[2];
                        cell[r] += wt * (p[i].v[r] + q);
                    }
                    cell[3] += wt;
                }
    }
    /* p2g_2 */
    double *mom = calloc((size_t)G * 3, sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
        double ci[3], d[3], w[3][3];
        for (int a = 0; a < 3; ++a) weights64(p[i].x[a], &ci[a], &d[a], w[a]);
        double rho = 0.0;
        for (int gx = 0; gx < 3; ++gx)
            for (int gy = 0; gy < 3; ++gy)
                for (int gz = 0; gz < 3; ++gz) {
                    const int g[3] = {gx, gy, gz};
                    int64_t c[3];
                    for (int a = 0; a < 3; ++a) c[a] = (int64_t)(ci[a] + g[a] - 1.0);
                    int64_t idx = c[0] * Y * Z + c[1] * Z + c[2];
                    if (idx < 0 || idx >= G) continue;
                    rho += grid[4 * idx + 3] * w[0][gx] * w[1][gy] * w[2][gz];
                }
        double vol = 1.0 / rho;
        double r = rho / rho0;
        double pr = k * (r * r * r * r * r - 1.0);
        if (pr < 0.0) pr = 0.0;
        double M[3][3];
        for (int c = 0; c < 3; ++c)
            for (int rr = 0; rr < 3; ++rr) {
                double s = (c == rr ? -pr : 0.0) + mu * (p[i].C[c][rr] + p[i].C[rr][c]);
                M[c][rr] = -vol * 4.0 * s * dt;
            }
        for (int gx = 0; gx < 3; ++gx)
            for (int gy = 0; gy < 3; ++gy)
                for (int gz = 0; gz < 3; ++gz) {
                    const int g[3] = {gx, gy, gz};
                    double wt = w[0][gx] * w[1][gy] * w[2][gz], dist[3];
                    int64_t c[3];
                    for (int a = 0; a < 3; ++a) {
                        double cx = ci[a] + g[a] - 1.0;
                        dist[a] = cx + 0.5 - p[i].x[a];
                        c[a] = (int64_t)cx;
                    }
                    int64_t idx = c[0] * Y * Z + c[1] * Z + c[2];
                    if (idx < 0 || idx >= G) continue;
                    for (int rr = 0; rr < 3; ++rr)
                        mom[3 * idx + rr] += wt * (M[0][rr] * dist[0] + M[1][rr] * dist[1] +
                                                   M[2][rr] * dist[2]);
                }
    }
    /* updateGrid */
    int32_t hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (int32_t)(ceil(rb[a]) - 3.0);
    for (int64_t id = 0; id < G; ++id) {
        double *cell = grid + 4 * id;
        for (int r = 0; r < 3; ++r) cell[r] += mom[3 * id + r];
        if (cell[3] > 0.0) {
            for (int r = 0; r < 3; ++r) cell[r] /= cell[3];
            cell[1] += -0.3 * dt;
            int64_t x = id / Z / Y, y = (id / Z) % Y, z = id % Z;
            if (x < 2 || x > hi[0]) cell[0] = 0.0;
            if (y < 2 || y > hi[1]) cell[1] = 0.0;
            if (z < 2 || z > hi[2]) cell[2] = 0.0;
        }
    }
    free(mom);
    /* g2p */
    for (uint32_t i = 0; i < n; ++i) {
        double ci[3], d[3], w[3][3], v[3] = {0, 0, 0}, B[3][3];
        memset(B, 0, sizeof(B));
        for (int a = 0; a < 3; ++a) weights64(p[i].x[a], &ci[a], &d[a], w[a]);
        for (int gx = 0; gx < 3; ++gx)
            for (int gy = 0; gy < 3; ++gy)
                for (int gz = 0; gz < 3; ++gz) {
                    const int g[3] = {gx, gy, gz};
                    double wt = w[0][gx] * w[1][gy] * w[2][gz], dist[3];
                    int64_t c[3];
                    for (int a = 0; a < 3; ++a) {
                        double cx = ci[a] + g[a] - 1.0;
                        dist[a] = cx + 0.5 - p[i].x[a];
                        c[a] = (int64_t)cx;
                    }
                    int64_t idx = c[0] * Y * Z + c[1] * Z + c[2];
                    if (idx < 0 || idx >= G) continue;
                    for (int r = 0; r < 3; ++r) {
                        double wv = grid[4 * idx + r] * wt;
                        v[r] += wv;
                        for (int cc = 0; cc < 3; ++cc) B[cc][r] += wv * dist[cc];
                    }
                }
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) p[i].C[c][r] = 4.0 * B[c][r];
        for (int a = 0; a < 3; ++a) {
            double xa = p[i].x[a] + v[a] * dt;
            if (xa < 1.0) xa = 1.0;
            if (xa > rb[a] - 2.0) xa = rb[a] - 2.0;
            p[i].x[a] = xa;
            double xn = xa + v[a] * dt * 3.0, wall_max = rb[a] - 4.0;
            if (xn < 3.0) v[a] += 0.3 * (3.0 - xn);
            if (xn > wall_max) v[a] += 0.3 * (wall_max - xn);
            p[i].v[a] = v[a];
        }
    }
}

/* Loss of mls-mpm rollouts used by the differentiable mode:
 * |centre of mass - target|^2 + ke_weight * mean |v|^2 at the final step. */
double oracle_rollout_loss_f64(const float *particles, const double *x0, const double *v0,
                               uint32_t n, const float init_box[3], double stiffness,
                               double viscosity,
                               double rest_density, double dt, uint32_t substeps,
                               const double target[3], double ke_weight, float *final_state)
{
    int32_t dims[3];
    double rb[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = (int32_t)init_box[a];
        rb[a] = init_box[a];
    }
    const int64_t G = (int64_t)ceil(init_box[0]) * (int64_t)ceil(init_box[1]) *
                      (int64_t)ceil(init_box[2]);
    p64 *p = malloc(sizeof(p64) * (n ? n : 1));
    double *grid = malloc(sizeof(double) * 4 * (size_t)G);
    for (uint32_t i = 0; i < n; ++i) {
        const float *rec = particles + 20 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
            p[i].x[a] = x0 ? x0[3 * i + a] : rec[a];
            p[i].v[a] = v0 ? v0[3 * i + a] : rec[4 + a];
        }
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) p[i].C[c][r] = rec[8 + 4 * c + r];
    }
    for (uint32_t s = 0; s < substeps; ++s)
        substep64(p, n, grid, G, dims, rb, stiffness, viscosity, rest_density, dt);
    double com[3] = {0, 0, 0}, ke = 0.0;
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            com[a] += p[i].x[a] / n;
            ke += p[i].v[a] * p[i].v[a] / n;
        }
    double loss = ke_weight * ke;
    for (int a = 0; a < 3; ++a) loss += (com[a] - target[a]) * (com[a] - target[a]);
    if (final_state)
        for (uint32_t i = 0; i < n; ++i) {
            float *rec = final_state + 20 * (size_t)i;
            memset(rec, 0, 80);
            for (int a = 0; a < 3; ++a) {
                rec[a] = (float)p[i].x[a];
                rec[4 + a] = (float)p[i].v[a];
            }
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) rec[8 + 4 * c + r] = (float)p[i].C[c][r];
        }
    free(p);
    free(grid);
    return loss;
}
