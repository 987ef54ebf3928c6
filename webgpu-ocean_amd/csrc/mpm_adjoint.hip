// mpm_adjoint.hip -- hand-written adjoint of the MLS-MPM substep and the
// checkpointed backward pass (mpm_backward).
//
// Forward substep (per particle p, grid node i, weights w_ip, dist_ip =
// x_i - x_p; mls-mpm/p2g_1.wgsl, p2g_2.wgsl, updateGrid.wgsl, g2p.wgsl):
//   m_i  = sum_p w_ip                    mv_i = sum_p w_ip (v_p + C_p dist_ip)
//   rho_p = sum_i m_i w_ip,  M_p = -4 dt / rho_p * (-P(rho_p) I + mu (C_p + C_p^T))
//   mv_i += sum_p w_ip M_p dist_ip
//   v_i  = mv_i / m_i + g dt e_y   (components zeroed at the walls)
//   v_p' = sum_i w_ip v_i,  C_p' = 4 sum_i w_ip v_i (x) dist_ip,
//   x_p' = clamp(x_p + v_p' dt), then the wall penalty on v_p'.
// The adjoint runs those steps backwards: G2P becomes a float scatter onto
// the grid, the update is differentiated per cell, and each P2G becomes a
// gather at the particles (plus a float scatter for the density's mass
// dependence).  Fixed-point encode/decode are taken as the identity, so the
// gradients are those of the unquantized model evaluated along the forward
// trajectory.  All per-particle adjoints are indexed by the caller's particle
// index, so the binning permutation of the forward never matters.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "mpm_batch.h"
#include "mpm_particle.h"

#pragma clang fp contract(off)

namespace mpm {
namespace {

constexpr int kAdjBlock = 256;

// Per-cell linearisation point of the grid update.
struct GridLin {
    float m;           // decoded mass
    float vpre[3];     // momentum / mass (or raw momentum where mass <= 0)
    float vg[3];       // velocity G2P read: decode(encode(updated))
    uint32_t flags;    // bit a: component a kept (not zeroed at a wall); bit 3: mass > 0
};

// Stencil weights and their derivatives for one particle.
struct StencilD {
    float w[3][3];   // w[g][axis]
    float dw[3][3];  // d w[g][axis] / d x_axis
    float dist[3][3];
    int32_t ic[3][3];
};

__device__ __forceinline__ void stencil_d(const float x[3], StencilD &s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ci = floorf(x[a]);
        const float d = x[a] - (ci + 0.5f);
        s.w[0][a] = (0.5f * (0.5f - d)) * (0.5f - d);
        s.w[1][a] = 0.75f - d * d;
        s.w[2][a] = (0.5f * (0.5f + d)) * (0.5f + d);
        s.dw[0][a] = d - 0.5f;
        s.dw[1][a] = -2.0f * d;
        s.dw[2][a] = 0.5f + d;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const float cx = (ci + (float)g) - 1.0f;
            s.dist[a][g] = (cx + 0.5f) - x[a];
            s.ic[a][g] = f2i_sat(cx);
        }
    }
}

__device__ __forceinline__ int32_t sidx(const StencilD &s, int gx, int gy, int gz, int32_t Y,
                                        int32_t Z) {
    return (int32_t)((uint32_t)s.ic[0][gx] * (uint32_t)Y * (uint32_t)Z +
                     (uint32_t)s.ic[1][gy] * (uint32_t)Z + (uint32_t)s.ic[2][gz]);
}

// weight and its gradient at node (gx,gy,gz)
__device__ __forceinline__ void weight_grad(const StencilD &s, int gx, int gy, int gz, float &w,
                                            float dw[3]) {
    const float a = s.w[gx][0], b = s.w[gy][1], c = s.w[gz][2];
    w = a * b * c;
    dw[0] = s.dw[gx][0] * b * c;
    dw[1] = a * s.dw[gy][1] * c;
    dw[2] = a * b * s.dw[gz][2];
}

__device__ __forceinline__ void load_state(const float *st, uint32_t stride, uint32_t i,
                                           float x[3], float v[3], float C[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x[a] = st[(PX + a) * stride + i];
        v[a] = st[(VX + a) * stride + i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = st[(C00 + 3 * c + r) * stride + i];
}

// Per-particle adjoint record, caller order: gx[3] gv[3] gC[9] (15 floats, SoA).
enum : int { AX = 0, AV = 3, AC = 6, ANCOMP = 15 };

// ---- linearisation of the grid update --------------------------------------

__global__ __launch_bounds__(kAdjBlock) void k_grid_lin(const int4 *__restrict__ grid,
                                                        uint32_t G, const Uniforms *__restrict__ u,
                                                        Params prm, GridLin *__restrict__ lin) {
    const uint32_t id = blockIdx.x * kAdjBlock + threadIdx.x;
    if (id >= G) return;
    const int4 c = grid[id];
    GridLin L;
    L.m = decode(c.w, prm.mult);
    const float mv[3] = {decode(c.x, prm.mult), decode(c.y, prm.mult), decode(c.z, prm.mult)};
    const int32_t wall_hi[3] = {u->wall_hi[0], u->wall_hi[1], u->wall_hi[2]};
    const int4 up = update_cell_id(c, (int32_t)id, u->Y, u->Z, wall_hi, prm);
    L.vg[0] = decode(up.x, prm.mult);
    L.vg[1] = decode(up.y, prm.mult);
    L.vg[2] = decode(up.z, prm.mult);
    L.flags = 7u;
    if (c.w > 0) {
        L.flags |= 8u;
#pragma unroll
        for (int a = 0; a < 3; ++a) L.vpre[a] = mv[a] / L.m;
        const int32_t Y = u->Y, Z = u->Z, iid = (int32_t)id;
        const int32_t xyz[3] = {iid / Z / Y, (iid / Z) % Y, iid % Z};
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (xyz[a] < 2 || xyz[a] > wall_hi[a]) L.flags &= ~(1u << a);
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) L.vpre[a] = mv[a];
    }
    lin[id] = L;
}

// ---- G2P adjoint: particle adjoints -> grid velocity adjoint (float scatter)

__global__ __launch_bounds__(kAdjBlock) void k_adj_g2p(
    const float *__restrict__ st, const uint32_t *__restrict__ id, uint32_t stride, uint32_t n,
    const GridLin *__restrict__ lin, uint32_t G, const Uniforms *__restrict__ u, Params prm,
    const float *__restrict__ adj_in, float *__restrict__ adj_out, uint32_t astride,
    float *__restrict__ gvel) {
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = id[i];
    float x[3], vold[3], C[3][3];
    load_state(st, stride, i, x, vold, C);
    StencilD s;
    stencil_d(x, s);
    const int32_t Y = u->Y, Z = u->Z;
    // forward G2P recomputation
    float v[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                if (!in_grid(idx, G)) continue;
                const float w = (s.w[gx][0] * s.w[gy][1]) * s.w[gz][2];
                const GridLin &L = lin[idx];
#pragma unroll
                for (int r = 0; r < 3; ++r) v[r] = v[r] + L.vg[r] * w;
            }
    const float dt = prm.dt, k = 3.0f, ws = 0.3f;
    float gv[3], gxp[3], B_[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) B_[c][r] = 4.0f * adj_in[(AC + 3 * c + r) * astride + o];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float rb = u->real_box[a];
        const float x1 = x[a] + v[a] * dt;
        const float xc = fminf(fmaxf(x1, 1.0f), rb - 2.0f);
        const float xn = xc + (v[a] * dt) * k;
        const float gvo = adj_in[(AV + a) * astride + o];
        const float gxo = adj_in[(AX + a) * astride + o];
        float act = 0.0f;
        if (xn < 3.0f) act += 1.0f;
        if (xn > rb - 4.0f) act += 1.0f;
        const float gxn = -ws * gvo * act;
        const float gxc = gxo + gxn;
        const float gx1 = (x1 > 1.0f && x1 < rb - 2.0f) ? gxc : 0.0f;
        gv[a] = gvo + gxn * dt * k + gx1 * dt;
        gxp[a] = gx1;
    }
    // reverse gather: scatter to the grid and accumulate the weight terms
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                if (!in_grid(idx, G)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                const GridLin &L = lin[idx];
                float coef = 0.0f;  // d(scalar)/dw
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    float bd = 0.0f;
#pragma unroll
                    for (int c = 0; c < 3; ++c) bd += B_[c][r] * d[c];
                    const float g = gv[r] + bd;  // adjoint of v_i[r] per unit weight
                    unsafeAtomicAdd(&gvel[3 * (size_t)idx + r], w * g);
                    coef += L.vg[r] * g;
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float bv = 0.0f;  // d dist[a] / dx[a] = -1
#pragma unroll
                    for (int r = 0; r < 3; ++r) bv += B_[a][r] * L.vg[r];
                    gxp[a] += dw[a] * coef - w * bv;
                }
            }
#pragma unroll
    for (int a = 0; a < 3; ++a) adj_out[(AX + a) * astride + o] = gxp[a];
}

// ---- grid update adjoint -----------------------------------------------------

__global__ __launch_bounds__(kAdjBlock) void k_adj_update(const GridLin *__restrict__ lin,
                                                          uint32_t G, float *__restrict__ gvel,
                                                          float *__restrict__ gmv,
                                                          float *__restrict__ gm) {
    const uint32_t id = blockIdx.x * kAdjBlock + threadIdx.x;
    if (id >= G) return;
    const GridLin L = lin[id];
    float g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g[a] = (L.flags & (1u << a)) ? gvel[3 * (size_t)id + a] : 0.0f;
        gvel[3 * (size_t)id + a] = 0.0f;  // ready for the next substep
    }
    float m_adj = 0.0f;
    if (L.flags & 8u) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            gmv[3 * (size_t)id + a] = g[a] / L.m;
            m_adj -= g[a] * L.vpre[a] / L.m;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) gmv[3 * (size_t)id + a] = g[a];
    }
    gm[id] = m_adj;
}

// ---- P2G-2 adjoint: gather of the momentum adjoint, scatter of the density's

__global__ __launch_bounds__(kAdjBlock) void k_adj_p2g2(
    const float *__restrict__ st, const uint32_t *__restrict__ id, uint32_t stride, uint32_t n,
    const GridLin *__restrict__ lin, uint32_t G, const Uniforms *__restrict__ u, Params prm,
    const float *__restrict__ gmv, float *__restrict__ gm, float *__restrict__ adj_out,
    uint32_t astride, float *__restrict__ gtheta) {
    __shared__ float s_red[2][kAdjBlock / 64];
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    float gk = 0.0f, gmu = 0.0f;
    if (i < n) {
        const uint32_t o = id[i];
        float x[3], v[3], C[3][3];
        load_state(st, stride, i, x, v, C);
        StencilD s;
        stencil_d(x, s);
        const int32_t Y = u->Y, Z = u->Z;
        float rho = 0.0f;
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy)
#pragma unroll
                for (int gz = 0; gz < 3; ++gz) {
                    const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                    if (!in_grid(idx, G)) continue;
                    rho = rho + lin[idx].m * ((s.w[gx][0] * s.w[gy][1]) * s.w[gz][2]);
                }
        const float vol = 1.0f / rho;
        const float r = rho / prm.rest_density;
        const float r4 = (r * r) * (r * r);
        const float pw = prm.stiffness * (r4 * r - 1.0f);
        const bool pos = pw > 0.0f;
        const float P = pos ? pw : 0.0f;
        float S[3][3], M[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                S[c][rr] = (c == rr ? -P : 0.0f) + prm.viscosity * (C[c][rr] + C[rr][c]);
                M[c][rr] = -4.0f * vol * prm.dt * S[c][rr];
            }
        // gather: M adjoint and the weight / dist terms
        float gM[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        float gxp[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy)
#pragma unroll
                for (int gz = 0; gz < 3; ++gz) {
                    const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                    if (!in_grid(idx, G)) continue;
                    float w, dw[3];
                    weight_grad(s, gx, gy, gz, w, dw);
                    const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                    const float g[3] = {gmv[3 * (size_t)idx], gmv[3 * (size_t)idx + 1],
                                        gmv[3 * (size_t)idx + 2]};
                    float coef = 0.0f;
#pragma unroll
                    for (int rr = 0; rr < 3; ++rr) {
                        const float md = M[0][rr] * d[0] + M[1][rr] * d[1] + M[2][rr] * d[2];
                        coef += g[rr] * md;
#pragma unroll
                        for (int c = 0; c < 3; ++c) gM[c][rr] += w * g[rr] * d[c];
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        float mg = 0.0f;
#pragma unroll
                        for (int rr = 0; rr < 3; ++rr) mg += M[a][rr] * g[rr];
                        gxp[a] += dw[a] * coef - w * mg;
                    }
                }
        // M = -4 dt vol S
        float gS[3][3], gvol = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                gS[c][rr] = -4.0f * vol * prm.dt * gM[c][rr];
                gvol += -4.0f * prm.dt * S[c][rr] * gM[c][rr];
            }
        float gC[3][3];
        const float gP = -(gS[0][0] + gS[1][1] + gS[2][2]);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                gC[c][rr] = prm.viscosity * (gS[c][rr] + gS[rr][c]);
                gmu += gS[c][rr] * (C[c][rr] + C[rr][c]);
            }
        float grho = -gvol / (rho * rho);
        if (pos) {
            grho += gP * prm.stiffness * 5.0f * r4 / prm.rest_density;
            gk += gP * (r4 * r - 1.0f);
        }
        // density = sum m_i w: scatter to the masses, weight terms to x
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy)
#pragma unroll
                for (int gz = 0; gz < 3; ++gz) {
                    const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                    if (!in_grid(idx, G)) continue;
                    float w, dw[3];
                    weight_grad(s, gx, gy, gz, w, dw);
                    unsafeAtomicAdd(&gm[idx], grho * w);
                    const float m = lin[idx].m;
#pragma unroll
                    for (int a = 0; a < 3; ++a) gxp[a] += grho * m * dw[a];
                }
#pragma unroll
        for (int a = 0; a < 3; ++a) adj_out[(AX + a) * astride + o] += gxp[a];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) adj_out[(AC + 3 * c + rr) * astride + o] = gC[c][rr];
    }
    // block reduction of the scalar parameter adjoints
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        gk += __shfl_xor(gk, off, 64);
        gmu += __shfl_xor(gmu, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_red[0][threadIdx.x / 64] = gk;
        s_red[1][threadIdx.x / 64] = gmu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f, b = 0.0f;
        for (int wv = 0; wv < kAdjBlock / 64; ++wv) {
            a += s_red[0][wv];
            b += s_red[1][wv];
        }
        unsafeAtomicAdd(&gtheta[0], a);
        unsafeAtomicAdd(&gtheta[1], b);
    }
}

// ---- P2G-1 adjoint: gather of mass and momentum adjoints ---------------------

__global__ __launch_bounds__(kAdjBlock) void k_adj_p2g1(
    const float *__restrict__ st, const uint32_t *__restrict__ id, uint32_t stride, uint32_t n,
    uint32_t G, const Uniforms *__restrict__ u, const float *__restrict__ gmv,
    const float *__restrict__ gm, float *__restrict__ adj_out, uint32_t astride) {
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = id[i];
    float x[3], v[3], C[3][3];
    load_state(st, stride, i, x, v, C);
    StencilD s;
    stencil_d(x, s);
    const int32_t Y = u->Y, Z = u->Z;
    float gvp[3] = {0.0f, 0.0f, 0.0f}, gxp[3] = {0.0f, 0.0f, 0.0f};
    float gC[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                const int32_t idx = sidx(s, gx, gy, gz, Y, Z);
                if (!in_grid(idx, G)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                const float g[3] = {gmv[3 * (size_t)idx], gmv[3 * (size_t)idx + 1],
                                    gmv[3 * (size_t)idx + 2]};
                float coef = gm[idx];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float q = C[0][r] * d[0] + C[1][r] * d[1] + C[2][r] * d[2];
                    coef += g[r] * (v[r] + q);
                    gvp[r] += w * g[r];
#pragma unroll
                    for (int c = 0; c < 3; ++c) gC[c][r] += w * g[r] * d[c];
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float cg = 0.0f;
#pragma unroll
                    for (int r = 0; r < 3; ++r) cg += C[a][r] * g[r];
                    gxp[a] += dw[a] * coef - w * cg;
                }
            }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        adj_out[(AX + a) * astride + o] += gxp[a];
        adj_out[(AV + a) * astride + o] = gvp[a];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) adj_out[(AC + 3 * c + r) * astride + o] += gC[c][r];
}

// ---- loss and its gradient at the final state ---------------------------------

// partial sums of x (3) and |v|^2 (1) per block
__global__ __launch_bounds__(kAdjBlock) void k_loss_partial(const float *__restrict__ st,
                                                            uint32_t stride, uint32_t n,
                                                            double *__restrict__ acc) {
    __shared__ double s_acc[4][kAdjBlock / 64];
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            a[ax] = st[(PX + ax) * stride + i];
            const double vv = st[(VX + ax) * stride + i];
            a[3] += vv * vv;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a[k] += __shfl_xor(a[k], off, 64);
        if ((threadIdx.x & 63) == 0) s_acc[k][threadIdx.x / 64] = a[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int wv = 0; wv < kAdjBlock / 64; ++wv) t += s_acc[threadIdx.x][wv];
        atomicAdd(&acc[threadIdx.x], t);
    }
}

__global__ __launch_bounds__(kAdjBlock) void k_loss_grad(const float *__restrict__ st,
                                                         const uint32_t *__restrict__ id,
                                                         uint32_t stride, uint32_t n,
                                                         float gcx, float gcy, float gcz,
                                                         float gke, float *__restrict__ adj,
                                                         uint32_t astride) {
    const uint32_t i = blockIdx.x * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = id[i];
    const float gc[3] = {gcx, gcy, gcz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        adj[(AX + a) * astride + o] = gc[a];
        adj[(AV + a) * astride + o] = gke * st[(VX + a) * stride + i];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) adj[(AC + c) * astride + o] = 0.0f;
}

inline unsigned nblocks(uint64_t count) {
    return (unsigned)((count + kAdjBlock - 1) / kAdjBlock);
}

#define TRY(call)                        \
    do {                                 \
        hipError_t e_ = (call);          \
        if (e_ != hipSuccess) return e_; \
    } while (0)

}  // namespace

hipError_t Batch::backward(const AdjointConfig &cfg, hipStream_t st, AdjointResult &res) {
    if (h_scenes_.size() != 1 || cfg.substeps == 0 || cfg.checkpoint_every == 0)
        return hipErrorInvalidValue;
    const SceneDesc &d = h_scenes_[0];
    const uint32_t n = d.n, G = d.G, T = cfg.substeps, K = cfg.checkpoint_every;
    const uint32_t nseg = (T + K - 1) / K;
    const size_t sbytes = sizeof(float) * NCOMP * stride_, ibytes = sizeof(uint32_t) * stride_;
    const size_t gbytes = sizeof(int4) * (size_t)G;
    const Tuning saved_tuning = tun_;
    if (tun_.path != 0) {
        Tuning t = tun_;
        t.path = 0;
        TRY(set_tuning(t));
    }
    // Device buffers: checkpoints (state + id) per segment start, the segment
    // tape (state, id, raw P2G grid per substep), grid linearisation,
    // adjoints.
    char *ckpt = nullptr, *tape = nullptr;
    GridLin *lin = nullptr;
    float *adjA = nullptr, *adjB = nullptr, *gvel = nullptr, *gmv = nullptr, *gm = nullptr,
          *gtheta = nullptr;
    double *lacc = nullptr;
    auto cleanup = [&]() {
        if (ckpt) (void)hipFree(ckpt);
        if (tape) (void)hipFree(tape);
        if (lin) (void)hipFree(lin);
        if (adjA) (void)hipFree(adjA);
        if (adjB) (void)hipFree(adjB);
        if (gvel) (void)hipFree(gvel);
        if (gmv) (void)hipFree(gmv);
        if (gm) (void)hipFree(gm);
        if (gtheta) (void)hipFree(gtheta);
        if (lacc) (void)hipFree(lacc);
    };
    hipError_t e = hipSuccess;
    const size_t per_ck = sbytes + ibytes, per_tape = sbytes + ibytes + gbytes;
    const uint32_t astride = std::max<uint32_t>(n, 1);
    do {
        if ((e = hipMalloc(&ckpt, per_ck * (nseg + 1))) != hipSuccess) break;
        if ((e = hipMalloc(&tape, per_tape * K)) != hipSuccess) break;
        if ((e = hipMalloc(&lin, sizeof(GridLin) * std::max<uint32_t>(G, 1))) != hipSuccess) break;
        if ((e = hipMalloc(&adjA, sizeof(float) * ANCOMP * astride)) != hipSuccess) break;
        if ((e = hipMalloc(&adjB, sizeof(float) * ANCOMP * astride)) != hipSuccess) break;
        if ((e = hipMalloc(&gvel, sizeof(float) * 3 * std::max<uint32_t>(G, 1))) != hipSuccess) break;
        if ((e = hipMalloc(&gmv, sizeof(float) * 3 * std::max<uint32_t>(G, 1))) != hipSuccess) break;
        if ((e = hipMalloc(&gm, sizeof(float) * std::max<uint32_t>(G, 1))) != hipSuccess) break;
        if ((e = hipMalloc(&gtheta, sizeof(float) * 2)) != hipSuccess) break;
        if ((e = hipMalloc(&lacc, sizeof(double) * 4)) != hipSuccess) break;
        if ((e = hipMemsetAsync(gvel, 0, sizeof(float) * 3 * G, st)) != hipSuccess) break;
        if ((e = hipMemsetAsync(gtheta, 0, sizeof(float) * 2, st)) != hipSuccess) break;
        if ((e = hipMemsetAsync(lacc, 0, sizeof(double) * 4, st)) != hipSuccess) break;

        auto save = [&](char *dst) -> hipError_t {
            TRY(hipMemcpyAsync(dst, state_, sbytes, hipMemcpyDeviceToDevice, st));
            return hipMemcpyAsync(dst + sbytes, id_, ibytes, hipMemcpyDeviceToDevice, st);
        };
        auto restore = [&](const char *src) -> hipError_t {
            TRY(hipMemcpyAsync(state_, src, sbytes, hipMemcpyDeviceToDevice, st));
            keys_ready_ = false;
            return hipMemcpyAsync(id_, src + sbytes, ibytes, hipMemcpyDeviceToDevice, st);
        };
        // forward rollout with checkpoints every K substeps
        for (uint32_t t = 0; t < T && e == hipSuccess; ++t) {
            if (t % K == 0) e = save(ckpt + per_ck * (t / K));
            if (e == hipSuccess) e = substep(st);
        }
        if (e != hipSuccess) break;
        if ((e = save(ckpt + per_ck * nseg)) != hipSuccess) break;  // final state
        // loss and its gradient
        if (n) {
            k_loss_partial<<<nblocks(n), kAdjBlock, 0, st>>>(state_, stride_, n, lacc);
            if ((e = hipGetLastError()) != hipSuccess) break;
        }
        double acc[4];
        if ((e = hipMemcpyAsync(acc, lacc, sizeof(acc), hipMemcpyDeviceToHost, st)) != hipSuccess)
            break;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
        const double inv = n ? 1.0 / n : 0.0;
        double loss = 0.0, gc[3];
        for (int a = 0; a < 3; ++a) {
            const double com = acc[a] * inv;
            loss += (com - cfg.target[a]) * (com - cfg.target[a]);
            gc[a] = 2.0 * (com - cfg.target[a]) * inv;
        }
        loss += cfg.ke_weight * acc[3] * inv;
        res.loss = loss;
        if (n) {
            k_loss_grad<<<nblocks(n), kAdjBlock, 0, st>>>(state_, id_, stride_, n, (float)gc[0],
                                                          (float)gc[1], (float)gc[2],
                                                          (float)(2.0 * cfg.ke_weight * inv), adjA,
                                                          astride);
            if ((e = hipGetLastError()) != hipSuccess) break;
        }
        // backward over segments, recomputing each segment's tape
        float *adj_in = adjA, *adj_out = adjB;
        const Uniforms *uni = reinterpret_cast<const Uniforms *>(
            reinterpret_cast<const char *>(d_scenes_) + offsetof(SceneDesc, uni));
        for (int32_t sgi = (int32_t)nseg - 1; sgi >= 0 && e == hipSuccess; --sgi) {
            const uint32_t t0 = (uint32_t)sgi * K, t1 = std::min(T, t0 + K);
            if ((e = restore(ckpt + per_ck * sgi)) != hipSuccess) break;
            for (uint32_t t = t0; t < t1 && e == hipSuccess; ++t) {
                char *tp = tape + per_tape * (t - t0);
                if ((e = save(tp)) != hipSuccess) break;
                if ((e = substep(st)) != hipSuccess) break;
                e = hipMemcpyAsync(tp + sbytes + ibytes, grid_ + d.goff, gbytes,
                                   hipMemcpyDeviceToDevice, st);
            }
            for (int32_t t = (int32_t)t1 - 1; t >= (int32_t)t0 && e == hipSuccess; --t) {
                const char *tp = tape + per_tape * (t - t0);
                const float *tst = reinterpret_cast<const float *>(tp);
                const uint32_t *tid = reinterpret_cast<const uint32_t *>(tp + sbytes);
                const int4 *tgrid = reinterpret_cast<const int4 *>(tp + sbytes + ibytes);
                if (G) {
                    k_grid_lin<<<nblocks(G), kAdjBlock, 0, st>>>(tgrid, G, uni, d.prm, lin);
                }
                if (n) {
                    k_adj_g2p<<<nblocks(n), kAdjBlock, 0, st>>>(tst, tid, stride_, n, lin, G, uni,
                                                                 d.prm, adj_in, adj_out, astride,
                                                                 gvel);
                }
                if (G) k_adj_update<<<nblocks(G), kAdjBlock, 0, st>>>(lin, G, gvel, gmv, gm);
                if (n) {
                    k_adj_p2g2<<<nblocks(n), kAdjBlock, 0, st>>>(tst, tid, stride_, n, lin, G, uni,
                                                                  d.prm, gmv, gm, adj_out, astride,
                                                                  gtheta);
                    k_adj_p2g1<<<nblocks(n), kAdjBlock, 0, st>>>(tst, tid, stride_, n, G, uni, gmv,
                                                                  gm, adj_out, astride);
                }
                e = hipGetLastError();
                std::swap(adj_in, adj_out);
            }
        }
        if (e != hipSuccess) break;
        // outputs: adjoint of the initial state (caller order)
        if (res.grad_v0 && n) {
            std::vector<float> tmp(3 * (size_t)n);
            for (int a = 0; a < 3; ++a)
                if ((e = hipMemcpyAsync(tmp.data() + (size_t)a * n, adj_in + (AV + a) * astride,
                                        sizeof(float) * n, hipMemcpyDeviceToHost, st)) != hipSuccess)
                    break;
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e == hipSuccess)
                for (uint32_t p = 0; p < n; ++p)
                    for (int a = 0; a < 3; ++a) res.grad_v0[3 * (size_t)p + a] = tmp[(size_t)a * n + p];
        }
        if (e == hipSuccess && res.grad_x0 && n) {
            std::vector<float> tmp(3 * (size_t)n);
            for (int a = 0; a < 3; ++a)
                if ((e = hipMemcpyAsync(tmp.data() + (size_t)a * n, adj_in + (AX + a) * astride,
                                        sizeof(float) * n, hipMemcpyDeviceToHost, st)) != hipSuccess)
                    break;
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e == hipSuccess)
                for (uint32_t p = 0; p < n; ++p)
                    for (int a = 0; a < 3; ++a) res.grad_x0[3 * (size_t)p + a] = tmp[(size_t)a * n + p];
        }
        float th[2] = {0.0f, 0.0f};
        if (e == hipSuccess) e = hipMemcpyAsync(th, gtheta, sizeof(th), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        res.grad_stiffness = th[0];
        res.grad_viscosity = th[1];
        // leave the simulator at the final state of the rollout (PosVel too)
        if (e == hipSuccess) e = restore(ckpt + per_ck * nseg);
        if (e == hipSuccess) e = launch_stage(5, stage_view(), st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    } while (false);
    cleanup();
    if (saved_tuning.path != tun_.path) (void)set_tuning(saved_tuning);
    return e;
}

}  // namespace mpm
