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
#include "mpm_index.h"
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

// ---- per-particle adjoint bodies ----------------------------------------------
// The adjoint is compared with finite differences, not bit for bit, so its
// own arithmetic may contract multiply-adds (the forward recomputation and
// the weights above keep the forward's rounding).
#pragma clang fp contract(fast)

// Written once over an accessor A that is either the chunk's LDS tile window
// (every node inside the grid) or the global grid (wild particles: nodes
// outside [0, G) are skipped, like the forward).

struct AdjArgs {
    const SceneDesc *sd;           // the scene (device)
    const float *st;               // taped binned state of the substep
    const uint32_t *id;
    uint32_t stride;
    const TileChunk *tchunks;      // taped tile chunks of the substep
    const uint32_t *ntchunks;
    const uint2 *tile_chunk;
    uint32_t tchunk_cap;
    const uint32_t *nbr;
    const uint4 *tile_info;
    uint32_t stiles;
    const GridLin *lin;
    uint32_t G;
    float *gvel, *gmv, *gm;        // per cell: 3, 3, 1 floats
    float4 *partial;               // kWinCells per tile chunk
    const float *adj_in;
    float *adj_out;
    uint32_t astride;
    float *gtheta;
    Params prm;
    Uniforms uni;
    int xcd_remap;
    uint32_t walk_grid;            // blocks of the chunk kernels (= their gridDim.x)
};

// G2P adjoint of one particle, in two phases: (a) the forward recomputation
// and the particle-side adjoints, (b) the reverse gather (scatter onto the
// grid).  Between them the chunk kernel bounds the scatter values to pick a
// fixed-point scale.
struct G2PAdj {
    StencilD s;
    float gv[3], gxp[3], B_[3][3];
    uint32_t o;
};

template <class A>
__device__ __forceinline__ void adj_g2p_a(const AdjArgs &q, uint32_t i, const A &acc, G2PAdj &p) {
    p.o = q.id[i];
    float x[3], vold[3], C[3][3];
    load_state(q.st, q.stride, i, x, vold, C);
    stencil_d(x, p.s);
    const StencilD &s = p.s;
    float v[3] = {0.0f, 0.0f, 0.0f};  // forward G2P recomputation
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                const float w = (s.w[gx][0] * s.w[gy][1]) * s.w[gz][2];
                const float3 vg = acc.vg(s, gx, gy, gz);
                v[0] = v[0] + vg.x * w;
                v[1] = v[1] + vg.y * w;
                v[2] = v[2] + vg.z * w;
            }
    const float dt = q.prm.dt, k = 3.0f, ws = 0.3f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            p.B_[c][r] = 4.0f * q.adj_in[(AC + 3 * c + r) * q.astride + p.o];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float rb = q.uni.real_box[a];
        const float x1 = x[a] + v[a] * dt;
        const float xc = fminf(fmaxf(x1, 1.0f), rb - 2.0f);
        const float xn = xc + (v[a] * dt) * k;
        const float gvo = q.adj_in[(AV + a) * q.astride + p.o];
        const float gxo = q.adj_in[(AX + a) * q.astride + p.o];
        float act = 0.0f;
        if (xn < 3.0f) act += 1.0f;
        if (xn > rb - 4.0f) act += 1.0f;
        const float gxn = -ws * gvo * act;
        const float gxc = gxo + gxn;
        const float gx1 = (x1 > 1.0f && x1 < rb - 2.0f) ? gxc : 0.0f;
        p.gv[a] = gvo + gxn * dt * k + gx1 * dt;
        p.gxp[a] = gx1;
    }
}

// Bound on |w g_r| over the particle's nodes: w <= 1, |dist| <= 1.5.
__device__ __forceinline__ float adj_g2p_bound(const G2PAdj &p) {
    float b = 0.0f;
#pragma unroll
    for (int r = 0; r < 3; ++r)
        b = fmaxf(b, fabsf(p.gv[r]) +
                         1.5f * (fabsf(p.B_[0][r]) + fabsf(p.B_[1][r]) + fabsf(p.B_[2][r])));
    return b;
}

template <class A>
__device__ __forceinline__ void adj_g2p_b(const AdjArgs &q, const A &acc, G2PAdj &p) {
    const StencilD &s = p.s;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                const float3 vg3 = acc.vg(s, gx, gy, gz);
                const float vg[3] = {vg3.x, vg3.y, vg3.z};
                float coef = 0.0f;  // d(scalar)/dw
                float add[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    float bd = 0.0f;
#pragma unroll
                    for (int c = 0; c < 3; ++c) bd += p.B_[c][r] * d[c];
                    const float g = p.gv[r] + bd;  // adjoint of v_i[r] per unit weight
                    add[r] = w * g;
                    coef += vg[r] * g;
                }
                acc.add_gvel(s, gx, gy, gz, add);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float bv = 0.0f;  // d dist[a] / dx[a] = -1
#pragma unroll
                    for (int r = 0; r < 3; ++r) bv += p.B_[a][r] * vg[r];
                    p.gxp[a] += dw[a] * coef - w * bv;
                }
            }
#pragma unroll
    for (int a = 0; a < 3; ++a) q.adj_out[(AX + a) * q.astride + p.o] = p.gxp[a];
}

// P2G-2 adjoint of one particle: (a) everything up to the density adjoint
// grho (stores the C adjoint), (b) the density scatter onto the masses.
struct P2G2Adj {
    StencilD s;
    float grho, gxp[3];
    uint32_t o;
};

template <class A>
__device__ __forceinline__ void adj_p2g2_a(const AdjArgs &q, uint32_t i, const A &acc, P2G2Adj &p,
                                           float &gk, float &gmu) {
    p.o = q.id[i];
    const Params prm = q.prm;
    float x[3], v[3], C[3][3];
    load_state(q.st, q.stride, i, x, v, C);
    stencil_d(x, p.s);
    const StencilD &s = p.s;
    float rho = 0.0f;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                rho = rho + acc.m(s, gx, gy, gz) * ((s.w[gx][0] * s.w[gy][1]) * s.w[gz][2]);
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
    p.gxp[0] = p.gxp[1] = p.gxp[2] = 0.0f;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                const float3 g3 = acc.gmv3(s, gx, gy, gz);
                const float g[3] = {g3.x, g3.y, g3.z};
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
                    p.gxp[a] += dw[a] * coef - w * mg;
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
    const float gP = -(gS[0][0] + gS[1][1] + gS[2][2]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            q.adj_out[(AC + 3 * c + rr) * q.astride + p.o] = prm.viscosity * (gS[c][rr] + gS[rr][c]);
            gmu += gS[c][rr] * (C[c][rr] + C[rr][c]);
        }
    p.grho = -gvol / (rho * rho);
    if (pos) {
        p.grho += gP * prm.stiffness * 5.0f * r4 / prm.rest_density;
        gk += gP * (r4 * r - 1.0f);
    }
}

// density = sum m_i w: scatter to the masses, weight terms to x
template <class A>
__device__ __forceinline__ void adj_p2g2_b(const AdjArgs &q, const A &acc, P2G2Adj &p) {
    const StencilD &s = p.s;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                acc.add_gm(s, gx, gy, gz, p.grho * w);
                const float m = acc.m(s, gx, gy, gz);
#pragma unroll
                for (int a = 0; a < 3; ++a) p.gxp[a] += p.grho * m * dw[a];
            }
#pragma unroll
    for (int a = 0; a < 3; ++a) q.adj_out[(AX + a) * q.astride + p.o] += p.gxp[a];
}

template <class A>
__device__ __forceinline__ void adj_p2g1_particle(const AdjArgs &q, uint32_t i, const A &acc) {
    const uint32_t o = q.id[i];
    float x[3], v[3], C[3][3];
    load_state(q.st, q.stride, i, x, v, C);
    StencilD s;
    stencil_d(x, s);
    float gvp[3] = {0.0f, 0.0f, 0.0f}, gxp[3] = {0.0f, 0.0f, 0.0f};
    float gC[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                if (!acc.ok(s, gx, gy, gz)) continue;
                float w, dw[3];
                weight_grad(s, gx, gy, gz, w, dw);
                const float d[3] = {s.dist[0][gx], s.dist[1][gy], s.dist[2][gz]};
                const float3 g3 = acc.gmv3(s, gx, gy, gz);
                const float g[3] = {g3.x, g3.y, g3.z};
                float coef = acc.gm1(s, gx, gy, gz);
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float qd = C[0][r] * d[0] + C[1][r] * d[1] + C[2][r] * d[2];
                    coef += g[r] * (v[r] + qd);
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
        q.adj_out[(AX + a) * q.astride + o] += gxp[a];
        q.adj_out[(AV + a) * q.astride + o] = gvp[a];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) q.adj_out[(AC + 3 * c + r) * q.astride + o] += gC[c][r];
}

// ---- accessors ----------------------------------------------------------------

struct GlobalAcc {  // wild particles: the whole grid, float atomics
    const GridLin *lin;
    float *gvel, *gmv, *gm;
    uint32_t G;
    int32_t Y, Z;
    __device__ __forceinline__ explicit GlobalAcc(const AdjArgs &q)
        : lin(q.lin), gvel(q.gvel), gmv(q.gmv), gm(q.gm), G(q.G), Y(q.uni.Y), Z(q.uni.Z) {}
    __device__ __forceinline__ int32_t idx(const StencilD &s, int gx, int gy, int gz) const {
        return sidx(s, gx, gy, gz, Y, Z);
    }
    __device__ __forceinline__ bool ok(const StencilD &s, int gx, int gy, int gz) const {
        return in_grid(idx(s, gx, gy, gz), G);
    }
    __device__ __forceinline__ float3 vg(const StencilD &s, int gx, int gy, int gz) const {
        const GridLin &L = lin[idx(s, gx, gy, gz)];
        return make_float3(L.vg[0], L.vg[1], L.vg[2]);
    }
    __device__ __forceinline__ float m(const StencilD &s, int gx, int gy, int gz) const {
        return lin[idx(s, gx, gy, gz)].m;
    }
    __device__ __forceinline__ float3 gmv3(const StencilD &s, int gx, int gy, int gz) const {
        const size_t i = 3 * (size_t)idx(s, gx, gy, gz);
        return make_float3(gmv[i], gmv[i + 1], gmv[i + 2]);
    }
    __device__ __forceinline__ float gm1(const StencilD &s, int gx, int gy, int gz) const {
        return gm[idx(s, gx, gy, gz)];
    }
    __device__ __forceinline__ void add_gvel(const StencilD &s, int gx, int gy, int gz,
                                             const float a[3]) const {
        const size_t i = 3 * (size_t)idx(s, gx, gy, gz);
        unsafeAtomicAdd(&gvel[i], a[0]);
        unsafeAtomicAdd(&gvel[i + 1], a[1]);
        unsafeAtomicAdd(&gvel[i + 2], a[2]);
    }
    __device__ __forceinline__ void add_gm(const StencilD &s, int gx, int gy, int gz,
                                           float a) const {
        unsafeAtomicAdd(&gm[idx(s, gx, gy, gz)], a);
    }
};

// The chunk's 6^3 tile window in LDS: w4 holds per-cell inputs (vg or gmv in
// xyz, mass or gm in w), acc per-cell accumulators (x, y, z) in i32 fixed
// point at `scale` (LDS float atomics are far slower than integer ones).
struct WinAcc {
    const float4 *w4;
    int *acc;  // 3 x kWinCells
    int l0;
    float scale;
    __device__ __forceinline__ int fx(float a) const { return (int)rintf(a * scale); }
    __device__ __forceinline__ int l(int gx, int gy, int gz) const {
        return l0 + (gx * kWinEdge + gy) * kWinEdge + gz;
    }
    __device__ __forceinline__ bool ok(const StencilD &, int, int, int) const { return true; }
    __device__ __forceinline__ float3 vg(const StencilD &, int gx, int gy, int gz) const {
        const float4 c = w4[l(gx, gy, gz)];
        return make_float3(c.x, c.y, c.z);
    }
    __device__ __forceinline__ float3 gmv3(const StencilD &s, int gx, int gy, int gz) const {
        return vg(s, gx, gy, gz);
    }
    __device__ __forceinline__ float m(const StencilD &, int gx, int gy, int gz) const {
        return w4[l(gx, gy, gz)].w;
    }
    __device__ __forceinline__ float gm1(const StencilD &s, int gx, int gy, int gz) const {
        return m(s, gx, gy, gz);
    }
    __device__ __forceinline__ void add_gvel(const StencilD &, int gx, int gy, int gz,
                                             const float a[3]) const {
        const int c = l(gx, gy, gz);
        atomicAdd(&acc[c], fx(a[0]));
        atomicAdd(&acc[kWinCells + c], fx(a[1]));
        atomicAdd(&acc[2 * kWinCells + c], fx(a[2]));
    }
    __device__ __forceinline__ void add_gm(const StencilD &, int gx, int gy, int gz,
                                           float a) const {
        atomicAdd(&acc[l(gx, gy, gz)], fx(a));
    }
};

// Fixed-point scale for one round of at most kAdjBlock particles whose
// scatter values are bounded by `bound`: each i32 cell sum stays below 2^30
// and the rounding error of a value is at most 2^-23 of the bound.
__device__ __forceinline__ float round_scale(float bound) {
    if (!(bound > 0.0f) || !isfinite(bound)) return 1.0f;
    int e;
    (void)frexpf(bound, &e);  // bound < 2^e
    return ldexpf(1.0f, min(126, max(-126, 22 - e)));
}

// Block-wide maximum of non-negative values (their bit patterns order like
// unsigned integers); *slot must be zero before the first call.
__device__ __forceinline__ float block_max(float v, uint32_t *slot) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(slot, __float_as_uint(v));
    __syncthreads();
    return __uint_as_float(*slot);
}

__device__ __forceinline__ int win_l0_of(const float *st, uint32_t stride, uint32_t i) {
    const int m = kTileEdge - 1;
    const int cx = (int)floorf(st[(PX + 0) * stride + i]);
    const int cy = (int)floorf(st[(PX + 1) * stride + i]);
    const int cz = (int)floorf(st[(PX + 2) * stride + i]);
    return ((cx & m) * kWinEdge + (cy & m)) * kWinEdge + (cz & m);
}

// Grid cell of window cell l of tile t, or -1 outside the scene.
__device__ __forceinline__ int64_t adj_win_cell(const AdjArgs &q, int l, const int t[3]) {
    const SceneDesc *sd = q.sd;
    const int x = t[0] * kTileEdge - 1 + l / (kWinEdge * kWinEdge);
    const int y = t[1] * kTileEdge - 1 + (l / kWinEdge) % kWinEdge;
    const int z = t[2] * kTileEdge - 1 + l % kWinEdge;
    if ((uint32_t)x >= (uint32_t)sd->kdim[0] || (uint32_t)y >= (uint32_t)sd->kdim[1] ||
        (uint32_t)z >= (uint32_t)sd->kdim[2])
        return -1;
    return (int64_t)((uint32_t)x * (uint32_t)q.uni.Y * (uint32_t)q.uni.Z +
                     (uint32_t)y * (uint32_t)q.uni.Z + (uint32_t)z);
}

// Persistent walk over the taped tile chunks: runs of 16 consecutive chunks
// dealt round robin over the XCDs (block b runs on XCD b mod 8), or with
// xcd_remap one contiguous range per XCD.
struct AdjWalk {
    uint32_t c, end, step;
};
__device__ __forceinline__ AdjWalk adj_walk(const AdjArgs &q) {
    const uint32_t n = min(*q.ntchunks, q.tchunk_cap);
    AdjWalk w;
    if (q.xcd_remap && (gridDim.x % 8) == 0) {
        const uint32_t per = (n + 7) / 8, x = blockIdx.x % 8;
        w.c = x * per + blockIdx.x / 8;
        w.end = min(n, (x + 1) * per);
        w.step = gridDim.x / 8;
    } else {
        w.c = walk_first_of(blockIdx.x, gridDim.x);
        w.end = n;
        w.step = gridDim.x;
    }
    return w;
}

// The walk of a chunk kernel's window loop with the descriptor of its current
// chunk.  In the default walk the first chunk's index does not depend on the
// chunk count, so the count, that descriptor and the scene's tile counts are
// loaded together and waited for once (scalar loads: the tape and the scene
// are not written while the kernel runs); each iteration loads the next
// descriptor before its own loads.
struct AdjChunkWalk {
    AdjWalk w;
    TileChunk ch;   // descriptor of chunk w.c (meaningful while w.c < w.end)
    uint32_t knt1, knt2;

    __device__ __forceinline__ explicit AdjChunkWalk(const AdjArgs &q) {
        const uint32_t grid = q.walk_grid;  // = gridDim.x
        w.c = walk_first_of(blockIdx.x, grid);
        ch = load_const(q.tchunks + spec_chunk_of(w.c, q.tchunk_cap));
        const uint32_t n = load_const(q.ntchunks);
        knt1 = load_const(&q.sd->knt[1]);
        knt2 = load_const(&q.sd->knt[2]);
        // (an empty asm that needs the values: their loads complete here together)
        asm volatile("" ::"s"(n), "s"(ch.x), "s"(ch.y), "s"(ch.z), "s"(ch.w), "s"(knt1),
                     "s"(knt2));
        w.end = min(n, q.tchunk_cap);
        w.step = grid;
        if (q.xcd_remap && (grid % 8) == 0) {  // contiguous ranges: another round of loads
            const uint32_t per = (w.end + 7) / 8, x = blockIdx.x % 8;
            w.c = x * per + blockIdx.x / 8;
            w.end = min(w.end, (x + 1) * per);
            w.step = grid / 8;
            ch = load_const(q.tchunks + spec_chunk_of(w.c, q.tchunk_cap));
        }
    }
    // Returns the current chunk's descriptor and starts loading the next one.
    __device__ __forceinline__ TileChunk take(const AdjArgs &q) {
        const TileChunk cur = ch;
        const uint32_t nc = w.c + w.step;
        if (nc < w.end) ch = load_const(q.tchunks + nc);
        return cur;
    }
    __device__ __forceinline__ void tile_coords(uint32_t xm, int t[3]) const {
        t[0] = (int)(xm / (knt1 * knt2));
        t[1] = (int)((xm / knt2) % knt1);
        t[2] = (int)(xm % knt2);
    }
};

// ---- chunk kernels ---------------------------------------------------------------

// G2P adjoint: window of vg in, window of grid-velocity adjoints out
// (partial).  Rounds of up to 256 particles, one per thread; each round
// accumulates in i32 fixed point at its own scale and is folded into the
// float window by the cell owners.
__global__ __launch_bounds__(kAdjBlock) void k_adj_g2p_t(AdjArgs q) {
    __shared__ float4 s_w4[kWinCells];
    __shared__ int s_acc[3 * kWinCells];
    __shared__ float s_sum[3 * kWinCells];
    __shared__ uint32_t s_max;
    bool wild_seen = false;  // block-uniform: a wild chunk among this block's
    for (AdjChunkWalk cw(q); cw.w.c < cw.w.end; cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(q);
        if (ch.y == kWildTile) {  // second loop
            wild_seen = true;
            continue;
        }
        int t[3];
        cw.tile_coords(ch.y, t);
        for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock) {
            const int64_t gi = adj_win_cell(q, l, t);
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gi >= 0) {
                const GridLin &L = q.lin[gi];
                c = make_float4(L.vg[0], L.vg[1], L.vg[2], 0.0f);
            }
            s_w4[l] = c;
        }
        for (int l = threadIdx.x; l < 3 * kWinCells; l += kAdjBlock) {
            s_acc[l] = 0;
            s_sum[l] = 0.0f;
        }
        if (threadIdx.x == 0) s_max = 0u;
        __syncthreads();
        for (uint32_t r0 = 0; r0 < ch.w; r0 += kAdjBlock) {
            // wave tasks of 64 lanes spread over the round: in cell-sorted
            // storage the lanes of one LDS add rarely share a cell
            const uint32_t cnt = min((uint32_t)kAdjBlock, ch.w - r0);
            const uint32_t tasks = (cnt + 63) >> 6, task = threadIdx.x >> 6;
            const uint32_t j = r0 + task + tasks * (threadIdx.x & 63);
            const bool active = task < tasks && j < r0 + cnt;
            const uint32_t i = ch.z + (active ? j : r0);
            const WinAcc pre{s_w4, s_acc, win_l0_of(q.st, q.stride, i), 1.0f};
            G2PAdj p;
            if (active) adj_g2p_a(q, i, pre, p);
            const float scale = round_scale(block_max(active ? adj_g2p_bound(p) : 0.0f, &s_max));
            if (active) adj_g2p_b(q, WinAcc{s_w4, s_acc, pre.l0, scale}, p);
            __syncthreads();
            const float inv = 1.0f / scale;
            for (int l = threadIdx.x; l < 3 * kWinCells; l += kAdjBlock) {
                s_sum[l] += (float)s_acc[l] * inv;
                s_acc[l] = 0;
            }
            if (threadIdx.x == 0) s_max = 0u;
            __syncthreads();
        }
        float4 *out = q.partial + (size_t)cw.w.c * kWinCells;
        for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock)
            out[l] = make_float4(s_sum[l], s_sum[kWinCells + l], s_sum[2 * kWinCells + l], 0.0f);
        __syncthreads();
    }
    // Wild chunks (global float atomics) in a loop of their own: in the
    // window loop they would set its register budget.
    if (!wild_seen) return;
    for (AdjWalk w = adj_walk(q); w.c < w.end; w.c += w.step) {
        const TileChunk ch = q.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const GlobalAcc acc(q);
        for (uint32_t j = threadIdx.x; j < ch.w; j += kAdjBlock) {
            G2PAdj p;
            adj_g2p_a(q, ch.z + j, acc, p);
            adj_g2p_b(q, acc, p);
        }
    }
}

// P2G-2 adjoint: window of (gmv, mass) in, window of mass adjoints out
// (fixed-point rounds as in k_adj_g2p_t).
__global__ __launch_bounds__(kAdjBlock) void k_adj_p2g2_t(AdjArgs q) {
    __shared__ float4 s_w4[kWinCells];
    __shared__ int s_acc[kWinCells];
    __shared__ float s_sum[kWinCells];
    __shared__ uint32_t s_max;
    __shared__ float s_red[2][kAdjBlock / 64];
    float gk = 0.0f, gmu = 0.0f;
    bool wild_seen = false;  // block-uniform: a wild chunk among this block's
    for (AdjChunkWalk cw(q); cw.w.c < cw.w.end; cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(q);
        if (ch.y == kWildTile) {  // second loop
            wild_seen = true;
            continue;
        }
        int t[3];
        cw.tile_coords(ch.y, t);
        for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock) {
            const int64_t gi = adj_win_cell(q, l, t);
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gi >= 0)
                c = make_float4(q.gmv[3 * gi], q.gmv[3 * gi + 1], q.gmv[3 * gi + 2], q.lin[gi].m);
            s_w4[l] = c;
            s_acc[l] = 0;
            s_sum[l] = 0.0f;
        }
        if (threadIdx.x == 0) s_max = 0u;
        __syncthreads();
        for (uint32_t r0 = 0; r0 < ch.w; r0 += kAdjBlock) {
            const uint32_t cnt = min((uint32_t)kAdjBlock, ch.w - r0);
            const uint32_t tasks = (cnt + 63) >> 6, task = threadIdx.x >> 6;
            const uint32_t j = r0 + task + tasks * (threadIdx.x & 63);
            const bool active = task < tasks && j < r0 + cnt;
            const uint32_t i = ch.z + (active ? j : r0);
            const WinAcc pre{s_w4, s_acc, win_l0_of(q.st, q.stride, i), 1.0f};
            P2G2Adj p;
            if (active) adj_p2g2_a(q, i, pre, p, gk, gmu);
            const float scale = round_scale(block_max(active ? fabsf(p.grho) : 0.0f, &s_max));
            if (active) adj_p2g2_b(q, WinAcc{s_w4, s_acc, pre.l0, scale}, p);
            __syncthreads();
            const float inv = 1.0f / scale;
            for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock) {
                s_sum[l] += (float)s_acc[l] * inv;
                s_acc[l] = 0;
            }
            if (threadIdx.x == 0) s_max = 0u;
            __syncthreads();
        }
        float4 *out = q.partial + (size_t)cw.w.c * kWinCells;
        for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock)
            out[l] = make_float4(s_sum[l], 0.0f, 0.0f, 0.0f);
        __syncthreads();
    }
    // Wild chunks (global float atomics) in a loop of their own: in the
    // window loop they would set its register budget.  (No early return:
    // the block reduction below needs every block.)
    for (AdjWalk w = adj_walk(q); wild_seen && w.c < w.end; w.c += w.step) {
        const TileChunk ch = q.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const GlobalAcc acc(q);
        for (uint32_t j = threadIdx.x; j < ch.w; j += kAdjBlock) {
            P2G2Adj p;
            adj_p2g2_a(q, ch.z + j, acc, p, gk, gmu);
            adj_p2g2_b(q, acc, p);
        }
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
        if (a != 0.0f) unsafeAtomicAdd(&q.gtheta[0], a);
        if (b != 0.0f) unsafeAtomicAdd(&q.gtheta[1], b);
    }
}

// P2G-1 adjoint: window of (gmv, gm) in, particle adjoints out.
__global__ __launch_bounds__(kAdjBlock) void k_adj_p2g1_t(AdjArgs q) {
    __shared__ float4 s_w4[kWinCells];
    bool wild_seen = false;  // block-uniform: a wild chunk among this block's
    for (AdjChunkWalk cw(q); cw.w.c < cw.w.end; cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(q);
        if (ch.y == kWildTile) {  // second loop
            wild_seen = true;
            continue;
        }
        int t[3];
        cw.tile_coords(ch.y, t);
        for (int l = threadIdx.x; l < kWinCells; l += kAdjBlock) {
            const int64_t gi = adj_win_cell(q, l, t);
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gi >= 0)
                c = make_float4(q.gmv[3 * gi], q.gmv[3 * gi + 1], q.gmv[3 * gi + 2], q.gm[gi]);
            s_w4[l] = c;
        }
        __syncthreads();
        // wave tasks of 64 lanes spread over the chunk: in cell-sorted storage
        // the lanes of one LDS add rarely share a cell
        const uint32_t tasks = (ch.w + 63) >> 6;
        for (uint32_t t = threadIdx.x >> 6; t < tasks; t += kAdjBlock / 64) {
            const uint32_t j = t + tasks * (threadIdx.x & 63);
            if (j >= ch.w) continue;
            const uint32_t i = ch.z + j;
            const WinAcc acc{s_w4, nullptr, win_l0_of(q.st, q.stride, i), 1.0f};
            adj_p2g1_particle(q, i, acc);
        }
        __syncthreads();
    }
    // Wild chunks (global float atomics) in a loop of their own: in the
    // window loop they would set its register budget.
    if (!wild_seen) return;
    for (AdjWalk w = adj_walk(q); w.c < w.end; w.c += w.step) {
        const TileChunk ch = q.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const GlobalAcc acc(q);
        for (uint32_t j = threadIdx.x; j < ch.w; j += kAdjBlock)
            adj_p2g1_particle(q, ch.z + j, acc);
    }
}

// Sum of the partial windows covering each of a tile's 64 cells (one wave
// per tile; the same cover as the forward reduce).
// kt: lane q < 27 holds the key of neighbour q of the tile (q.nbr), other
// lanes kNoTile.  All covering slots' loads of one round are in flight at once.
__device__ __forceinline__ float4 adj_sum_partials(const AdjArgs &q, uint32_t kt, const int l[3],
                                                   bool cell_ok) {
    // (no branch around the load: a missing neighbour reads key 0 and takes no chunks)
    const uint2 tc = q.tile_chunk[kt != kNoTile ? kt : 0u];
    const uint32_t nfirst = kt != kNoTile ? tc.x : 0u;
    const uint32_t ncnt =
        kt != kNoTile && tc.x <= q.tchunk_cap && tc.y <= q.tchunk_cap - tc.x ? tc.y : 0u;
    int ext[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) ext[ax] = l[ax] == 0 ? -1 : (l[ax] == kTileEdge - 1 ? 1 : 0);
    uint32_t base[8], cnt[8], rounds = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        bool valid = cell_ok;
        int d[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const bool bit = (s >> ax) & 1;
            d[ax] = bit ? ext[ax] : 0;
            valid = valid && !(bit && ext[ax] == 0);
        }
        const int qq = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1);
        const uint32_t f = __shfl(nfirst, qq, 64), c = __shfl(ncnt, qq, 64);
        const int widx = ((l[0] + 1 - kTileEdge * d[0]) * kWinEdge + (l[1] + 1 - kTileEdge * d[1])) *
                             kWinEdge +
                         (l[2] + 1 - kTileEdge * d[2]);
        base[s] = f * (uint32_t)kWinCells + (uint32_t)widx;
        cnt[s] = valid ? c : 0u;
        rounds = max(rounds, cnt[s]);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) rounds = max(rounds, (uint32_t)__shfl_xor(rounds, o, 64));
    // per slot (its chunks in order), then the slots in order: a fixed order,
    // so the result does not depend on the launch
    float4 sl[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) sl[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t j = 0; j < rounds; ++j) {
        float4 v[8];
#pragma unroll
        for (int s = 0; s < 8; ++s)
            v[s] = j < cnt[s] ? q.partial[(size_t)base[s] + (size_t)j * kWinCells]
                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (j < cnt[s]) {
                sl[s].x += v[s].x;
                sl[s].y += v[s].y;
                sl[s].z += v[s].z;
            }
    }
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        acc.x += sl[s].x;
        acc.y += sl[s].y;
        acc.z += sl[s].z;
    }
    return acc;
}

// The neighbour keys of tile b, one per lane (lanes past 26: kNoTile), loaded
// without a branch around the load.
__device__ __forceinline__ uint32_t adj_nbr_keys(const AdjArgs &q, uint32_t b) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k = q.nbr[(size_t)b * 27 + min(lane, 26u)];
    return lane < 27 ? k : kNoTile;
}

// Grid-update adjoint per tile: grid-velocity adjoint = global scatter (wild
// particles) + the G2P partials, then through the update to (gmv, gm).
__global__ __launch_bounds__(64) void k_adj_update_t(AdjArgs q) {
    const uint32_t b = blockIdx.x;
    if (b >= q.stiles) return;
    // the scene, the tile and its neighbour keys in one round of loads
    const SceneDesc sdv = load_const(q.sd);
    const SceneDesc *sd = &sdv;
    const uint4 ti = load_const(q.tile_info + b);
    const uint32_t kt = adj_nbr_keys(q, b);
    int l[3] = {(int)threadIdx.x >> 4, ((int)threadIdx.x >> 2) & 3, (int)threadIdx.x & 3};
    int64_t gi = -1;
    float4 part = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (sd->fast) {
        const int c[3] = {(int)ti.x * kTileEdge + l[0], (int)ti.y * kTileEdge + l[1],
                          (int)ti.z * kTileEdge + l[2]};
        const bool ok = (uint32_t)c[0] < (uint32_t)sd->kdim[0] &&
                        (uint32_t)c[1] < (uint32_t)sd->kdim[1] &&
                        (uint32_t)c[2] < (uint32_t)sd->kdim[2];
        part = adj_sum_partials(q, kt, l, ok);
        if (ok)
            gi = (int64_t)((uint32_t)c[0] * (uint32_t)q.uni.Y * (uint32_t)q.uni.Z +
                           (uint32_t)c[1] * (uint32_t)q.uni.Z + (uint32_t)c[2]);
    } else {
        const uint32_t li = (b - sd->soff) * kTileCells + threadIdx.x;
        if (li < q.G) gi = li;
    }
    if (gi < 0) return;
    const GridLin L = q.lin[gi];
    const float gsum[3] = {q.gvel[3 * gi] + part.x, q.gvel[3 * gi + 1] + part.y,
                           q.gvel[3 * gi + 2] + part.z};
    float g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g[a] = (L.flags & (1u << a)) ? gsum[a] : 0.0f;
        q.gvel[3 * gi + a] = 0.0f;  // ready for the next substep
    }
    float m_adj = 0.0f;
    if (L.flags & 8u) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            q.gmv[3 * gi + a] = g[a] / L.m;
            m_adj -= g[a] * L.vpre[a] / L.m;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) q.gmv[3 * gi + a] = g[a];
    }
    q.gm[gi] = m_adj;
}

// Adds the P2G-2 partials (mass adjoints) into gm, per tile.
__global__ __launch_bounds__(64) void k_adj_reduce_gm_t(AdjArgs q) {
    const uint32_t b = blockIdx.x;
    if (b >= q.stiles) return;
    // the scene, the tile and its neighbour keys in one round of loads
    const SceneDesc sdv = load_const(q.sd);
    const SceneDesc *sd = &sdv;
    const uint4 ti = load_const(q.tile_info + b);
    const uint32_t kt = adj_nbr_keys(q, b);
    if (!sd->fast) return;
    const int l[3] = {(int)threadIdx.x >> 4, ((int)threadIdx.x >> 2) & 3, (int)threadIdx.x & 3};
    const int c[3] = {(int)ti.x * kTileEdge + l[0], (int)ti.y * kTileEdge + l[1],
                      (int)ti.z * kTileEdge + l[2]};
    const bool ok = (uint32_t)c[0] < (uint32_t)sd->kdim[0] &&
                    (uint32_t)c[1] < (uint32_t)sd->kdim[1] && (uint32_t)c[2] < (uint32_t)sd->kdim[2];
    const float4 part = adj_sum_partials(q, kt, l, ok);
    if (!ok) return;
    const uint64_t gi = (uint32_t)c[0] * (uint32_t)q.uni.Y * (uint32_t)q.uni.Z +
                        (uint32_t)c[1] * (uint32_t)q.uni.Z + (uint32_t)c[2];
    if (part.x != 0.0f) q.gm[gi] += part.x;
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

// Byte offsets of one substep's tape record: the binned state and ids the
// substep's kernels read, its tile chunks, and its raw grid after P2G.
struct TapeLayout {
    size_t state, id, grid, tchunks, tile_chunk, ntchunks, bytes;
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Several device-to-device copies in one launch (sizes in 4-byte words).
struct CopySegs {
    const uint32_t *src[5];
    uint32_t *dst[5];
    uint64_t words[5];
    int count;
};

__global__ __launch_bounds__(256) void k_copy_segs(CopySegs c) {
    for (int s = 0; s < c.count; ++s)
        for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < c.words[s];
             i += (uint64_t)gridDim.x * 256)
            c.dst[s][i] = c.src[s][i];
}

hipError_t Batch::substep_taped(hipStream_t st, char *rec, const void *layout) {
    const TapeLayout &L = *static_cast<const TapeLayout *>(layout);
    BatchArgs a = args();
    TRY(launch_bin(a, bin_buffers(), !keys_ready_, st));
    // the tape holds this substep's input state in bin order
    TRY(launch_gather_binned(a, reinterpret_cast<float *>(rec + L.state),
                             reinterpret_cast<uint32_t *>(rec + L.id), st));
    CopySegs cs{};
    auto seg = [&](size_t at, const void *src, size_t bytes) {
        cs.src[cs.count] = static_cast<const uint32_t *>(src);
        cs.dst[cs.count] = reinterpret_cast<uint32_t *>(rec + at);
        cs.words[cs.count] = bytes / 4;
        ++cs.count;
    };
    seg(L.tchunks, tchunks_, sizeof(TileChunk) * cap_tchunks_);
    seg(L.tile_chunk, tile_chunk_, sizeof(uint2) * total_keys_);
    seg(L.ntchunks, ntchunks_, sizeof(uint32_t));
    k_copy_segs<<<1024, 256, 0, st>>>(cs);
    TRY(hipGetLastError());
    TRY(launch_p2g1(a, st));
    TRY(launch_reduce1(a, st));
    TRY(launch_p2g2(a, st));
    TRY(launch_reduce2(a, st));
    TRY(launch_g2p(a, st));
    TRY(hipMemcpyAsync(rec + L.grid, grid_ + h_scenes_[0].goff,
                       sizeof(int4) * (size_t)h_scenes_[0].G, hipMemcpyDeviceToDevice, st));
    after_substep(st);
    return hipSuccess;
}

hipError_t Batch::backward(const AdjointConfig &cfg, hipStream_t st, AdjointResult &res) {
    if (h_scenes_.size() != 1 || cfg.substeps == 0 || cfg.checkpoint_every == 0)
        return hipErrorInvalidValue;
    // the reverse kernels work on 4^3 tiles
    const Tuning saved_tuning = tun_;
    if (tun_.path != 0 || tun_.tile_bits != kTileBits) {
        Tuning t = tun_;
        t.path = 0;
        t.tile_bits = kTileBits;
        TRY(set_tuning(t, st));
    }
    const SceneDesc &d = h_scenes_[0];
    const uint32_t n = d.n, G = d.G, T = cfg.substeps, K = cfg.checkpoint_every;
    const uint32_t nseg = (T + K - 1) / K;
    const size_t sbytes = sizeof(float) * NCOMP * stride_, ibytes = sizeof(uint32_t) * stride_;
    TapeLayout L{};
    L.state = 0;
    L.id = align256(L.state + sbytes);
    L.grid = align256(L.id + ibytes);
    L.tchunks = align256(L.grid + sizeof(int4) * (size_t)G);
    L.tile_chunk = align256(L.tchunks + sizeof(TileChunk) * cap_tchunks_);
    L.ntchunks = align256(L.tile_chunk + sizeof(uint2) * total_keys_);
    L.bytes = align256(L.ntchunks + sizeof(uint32_t));
    // Device workspace (kept between calls): checkpoints (state + id) per
    // segment start, the segment tape, grid linearisation, adjoints.
    hipError_t e = hipSuccess;
    const size_t per_ck = sbytes + ibytes;
    const uint32_t astride = std::max<uint32_t>(n, 1);
    const size_t g1 = std::max<uint32_t>(G, 1);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    const size_t o_ckpt = carve(per_ck * (nseg + 1)), o_tape = carve(L.bytes * K),
                 o_lin = carve(sizeof(GridLin) * g1), o_adjA = carve(sizeof(float) * ANCOMP * astride),
                 o_adjB = carve(sizeof(float) * ANCOMP * astride), o_gvel = carve(sizeof(float) * 3 * g1),
                 o_gmv = carve(sizeof(float) * 3 * g1), o_gm = carve(sizeof(float) * g1),
                 o_gtheta = carve(sizeof(float) * 2), o_lacc = carve(sizeof(double) * 4);
    if (off > adj_ws_cap_) {
        if (adj_ws_) (void)hipFree(adj_ws_);
        adj_ws_ = nullptr;
        adj_ws_cap_ = 0;
        TRY(hipMalloc(&adj_ws_, off));
        adj_ws_cap_ = off;
    }
    char *const ckpt = adj_ws_ + o_ckpt, *const tape = adj_ws_ + o_tape;
    GridLin *const lin = reinterpret_cast<GridLin *>(adj_ws_ + o_lin);
    float *const adjA = reinterpret_cast<float *>(adj_ws_ + o_adjA);
    float *const adjB = reinterpret_cast<float *>(adj_ws_ + o_adjB);
    float *const gvel = reinterpret_cast<float *>(adj_ws_ + o_gvel);
    float *const gmv = reinterpret_cast<float *>(adj_ws_ + o_gmv);
    float *const gm = reinterpret_cast<float *>(adj_ws_ + o_gm);
    float *const gtheta = reinterpret_cast<float *>(adj_ws_ + o_gtheta);
    double *const lacc = reinterpret_cast<double *>(adj_ws_ + o_lacc);
    do {
        if ((e = hipMemsetAsync(gvel, 0, sizeof(float) * 3 * g1, st)) != hipSuccess) break;
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
        // forward rollout with checkpoints every K substeps; the last segment
        // is taped on the way, so the backward pass need not recompute it
        const uint32_t last0 = (nseg - 1) * K;
        for (uint32_t t = 0; t < T && e == hipSuccess; ++t) {
            if (t % K == 0) e = save(ckpt + per_ck * (t / K));
            if (e == hipSuccess)
                e = t >= last0 ? substep_taped(st, tape + L.bytes * (t - last0), &L) : substep(st);
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
        const Uniforms *uni_d = reinterpret_cast<const Uniforms *>(
            reinterpret_cast<const char *>(d_scenes_) + offsetof(SceneDesc, uni));
        AdjArgs q{};
        q.sd = d_scenes_;
        q.stride = stride_;
        q.tchunk_cap = tchunk_cap();  // (4^3 tiles here)
        q.nbr = d_nbr_;
        q.tile_info = d_tile_info_;
        q.stiles = total_tiles_;
        q.lin = lin;
        q.G = G;
        q.gvel = gvel;
        q.gmv = gmv;
        q.gm = gm;
        q.partial = reinterpret_cast<float4 *>(partial_);
        q.astride = astride;
        q.gtheta = gtheta;
        q.prm = d.prm;
        q.uni = d.uni;
        q.xcd_remap = tun_.xcd_remap;
        const unsigned nchunk_blocks = 2048;
        q.walk_grid = nchunk_blocks;
        for (int32_t sgi = (int32_t)nseg - 1; sgi >= 0 && e == hipSuccess; --sgi) {
            const uint32_t t0 = (uint32_t)sgi * K, t1 = std::min(T, t0 + K);
            if (sgi != (int32_t)nseg - 1) {  // (the last segment's tape is the forward's)
                if ((e = restore(ckpt + per_ck * sgi)) != hipSuccess) break;
                for (uint32_t t = t0; t < t1 && e == hipSuccess; ++t)
                    e = substep_taped(st, tape + L.bytes * (t - t0), &L);
            }
            for (int32_t t = (int32_t)t1 - 1; t >= (int32_t)t0 && e == hipSuccess; --t) {
                const char *rec = tape + L.bytes * (t - t0);
                q.st = reinterpret_cast<const float *>(rec + L.state);
                q.id = reinterpret_cast<const uint32_t *>(rec + L.id);
                q.tchunks = reinterpret_cast<const TileChunk *>(rec + L.tchunks);
                q.tile_chunk = reinterpret_cast<const uint2 *>(rec + L.tile_chunk);
                q.ntchunks = reinterpret_cast<const uint32_t *>(rec + L.ntchunks);
                q.adj_in = adj_in;
                q.adj_out = adj_out;
                if (G)
                    k_grid_lin<<<nblocks(G), kAdjBlock, 0, st>>>(
                        reinterpret_cast<const int4 *>(rec + L.grid), G, uni_d, d.prm, lin);
                if (n) k_adj_g2p_t<<<nchunk_blocks, kAdjBlock, 0, st>>>(q);
                if (total_tiles_) k_adj_update_t<<<total_tiles_, 64, 0, st>>>(q);
                if (n) k_adj_p2g2_t<<<nchunk_blocks, kAdjBlock, 0, st>>>(q);
                if (total_tiles_) k_adj_reduce_gm_t<<<total_tiles_, 64, 0, st>>>(q);
                if (n) k_adj_p2g1_t<<<nchunk_blocks, kAdjBlock, 0, st>>>(q);
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
    if (saved_tuning.path != tun_.path || saved_tuning.tile_bits != tun_.tile_bits)
        (void)set_tuning(saved_tuning, st);
    return e;
}

}  // namespace mpm
