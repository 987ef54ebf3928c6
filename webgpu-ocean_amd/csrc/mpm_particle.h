// mpm_particle.h -- per-particle bodies of p2g_1, p2g_2 and g2p, written
// once over a "sink" that is either an LDS window or the global grid.
//
// Per-axis quantities of the 3x3x3 stencil (weights, cell_dist, integer cell
// coordinates) and the per-axis products C[c][r] * cell_dist[c] are computed
// once per particle; the per-node values are then formed with the same
// operations, operands and order as the WGSL (mls-mpm/p2g_1.wgsl:41-60,
// p2g_2.wgsl:46-89, g2p.wgsl:43-67), so every rounding matches.
#pragma once
#include "mpm_device.h"

#pragma clang fp contract(off)

namespace mpm {

struct Node3 {
    float w[3][3];     // w[g][axis]                         (p2g_1.wgsl:32-34)
    float dist[3][3];  // dist[axis][g] = (cell_x + 0.5) - x   (p2g_1.wgsl:47)
    int32_t ic[3][3];  // ic[axis][g] = i32(cell_x)            (p2g_1.wgsl:53-56)
};

__device__ __forceinline__ void node3_init(const float x[3], Node3 &n) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ci = floorf(x[a]);
        const float d = x[a] - (ci + 0.5f);
        n.w[0][a] = (0.5f * (0.5f - d)) * (0.5f - d);
        n.w[1][a] = 0.75f - d * d;
        n.w[2][a] = (0.5f * (0.5f + d)) * (0.5f + d);
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const float cx = (ci + (float)g) - 1.0f;
            n.dist[a][g] = (cx + 0.5f) - x[a];
            n.ic[a][g] = f2i_sat(cx);
        }
    }
}

// One stencil node along one axis: weight, cell_dist, integer cell coordinate.
struct AxisNode {
    float w, dist;
    int32_t ic;
};

__device__ __forceinline__ AxisNode axis_node(const Node3 &n, int axis, int g) {
    return AxisNode{n.w[g][axis], n.dist[axis][g], n.ic[axis][g]};
}

// Linear cell index of node (ax, ay, gz) (p2g_1.wgsl:53-56).
__device__ __forceinline__ int32_t node_index(const Node3 &n, const AxisNode &ax,
                                              const AxisNode &ay, int gz, int32_t Y, int32_t Z) {
    return (int32_t)((uint32_t)ax.ic * (uint32_t)Y * (uint32_t)Z + (uint32_t)ay.ic * (uint32_t)Z +
                     (uint32_t)n.ic[2][gz]);
}

// Visits the 9 (x, y) rows of the stencil, x-major like the WGSL loops,
// calling row(ax, ay, gx, gy); each row then walks its 3 z nodes.  kRolled
// keeps both loops rolled, the node values rotating through registers rather
// than being indexed: far fewer live registers, for the rare global-memory
// paths that share a kernel with the LDS-window path.  Both forms perform the
// same operations in the same order.
template <bool kRolled, class Row>
__device__ __forceinline__ void for_rows(const Node3 &n, Row &&row) {
    if constexpr (kRolled) {
        AxisNode x0 = axis_node(n, 0, 0), x1 = axis_node(n, 0, 1), x2 = axis_node(n, 0, 2);
#pragma unroll 1
        for (int gx = 0; gx < 3; ++gx) {
            AxisNode y0 = axis_node(n, 1, 0), y1 = axis_node(n, 1, 1), y2 = axis_node(n, 1, 2);
#pragma unroll 1
            for (int gy = 0; gy < 3; ++gy) {
                row(x0, y0, gx, gy);
                y0 = y1;
                y1 = y2;
            }
            x0 = x1;
            x1 = x2;
        }
    } else {
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy) row(axis_node(n, 0, gx), axis_node(n, 1, gy), gx, gy);
    }
}

// ---- p2g_1 (mls-mpm/p2g_1.wgsl:36-63) ---------------------------------------

template <bool kRolled = false, class Sink>
__device__ __forceinline__ void p2g1_particle(const Node3 &n, const float v[3],
                                              const float C[3][3], float mult, const Sink &sink) {
    float P2[3][3];  // P2[gz][r] = C[2][r] * dist[2][gz]
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int r = 0; r < 3; ++r) P2[g][r] = C[2][r] * n.dist[2][g];
    for_rows<kRolled>(n, [&](const AxisNode &ax, const AxisNode &ay, int gx, int gy) {
        const float wxy = ax.w * ay.w;
        float PQ[3];  // C[0][r] * dist_x + C[1][r] * dist_y
#pragma unroll
        for (int r = 0; r < 3; ++r) PQ[r] = C[0][r] * ax.dist + C[1][r] * ay.dist;
#pragma unroll
        for (int gz = 0; gz < 3; ++gz) {
            const float weight = wxy * n.w[gz][2];
            float vel[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) vel[r] = weight * (v[r] + (PQ[r] + P2[gz][r]));
            sink.add4(n, ax, ay, gx, gy, gz, encode(weight, mult), encode(vel[0], mult),
                      encode(vel[1], mult), encode(vel[2], mult));
        }
    });
}

// ---- p2g_2 (mls-mpm/p2g_2.wgsl:42-92) ---------------------------------------

template <bool kRolled = false, class Sink>
__device__ __forceinline__ void p2g2_particle(const Node3 &n, const float C[3][3],
                                              const Params &prm, const Sink &sink) {
    float density = 0.0f;
    for_rows<kRolled>(n, [&](const AxisNode &ax, const AxisNode &ay, int gx, int gy) {
        const float wxy = ax.w * ay.w;
#pragma unroll
        for (int gz = 0; gz < 3; ++gz)
            density = density + sink.mass(n, ax, ay, gx, gy, gz) * (wxy * n.w[gz][2]);
    });
    const float volume = 1.0f / density;
    const float r = density / prm.rest_density;
    const float r2 = r * r;
    const float r5 = (r2 * r2) * r;  // pow(r, 5.)
    const float pw = prm.stiffness * (r5 - 1.0f);
    const float pressure = (pw > -0.0f) ? pw : -0.0f;  // max(-0.0, .)
    float M[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            float stress = (c == rr) ? -pressure : 0.0f;
            const float strain = C[c][rr] + C[rr][c];
            stress = stress + prm.viscosity * strain;
            M[c][rr] = ((-volume * 4.0f) * stress) * prm.dt;
        }
    for_rows<kRolled>(n, [&](const AxisNode &ax, const AxisNode &ay, int gx, int gy) {
        const float wxy = ax.w * ay.w;
#pragma unroll
        for (int gz = 0; gz < 3; ++gz) {
            const float weight = wxy * n.w[gz][2];
            float mom[3];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
                mom[rr] = ((M[0][rr] * weight) * ax.dist + (M[1][rr] * weight) * ay.dist) +
                          (M[2][rr] * weight) * n.dist[2][gz];
            sink.add3(n, ax, ay, gx, gy, gz, encode(mom[0], prm.mult), encode(mom[1], prm.mult),
                      encode(mom[2], prm.mult));
        }
    });
}

// ---- g2p gather (mls-mpm/g2p.wgsl:29-72) -------------------------------------

template <bool kRolled = false, class Sink>
__device__ __forceinline__ void g2p_gather(const Node3 &n, const Sink &sink, float v[3],
                                           float B[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        v[r] = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) B[c][r] = 0.0f;
    }
    for_rows<kRolled>(n, [&](const AxisNode &ax, const AxisNode &ay, int gx, int gy) {
        const float wxy = ax.w * ay.w;
#pragma unroll
        for (int gz = 0; gz < 3; ++gz) {
            const float weight = wxy * n.w[gz][2];
            float cv[3];
            sink.vel(n, ax, ay, gx, gy, gz, cv);
            const float wv[3] = {cv[0] * weight, cv[1] * weight, cv[2] * weight};
            const float d[3] = {ax.dist, ay.dist, n.dist[2][gz]};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) B[c][r] = B[c][r] + wv[r] * d[c];
#pragma unroll
            for (int r = 0; r < 3; ++r) v[r] = v[r] + wv[r];
        }
    });
}

// g2p tail: C, advection, clamp, wall penalty (mls-mpm/g2p.wgsl:72-90).
__device__ __forceinline__ void g2p_finish(const float x[3], const float rb[3], float dt,
                                           float v[3], const float B[3][3], float xo[3],
                                           float C[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = B[c][r] * 4.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xa = x[a] + v[a] * dt;
        xo[a] = fminf(fmaxf(xa, 1.0f), rb[a] - 2.0f);
    }
    const float k = 3.0f, wall_stiffness = 0.3f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xn = xo[a] + (v[a] * dt) * k;
        const float wall_max = rb[a] - 4.0f;
        float va = v[a];
        if (xn < 3.0f) va = va + wall_stiffness * (3.0f - xn);
        if (xn > wall_max) va = va + wall_stiffness * (wall_max - xn);
        v[a] = va;
    }
}

// updateGrid on one cell (mls-mpm/updateGrid.wgsl:26-44) with the cell's
// (x, y, z) already decomposed from its linear index.
__device__ __forceinline__ int4 update_cell_xyz(int4 c, int32_t x, int32_t y, int32_t z,
                                                const int32_t wall_hi[3], const Params &prm) {
    if (c.w > 0) {
        float fx = decode(c.x, prm.mult), fy = decode(c.y, prm.mult), fz = decode(c.z, prm.mult);
        const float m = decode(c.w, prm.mult);
        fx = fx / m;
        fy = fy / m;
        fz = fz / m;
        const float gdt = -0.3f * prm.dt;
        c.x = encode(fx, prm.mult);
        c.y = encode(fy + gdt, prm.mult);
        c.z = encode(fz, prm.mult);
        if (x < 2 || x > wall_hi[0]) c.x = 0;
        if (y < 2 || y > wall_hi[1]) c.y = 0;
        if (z < 2 || z > wall_hi[2]) c.z = 0;
    }
    return c;
}

// Same, decomposing the linear index like updateGrid.wgsl:37-39.
__device__ __forceinline__ int4 update_cell_id(int4 c, int32_t id, int32_t Y, int32_t Z,
                                               const int32_t wall_hi[3], const Params &prm) {
    if (c.w <= 0) return c;
    return update_cell_xyz(c, id / Z / Y, (id / Z) % Y, id % Z, wall_hi, prm);
}

}  // namespace mpm
