// mpm_stage_kernels.hip -- the six reference kernels, one gfx950 launch each.
//
// Same decomposition as the reference (one thread per particle or cell,
// 27-node stencils, i32 fixed-point global atomics) so that every stage can
// be run and checked on its own (mpm_run_stage).  The production substep in
// mpm_fused_kernels.hip is checked against these and against the CPU oracle.
#include "mpm_internal.h"

#pragma clang fp contract(off)

namespace mpm {
namespace {

constexpr int kBlock = 256;

// Where component k (PX..C22) of particle i lives.
struct SoALayout {  // the library's own state: state[k * stride + i]
    float *st;
    uint32_t stride;
    __device__ __forceinline__ float &at(int k, uint32_t i) const { return st[k * stride + i]; }
};

struct RecordLayout {  // the reference's 80-B records (mls-mpm/p2g_1.wgsl:1-5): position @0,
                       // v @16, C column c @32 + 16c, each vec3 followed by 4 B of pad
    float *rec;
    __device__ __forceinline__ float &at(int k, uint32_t i) const {
        const int off = k < VX ? k : (k < C00 ? k + 1 : 8 + 4 * ((k - C00) / 3) + (k - C00) % 3);
        return rec[20 * (size_t)i + off];
    }
};

template <class L>
__device__ __forceinline__ void load_particle(const L &p, uint32_t i, float x[3], float v[3],
                                              float C[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x[a] = p.at(PX + a, i);
        v[a] = p.at(VX + a, i);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = p.at(C00 + 3 * c + r, i);
}

// clearGrid -- mls-mpm/clearGrid.wgsl:10-18
__global__ __launch_bounds__(kBlock) void k_clear_grid(int4 *__restrict__ grid, uint32_t G) {
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id < G) grid[id] = make_int4(0, 0, 0, 0);
}

// p2g_1 -- mls-mpm/p2g_1.wgsl:24-65
template <class L>
__global__ __launch_bounds__(kBlock) void k_p2g_1(L p, uint32_t n, int32_t *__restrict__ grid,
                                                  uint32_t G, const Uniforms *__restrict__ u,
                                                  Params prm) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[3], v[3], C[3][3];
    load_particle(p, i, x, v, C);
    Stencil s;
    stencil_init(x, s);
    const int32_t Y = u->Y, Z = u->Z;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(s, gx, gy, gz);
                float dist[3];
                int32_t idx = node_index(s, x, gx, gy, gz, Y, Z, dist);
                float mass_contrib = weight * 1.0f;
                float vel[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    float q = (C[0][r] * dist[0] + C[1][r] * dist[1]) + C[2][r] * dist[2];
                    vel[r] = mass_contrib * (v[r] + q);
                }
                if (!in_grid(idx, G)) continue;
                int32_t *cell = grid + 4 * (size_t)idx;
                atomicAdd(cell + 3, encode(mass_contrib, prm.mult));
                atomicAdd(cell + 0, encode(vel[0], prm.mult));
                atomicAdd(cell + 1, encode(vel[1], prm.mult));
                atomicAdd(cell + 2, encode(vel[2], prm.mult));
            }
}

// p2g_2 -- mls-mpm/p2g_2.wgsl:30-94
template <class L>
__global__ __launch_bounds__(kBlock) void k_p2g_2(L p, uint32_t n, int32_t *__restrict__ grid,
                                                  uint32_t G, const Uniforms *__restrict__ u,
                                                  Params prm) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[3], v[3], C[3][3];
    load_particle(p, i, x, v, C);
    Stencil s;
    stencil_init(x, s);
    const int32_t Y = u->Y, Z = u->Z;
    float density = 0.0f;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(s, gx, gy, gz);
                float dist[3];
                int32_t idx = node_index(s, x, gx, gy, gz, Y, Z, dist);
                int32_t m = in_grid(idx, G) ? grid[4 * (size_t)idx + 3] : 0;
                density = density + decode(m, prm.mult) * weight;
            }
    float volume = 1.0f / density;
    float r = density / prm.rest_density;
    float r2 = r * r;
    float r5 = (r2 * r2) * r;
    float pw = prm.stiffness * (r5 - 1.0f);
    float pressure = (pw > -0.0f) ? pw : -0.0f;
    float M[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            float stress = (c == rr) ? -pressure : 0.0f;
            float strain = C[c][rr] + C[rr][c];
            stress = stress + prm.viscosity * strain;
            M[c][rr] = ((-volume * 4.0f) * stress) * prm.dt;
        }
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(s, gx, gy, gz);
                float dist[3];
                int32_t idx = node_index(s, x, gx, gy, gz, Y, Z, dist);
                float mom[3];
#pragma unroll
                for (int rr = 0; rr < 3; ++rr)
                    mom[rr] = ((M[0][rr] * weight) * dist[0] + (M[1][rr] * weight) * dist[1]) +
                              (M[2][rr] * weight) * dist[2];
                if (!in_grid(idx, G)) continue;
                int32_t *cell = grid + 4 * (size_t)idx;
                atomicAdd(cell + 0, encode(mom[0], prm.mult));
                atomicAdd(cell + 1, encode(mom[1], prm.mult));
                atomicAdd(cell + 2, encode(mom[2], prm.mult));
            }
}

// updateGrid -- mls-mpm/updateGrid.wgsl:23-46
__global__ __launch_bounds__(kBlock) void k_update_grid(int4 *__restrict__ grid, uint32_t G,
                                                        const Uniforms *__restrict__ u,
                                                        Params prm) {
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= G) return;
    int4 c = grid[id];
    if (c.w > 0) {
        float fx = decode(c.x, prm.mult), fy = decode(c.y, prm.mult), fz = decode(c.z, prm.mult);
        float m = decode(c.w, prm.mult);
        fx = fx / m;
        fy = fy / m;
        fz = fz / m;
        float gdt = -0.3f * prm.dt;
        c.x = encode(fx, prm.mult);
        c.y = encode(fy + gdt, prm.mult);
        c.z = encode(fz, prm.mult);
        const int32_t Y = u->Y, Z = u->Z, iid = (int32_t)id;
        int32_t x = iid / Z / Y;
        int32_t y = (iid / Z) % Y;
        int32_t z = iid % Z;
        if (x < 2 || x > u->wall_hi[0]) c.x = 0;
        if (y < 2 || y > u->wall_hi[1]) c.y = 0;
        if (z < 2 || z > u->wall_hi[2]) c.z = 0;
        grid[id] = c;
    }
}

// g2p -- mls-mpm/g2p.wgsl:26-92
template <class L>
__global__ __launch_bounds__(kBlock) void k_g2p(L p, uint32_t n, const int4 *__restrict__ grid,
                                                uint32_t G, const Uniforms *__restrict__ u,
                                                Params prm) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = p.at(PX + a, i);
    Stencil s;
    stencil_init(x, s);
    const int32_t Y = u->Y, Z = u->Z;
    float v[3] = {0.0f, 0.0f, 0.0f};
    float B[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy)
#pragma unroll
            for (int gz = 0; gz < 3; ++gz) {
                float weight = node_weight(s, gx, gy, gz);
                float dist[3];
                int32_t idx = node_index(s, x, gx, gy, gz, Y, Z, dist);
                int4 c = in_grid(idx, G) ? grid[idx] : make_int4(0, 0, 0, 0);
                float wv[3] = {decode(c.x, prm.mult) * weight, decode(c.y, prm.mult) * weight,
                               decode(c.z, prm.mult) * weight};
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                    for (int r = 0; r < 3; ++r) B[cc][r] = B[cc][r] + wv[r] * dist[cc];
#pragma unroll
                for (int r = 0; r < 3; ++r) v[r] = v[r] + wv[r];
            }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) p.at(C00 + 3 * c + r, i) = B[c][r] * 4.0f;
    const float dt = prm.dt;
    float xn_[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xa = x[a] + v[a] * dt;
        xa = fminf(fmaxf(xa, 1.0f), u->real_box[a] - 2.0f);
        xn_[a] = xa;
    }
    const float k = 3.0f, wall_stiffness = 0.3f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xn = xn_[a] + (v[a] * dt) * k;
        float wall_max = u->real_box[a] - 4.0f;
        float va = v[a];
        if (xn < 3.0f) va = va + wall_stiffness * (3.0f - xn);
        if (xn > wall_max) va = va + wall_stiffness * (wall_max - xn);
        p.at(PX + a, i) = xn_[a];
        p.at(VX + a, i) = va;
    }
}

// copyPosition -- mls-mpm/copyPosition.wgsl:15-21 (PosVel in caller order: particle i of
// the storage is caller particle id[i], or i itself without an id map)
template <class L>
__global__ __launch_bounds__(kBlock) void k_copy_position(L p, const uint32_t *__restrict__ id,
                                                          uint32_t n,
                                                          float4 *__restrict__ posvel) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t o = id ? id[i] : i;
    posvel[2 * o] = make_float4(p.at(PX, i), p.at(PY, i), p.at(PZ, i), 0.0f);
    posvel[2 * o + 1] = make_float4(p.at(VX, i), p.at(VY, i), p.at(VZ, i), 0.0f);
}

__global__ void k_set_uniforms(Uniforms *dst, Uniforms src) { *dst = src; }

// 80-B reference record (20 floats) <-> SoA state.
__global__ __launch_bounds__(kBlock) void k_unpack(const float *__restrict__ packed,
                                                   float *__restrict__ st,
                                                   uint32_t *__restrict__ id, uint32_t stride,
                                                   uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float *p = packed + 20 * (size_t)i;
    id[i] = i;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st[(PX + a) * stride + i] = p[a];
        st[(VX + a) * stride + i] = p[4 + a];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) st[(C00 + 3 * c + r) * stride + i] = p[8 + 4 * c + r];
}

__global__ __launch_bounds__(kBlock) void k_pack(float *__restrict__ packed,
                                                 const float *__restrict__ st,
                                                 const uint32_t *__restrict__ id,
                                                 uint32_t stride, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float *p = packed + 20 * (size_t)id[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        p[a] = st[(PX + a) * stride + i];
        p[4 + a] = st[(VX + a) * stride + i];
    }
    p[3] = 0.0f;
    p[7] = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) p[8 + 4 * c + r] = st[(C00 + 3 * c + r) * stride + i];
        p[8 + 4 * c + 3] = 0.0f;
    }
}

inline unsigned blocks_for(uint64_t count) { return (unsigned)((count + kBlock - 1) / kBlock); }

}  // namespace

namespace {

// One reference kernel over particles stored as L (mls-mpm/mls-mpm.ts:248-265).
template <class L>
hipError_t stage_on(int stage, L p, const uint32_t *id, uint32_t n, int4 *grid, uint32_t G,
                    const Uniforms *uni, const Params &prm, float4 *posvel, hipStream_t st) {
    switch (stage) {
    case 0:
        if (G) k_clear_grid<<<blocks_for(G), kBlock, 0, st>>>(grid, G);
        break;
    case 1:
        if (n) k_p2g_1<<<blocks_for(n), kBlock, 0, st>>>(p, n, (int32_t *)grid, G, uni, prm);
        break;
    case 2:
        if (n) k_p2g_2<<<blocks_for(n), kBlock, 0, st>>>(p, n, (int32_t *)grid, G, uni, prm);
        break;
    case 3:
        if (G) k_update_grid<<<blocks_for(G), kBlock, 0, st>>>(grid, G, uni, prm);
        break;
    case 4:
        if (n) k_g2p<<<blocks_for(n), kBlock, 0, st>>>(p, n, grid, G, uni, prm);
        break;
    case 5:
        if (n) k_copy_position<<<blocks_for(n), kBlock, 0, st>>>(p, id, n, posvel);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_stage(int stage, const DeviceScene &s, hipStream_t st) {
    return stage_on(stage, SoALayout{s.state, s.stride}, s.id, s.n, s.grid, s.G, s.uni, s.prm,
                    s.posvel, st);
}

hipError_t launch_stage_records(int stage, float *records, uint32_t n, int4 *grid, uint32_t G,
                                const Uniforms *uni, const Params &prm, float4 *posvel,
                                hipStream_t st) {
    return stage_on(stage, RecordLayout{records}, nullptr, n, grid, G, uni, prm, posvel, st);
}

hipError_t launch_set_uniforms(Uniforms *dst, const Uniforms &src, hipStream_t st) {
    k_set_uniforms<<<1, 1, 0, st>>>(dst, src);
    return hipGetLastError();
}

hipError_t launch_unpack_particles(const float *packed, const DeviceScene &s, hipStream_t st) {
    if (s.n) k_unpack<<<blocks_for(s.n), kBlock, 0, st>>>(packed, s.state, s.id, s.stride, s.n);
    return hipGetLastError();
}

hipError_t launch_pack_particles(float *packed, const DeviceScene &s, hipStream_t st) {
    if (s.n) k_pack<<<blocks_for(s.n), kBlock, 0, st>>>(packed, s.state, s.id, s.stride, s.n);
    return hipGetLastError();
}

}  // namespace mpm
