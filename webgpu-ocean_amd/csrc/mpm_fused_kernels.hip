// mpm_fused_kernels.hip -- the production substep for gfx950.
//
//   bin (radix/counting sort by tile-major cell key, every `sort_period`
//   substeps) -> clear -> p2g_1 -> p2g_2 -> g2p(+updateGrid +copyPosition)
//
// Particles live in SoA storage permuted into bin order; id[] maps a storage
// slot to the caller's particle index, so readout order never changes.
// Every P2G/G2P block takes a contiguous chunk of stored particles and stages
// the chunk's stencil bounding box ("window") in LDS:
//   p2g_1 / p2g_2: i32 fixed-point ds_add into the window, then one global
//                  atomic per non-zero window cell and component;
//   g2p:           window cells loaded once, updateGrid applied on load
//                  (a pure per-cell function), 27-node gathers from LDS.
// Integer adds commute, and every float expression is evaluated per particle
// exactly as in the stage kernels, so results are identical to them.
// Particles whose floor(x) is non-finite or beyond +-2^20, and whole chunks
// whose window exceeds the LDS budget, take the global-memory path of the
// stage kernels instead.
#include "mpm_fused.h"

#pragma clang fp contract(off)

namespace mpm {
namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr float kTameLimit = 1048576.0f;  // |floor(x)| <= 2^20: exact int coordinates

// ---------------------------------------------------------------------------
// small helpers

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// Inclusive wave scan of unsigned values.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan over the block (kBlock threads); returns the thread's
// exclusive prefix, *total receives the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t s_w[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) {
        uint32_t s = s_w[i];
        if (i < w) wbase += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
    uint32_t tot;
    (void)block_excl_scan(v, &tot);
    return tot;
}

__device__ __forceinline__ void load_xvC(const float *__restrict__ st, uint32_t stride,
                                         uint32_t i, float v[3], float C[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = st[(VX + a) * stride + i];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = st[(C00 + 3 * c + r) * stride + i];
}

// updateGrid on one cell (mls-mpm/updateGrid.wgsl:26-44); `id` is the cell's
// linear index.  Returns the updated (vx, vy, vz).
__device__ __forceinline__ int4 update_cell(int4 c, int32_t id, int32_t Y, int32_t Z,
                                            int32_t wx, int32_t wy, int32_t wz,
                                            const Params &prm) {
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
        int32_t x = id / Z / Y;
        int32_t y = (id / Z) % Y;
        int32_t z = id % Z;
        if (x < 2 || x > wx) c.x = 0;
        if (y < 2 || y > wy) c.y = 0;
        if (z < 2 || z > wz) c.z = 0;
    }
    return c;
}

// Linear cell index of integer cell coordinates, i32 wrap-around like WGSL
// (mls-mpm/p2g_1.wgsl:53-56).
__device__ __forceinline__ int32_t cell_linear(int32_t x, int32_t y, int32_t z, int32_t Y,
                                               int32_t Z) {
    return (int32_t)((uint32_t)x * (uint32_t)Y * (uint32_t)Z + (uint32_t)y * (uint32_t)Z +
                     (uint32_t)z);
}

// Window = bounding box of the chunk's stencils, [lo, lo + dim).
struct Window {
    int lo[3];
    int dim[3];
    int vol;   // 0: no window (global path for the whole chunk)
};

// Computes the block's window from the tame particles' integer base cells.
// Must be called by all threads of the block.
template <int PPT>
__device__ __forceinline__ Window block_window(const int ic[PPT][3], const bool tame[PPT],
                                               int wmax) {
    __shared__ int s_bb[6];
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int k = 0; k < PPT; ++k)
        if (tame[k])
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], ic[k][a]);
                hi[a] = max(hi[a], ic[k][a]);
            }
    if (threadIdx.x < 6) s_bb[threadIdx.x] = threadIdx.x < 3 ? INT_MAX : INT_MIN;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&s_bb[a], lo[a]);
            atomicMax(&s_bb[3 + a], hi[a]);
        }
    }
    __syncthreads();
    Window w;
    int64_t vol = 1;
    bool fits = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int l = s_bb[a], h = s_bb[3 + a];
        w.lo[a] = l - 1;
        // tame coordinates lie in [-2^20, 2^20], so h - l + 3 cannot overflow
        w.dim[a] = (l <= h) ? (h - l + 3) : 0;
        fits = fits && w.dim[a] > 0 && w.dim[a] <= wmax;
        vol *= fits ? w.dim[a] : 1;
    }
    w.vol = (fits && vol <= wmax) ? (int)vol : 0;
    return w;
}

__device__ __forceinline__ int win_index(const Window &w, int x, int y, int z) {
    return ((x - w.lo[0]) * w.dim[1] + (y - w.lo[1])) * w.dim[2] + (z - w.lo[2]);
}

__device__ __forceinline__ void win_coords(const Window &w, int l, int &x, int &y, int &z) {
    int lz = l % w.dim[2];
    int t = l / w.dim[2];
    int ly = t % w.dim[1];
    int lx = t / w.dim[1];
    x = w.lo[0] + lx;
    y = w.lo[1] + ly;
    z = w.lo[2] + lz;
}

// Loads the chunk's positions, integer base cells and "tame" flags.
template <int PPT>
__device__ __forceinline__ void load_chunk(const float *__restrict__ st, uint32_t stride,
                                           uint32_t n, uint32_t base, float x[PPT][3],
                                           int ic[PPT][3], bool valid[PPT], bool tame[PPT]) {
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        uint32_t i = base + threadIdx.x + k * kBlock;
        valid[k] = i < n;
        tame[k] = false;
        if (valid[k]) {
            bool t = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                x[k][a] = st[(PX + a) * stride + i];
                float ci = floorf(x[k][a]);
                t = t && (ci >= -kTameLimit) && (ci <= kTameLimit);
                ic[k][a] = t ? (int)ci : 0;
            }
            tame[k] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// binning sort

__global__ __launch_bounds__(kBlock) void k_bin_count(const float *__restrict__ st,
                                                      uint32_t stride, uint32_t n,
                                                      KeyGeom kg, uint32_t *__restrict__ key,
                                                      uint32_t *__restrict__ rank,
                                                      uint32_t *__restrict__ count) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float f = floorf(st[(PX + a) * stride + i]);
        int v = (f >= 0.0f) ? ((f < (float)kg.dim[a]) ? (int)f : kg.dim[a] - 1) : 0;
        c[a] = v;  // NaN compares false -> 0
    }
    uint32_t tile = ((uint32_t)(c[0] >> kTileBits) * kg.nt[1] + (uint32_t)(c[1] >> kTileBits)) *
                        kg.nt[2] +
                    (uint32_t)(c[2] >> kTileBits);
    constexpr int m = (1 << kTileBits) - 1;
    uint32_t local = ((uint32_t)(c[0] & m) << (2 * kTileBits)) |
                     ((uint32_t)(c[1] & m) << kTileBits) | (uint32_t)(c[2] & m);
    uint32_t k = (tile << (3 * kTileBits)) | local;
    key[i] = k;
    rank[i] = atomicAdd(&count[k], 1u);
}

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *__restrict__ count,
                                                        uint32_t nkeys,
                                                        uint32_t *__restrict__ block_sums) {
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < nkeys) s += count[base + j];
    uint32_t tot = block_sum(s);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t *__restrict__ block_sums,
                                                     uint32_t nblocks) {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kBlock) {
        uint32_t i = b0 + threadIdx.x;
        uint32_t v = i < nblocks ? block_sums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (i < nblocks) block_sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t *__restrict__ count,
                                                      uint32_t *__restrict__ offset,
                                                      uint32_t nkeys,
                                                      const uint32_t *__restrict__ block_sums) {
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < nkeys) ? count[base + j] : 0;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = block_sums[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < nkeys) {
            offset[base + j] = run;
            run += v[j];
            count[base + j] = 0;  // ready for the next binning pass
        }
}

__global__ __launch_bounds__(kBlock) void k_bin_scatter(
    const float *__restrict__ st, const uint32_t *__restrict__ id, uint32_t stride, uint32_t n,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ rank,
    const uint32_t *__restrict__ offset, float *__restrict__ st2, uint32_t *__restrict__ id2) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t d = offset[key[i]] + rank[i];
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) st2[c * stride + d] = st[c * stride + i];
    id2[d] = id[i];
}

// ---------------------------------------------------------------------------
// p2g_1: mass and APIC momentum (mls-mpm/p2g_1.wgsl:24-65)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g1_win(const float *__restrict__ st,
                                                     uint32_t stride, uint32_t n,
                                                     int32_t *__restrict__ grid, uint32_t G,
                                                     const Uniforms *__restrict__ u, Params prm,
                                                     int wmax) {
    extern __shared__ int lds[];  // [mass | vx | vy | vz] x vol
    const uint32_t base = blockIdx.x * (kBlock * PPT);
    const int32_t Y = u->Y, Z = u->Z;
    float x[PPT][3];
    int ic[PPT][3];
    bool valid[PPT], tame[PPT];
    load_chunk<PPT>(st, stride, n, base, x, ic, valid, tame);
    const Window w = block_window<PPT>(ic, tame, wmax);
    const int vol = w.vol;
    for (int l = threadIdx.x; l < 4 * vol; l += kBlock) lds[l] = 0;
    __syncthreads();

#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        const uint32_t i = base + threadIdx.x + k * kBlock;
        float v[3], C[3][3];
        load_xvC(st, stride, i, v, C);
        Stencil s;
        stencil_init(x[k], s);
        const bool in_win = vol > 0 && tame[k];
        const int l0 = in_win ? win_index(w, ic[k][0] - 1, ic[k][1] - 1, ic[k][2] - 1) : 0;
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy)
#pragma unroll
                for (int gz = 0; gz < 3; ++gz) {
                    float weight = node_weight(s, gx, gy, gz);
                    float dist[3];
                    int32_t idx = node_index(s, x[k], gx, gy, gz, Y, Z, dist);
                    float mass_contrib = weight * 1.0f;
                    float vel[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        float q = (C[0][r] * dist[0] + C[1][r] * dist[1]) + C[2][r] * dist[2];
                        vel[r] = mass_contrib * (v[r] + q);
                    }
                    const int32_t em = encode(mass_contrib, prm.mult);
                    const int32_t ex = encode(vel[0], prm.mult);
                    const int32_t ey = encode(vel[1], prm.mult);
                    const int32_t ez = encode(vel[2], prm.mult);
                    if (in_win) {
                        const int l = l0 + (gx * w.dim[1] + gy) * w.dim[2] + gz;
                        atomicAdd(&lds[l], em);
                        atomicAdd(&lds[vol + l], ex);
                        atomicAdd(&lds[2 * vol + l], ey);
                        atomicAdd(&lds[3 * vol + l], ez);
                    } else if (in_grid(idx, G)) {
                        int32_t *cell = grid + 4 * (size_t)idx;
                        atomicAdd(cell + 3, em);
                        atomicAdd(cell + 0, ex);
                        atomicAdd(cell + 1, ey);
                        atomicAdd(cell + 2, ez);
                    }
                }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < vol; l += kBlock) {
        int cx, cy, cz;
        win_coords(w, l, cx, cy, cz);
        const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
        if (!in_grid(gi, G)) continue;
        int32_t *cell = grid + 4 * (size_t)gi;
        const int32_t m = lds[l], vx = lds[vol + l], vy = lds[2 * vol + l],
                      vz = lds[3 * vol + l];
        if (m) atomicAdd(cell + 3, m);
        if (vx) atomicAdd(cell + 0, vx);
        if (vy) atomicAdd(cell + 1, vy);
        if (vz) atomicAdd(cell + 2, vz);
    }
}

// ---------------------------------------------------------------------------
// p2g_2: density, pressure, viscous stress momentum (mls-mpm/p2g_2.wgsl:30-94)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g2_win(const float *__restrict__ st,
                                                     uint32_t stride, uint32_t n,
                                                     int32_t *__restrict__ grid, uint32_t G,
                                                     const Uniforms *__restrict__ u, Params prm,
                                                     int wmax) {
    extern __shared__ int lds[];  // [decoded mass (f32) | vx | vy | vz] x vol
    float *dmass = reinterpret_cast<float *>(lds);
    const uint32_t base = blockIdx.x * (kBlock * PPT);
    const int32_t Y = u->Y, Z = u->Z;
    float x[PPT][3];
    int ic[PPT][3];
    bool valid[PPT], tame[PPT];
    load_chunk<PPT>(st, stride, n, base, x, ic, valid, tame);
    const Window w = block_window<PPT>(ic, tame, wmax);
    const int vol = w.vol;
    for (int l = threadIdx.x; l < vol; l += kBlock) {
        int cx, cy, cz;
        win_coords(w, l, cx, cy, cz);
        const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
        const int32_t m = in_grid(gi, G) ? grid[4 * (size_t)gi + 3] : 0;
        dmass[l] = decode(m, prm.mult);
        lds[vol + l] = 0;
        lds[2 * vol + l] = 0;
        lds[3 * vol + l] = 0;
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        const uint32_t i = base + threadIdx.x + k * kBlock;
        float v[3], C[3][3];
        load_xvC(st, stride, i, v, C);
        Stencil s;
        stencil_init(x[k], s);
        const bool in_win = vol > 0 && tame[k];
        const int l0 = in_win ? win_index(w, ic[k][0] - 1, ic[k][1] - 1, ic[k][2] - 1) : 0;
        float density = 0.0f;
#pragma unroll
        for (int gx = 0; gx < 3; ++gx)
#pragma unroll
            for (int gy = 0; gy < 3; ++gy)
#pragma unroll
                for (int gz = 0; gz < 3; ++gz) {
                    float weight = node_weight(s, gx, gy, gz);
                    float dm;
                    if (in_win) {
                        dm = dmass[l0 + (gx * w.dim[1] + gy) * w.dim[2] + gz];
                    } else {
                        float dist[3];
                        int32_t idx = node_index(s, x[k], gx, gy, gz, Y, Z, dist);
                        dm = decode(in_grid(idx, G) ? grid[4 * (size_t)idx + 3] : 0, prm.mult);
                    }
                    density = density + dm * weight;
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
                    int32_t idx = node_index(s, x[k], gx, gy, gz, Y, Z, dist);
                    float mom[3];
#pragma unroll
                    for (int rr = 0; rr < 3; ++rr)
                        mom[rr] = ((M[0][rr] * weight) * dist[0] + (M[1][rr] * weight) * dist[1]) +
                                  (M[2][rr] * weight) * dist[2];
                    const int32_t ex = encode(mom[0], prm.mult);
                    const int32_t ey = encode(mom[1], prm.mult);
                    const int32_t ez = encode(mom[2], prm.mult);
                    if (in_win) {
                        const int l = l0 + (gx * w.dim[1] + gy) * w.dim[2] + gz;
                        atomicAdd(&lds[vol + l], ex);
                        atomicAdd(&lds[2 * vol + l], ey);
                        atomicAdd(&lds[3 * vol + l], ez);
                    } else if (in_grid(idx, G)) {
                        int32_t *cell = grid + 4 * (size_t)idx;
                        atomicAdd(cell + 0, ex);
                        atomicAdd(cell + 1, ey);
                        atomicAdd(cell + 2, ez);
                    }
                }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < vol; l += kBlock) {
        int cx, cy, cz;
        win_coords(w, l, cx, cy, cz);
        const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
        if (!in_grid(gi, G)) continue;
        int32_t *cell = grid + 4 * (size_t)gi;
        const int32_t vx = lds[vol + l], vy = lds[2 * vol + l], vz = lds[3 * vol + l];
        if (vx) atomicAdd(cell + 0, vx);
        if (vy) atomicAdd(cell + 1, vy);
        if (vz) atomicAdd(cell + 2, vz);
    }
}

// ---------------------------------------------------------------------------
// g2p with updateGrid applied on window load and copyPosition fused
// (mls-mpm/updateGrid.wgsl:23-46, g2p.wgsl:26-92, copyPosition.wgsl:15-21)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_g2p_win(float *__restrict__ st,
                                                    const uint32_t *__restrict__ id,
                                                    uint32_t stride, uint32_t n,
                                                    const int4 *__restrict__ grid, uint32_t G,
                                                    float4 *__restrict__ posvel,
                                                    const Uniforms *__restrict__ u, Params prm,
                                                    int wmax) {
    extern __shared__ float4 lvel[];  // decoded updated velocity x vol
    const uint32_t base = blockIdx.x * (kBlock * PPT);
    const int32_t Y = u->Y, Z = u->Z;
    const int32_t wx = u->wall_hi[0], wy = u->wall_hi[1], wz = u->wall_hi[2];
    float x[PPT][3];
    int ic[PPT][3];
    bool valid[PPT], tame[PPT];
    load_chunk<PPT>(st, stride, n, base, x, ic, valid, tame);
    const Window w = block_window<PPT>(ic, tame, wmax);
    const int vol = w.vol;
    for (int l = threadIdx.x; l < vol; l += kBlock) {
        int cx, cy, cz;
        win_coords(w, l, cx, cy, cz);
        const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in_grid(gi, G)) {
            int4 c = update_cell(grid[gi], gi, Y, Z, wx, wy, wz, prm);
            o = make_float4(decode(c.x, prm.mult), decode(c.y, prm.mult), decode(c.z, prm.mult),
                            0.0f);
        }
        lvel[l] = o;
    }
    __syncthreads();

    const float dt = prm.dt;
    float rb[3] = {u->real_box[0], u->real_box[1], u->real_box[2]};
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        const uint32_t i = base + threadIdx.x + k * kBlock;
        Stencil s;
        stencil_init(x[k], s);
        const bool in_win = vol > 0 && tame[k];
        const int l0 = in_win ? win_index(w, ic[k][0] - 1, ic[k][1] - 1, ic[k][2] - 1) : 0;
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
                    int32_t idx = node_index(s, x[k], gx, gy, gz, Y, Z, dist);
                    float cv[3];
                    if (in_win) {
                        float4 o = lvel[l0 + (gx * w.dim[1] + gy) * w.dim[2] + gz];
                        cv[0] = o.x;
                        cv[1] = o.y;
                        cv[2] = o.z;
                    } else {
                        int4 c = make_int4(0, 0, 0, 0);
                        if (in_grid(idx, G)) c = update_cell(grid[idx], idx, Y, Z, wx, wy, wz, prm);
                        cv[0] = decode(c.x, prm.mult);
                        cv[1] = decode(c.y, prm.mult);
                        cv[2] = decode(c.z, prm.mult);
                    }
                    float wv[3] = {cv[0] * weight, cv[1] * weight, cv[2] * weight};
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
            for (int r = 0; r < 3; ++r) st[(C00 + 3 * c + r) * stride + i] = B[c][r] * 4.0f;
        float xo[3], vo[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float xa = x[k][a] + v[a] * dt;
            xo[a] = fminf(fmaxf(xa, 1.0f), rb[a] - 2.0f);
        }
        const float kk = 3.0f, wall_stiffness = 0.3f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float xn = xo[a] + (v[a] * dt) * kk;
            float wall_max = rb[a] - 4.0f;
            float va = v[a];
            if (xn < 3.0f) va = va + wall_stiffness * (3.0f - xn);
            if (xn > wall_max) va = va + wall_stiffness * (wall_max - xn);
            vo[a] = va;
            st[(PX + a) * stride + i] = xo[a];
            st[(VX + a) * stride + i] = va;
        }
        const uint32_t o = id[i];
        posvel[2 * (size_t)o] = make_float4(xo[0], xo[1], xo[2], 0.0f);
        posvel[2 * (size_t)o + 1] = make_float4(vo[0], vo[1], vo[2], 0.0f);
    }
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t *__restrict__ id, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) id[i] = i;
}

inline unsigned blocks_for(uint64_t count, uint64_t per) {
    return (unsigned)((count + per - 1) / per);
}

template <int PPT>
hipError_t launch_win_kernels(int which, const FusedArgs &a, hipStream_t st) {
    const unsigned nb = blocks_for(a.n, (uint64_t)kBlock * PPT);
    const size_t lds = 16 * (size_t)a.wmax;
    if (!a.n) return hipSuccess;
    switch (which) {
    case 1:
        k_p2g1_win<PPT><<<nb, kBlock, lds, st>>>(a.state, a.stride, a.n, (int32_t *)a.grid, a.G,
                                                  a.uni, a.prm, a.wmax);
        break;
    case 2:
        k_p2g2_win<PPT><<<nb, kBlock, lds, st>>>(a.state, a.stride, a.n, (int32_t *)a.grid, a.G,
                                                  a.uni, a.prm, a.wmax);
        break;
    case 3:
        k_g2p_win<PPT><<<nb, kBlock, lds, st>>>(a.state, a.id, a.stride, a.n, a.grid, a.G,
                                                 a.posvel, a.uni, a.prm, a.wmax);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

KeyGeom make_key_geom(const float init_box[3]) {
    KeyGeom kg{};
    uint64_t nkeys = 1;
    for (int a = 0; a < 3; ++a) {
        kg.dim[a] = (int32_t)std::ceil(init_box[a]);
        if (kg.dim[a] < 1) kg.dim[a] = 1;
        kg.nt[a] = (uint32_t)((kg.dim[a] + (1 << kTileBits) - 1) >> kTileBits);
        nkeys *= kg.nt[a];
    }
    kg.nkeys = nkeys << (3 * kTileBits);
    return kg;
}

hipError_t launch_bin(const FusedArgs &a, const BinArgs &b, hipStream_t st) {
    if (!a.n) return hipSuccess;
    k_bin_count<<<blocks_for(a.n, kBlock), kBlock, 0, st>>>(a.state, a.stride, a.n, b.kg, b.key,
                                                            b.rank, b.count);
    const unsigned nsb = blocks_for(b.kg.nkeys, kScanTile);
    k_scan_reduce<<<nsb, kBlock, 0, st>>>(b.count, (uint32_t)b.kg.nkeys, b.block_sums);
    k_scan_top<<<1, kBlock, 0, st>>>(b.block_sums, nsb);
    k_scan_down<<<nsb, kBlock, 0, st>>>(b.count, b.offset, (uint32_t)b.kg.nkeys, b.block_sums);
    k_bin_scatter<<<blocks_for(a.n, kBlock), kBlock, 0, st>>>(
        a.state, a.id, a.stride, a.n, b.key, b.rank, b.offset, b.state_out, b.id_out);
    return hipGetLastError();
}

hipError_t launch_p2g1(const FusedArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_win_kernels<1>(1, a, st);
    case 2: return launch_win_kernels<2>(1, a, st);
    case 4: return launch_win_kernels<4>(1, a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_p2g2(const FusedArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_win_kernels<1>(2, a, st);
    case 2: return launch_win_kernels<2>(2, a, st);
    case 4: return launch_win_kernels<4>(2, a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_g2p(const FusedArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_win_kernels<1>(3, a, st);
    case 2: return launch_win_kernels<2>(3, a, st);
    case 4: return launch_win_kernels<4>(3, a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_iota(uint32_t *id, uint32_t n, hipStream_t st) {
    if (n) k_iota<<<blocks_for(n, kBlock), kBlock, 0, st>>>(id, n);
    return hipGetLastError();
}

}  // namespace mpm
