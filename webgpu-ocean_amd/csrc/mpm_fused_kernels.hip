// mpm_fused_kernels.hip -- the production substep for gfx950.
//
//   bin (counting sort by tile-major cell key, every `sort_period` substeps)
//   -> clear -> p2g_1 -> p2g_2 -> g2p (+updateGrid +copyPosition)
//
// Each stage is one launch over a table of particle chunks covering every
// scene of the batch; blocks walk the table persistently.  Particles live in
// SoA storage permuted into bin order; id[] maps a storage slot to the
// caller's particle index, so readout order never changes.  A block stages
// the 6^3-cell windows (4^3 tile + halo) of the tiles its chunk occupies in
// LDS:
//   p2g_1 / p2g_2: i32 fixed-point ds_add into the windows, then one global
//                  atomic per non-zero window cell and component;
//   g2p:           window cells loaded once with updateGrid applied on load
//                  (a pure per-cell function), 27-node gathers from LDS.
// Integer adds commute, and every float expression is evaluated per particle
// exactly as in the reference-shaped stage kernels, so results are identical
// to them.  Particles whose floor(x) is non-finite or far outside any grid,
// and particles whose tile finds no free window slot, take the global-memory
// path of the stage kernels instead.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "mpm_fused.h"
#include "mpm_particle.h"

#pragma clang fp contract(off)

namespace mpm {
namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;
// Tame particles have |floor(x)| < 2040 cells, so their tiles pack into
// tile_key and their cell coordinates are exact integers.
constexpr float kTameLimit = 2040.0f;

// ---------------------------------------------------------------------------
// chunk walk: blocks stride over the chunk table; with xcd_remap, the chunks
// are dealt so that the blocks of one XCD (block id mod 8) take a contiguous
// range, keeping neighbouring windows in the same L2.

struct ChunkWalk {
    uint32_t t, end, step, per;
    bool remap;
};

__device__ __forceinline__ ChunkWalk walk_begin(const BatchArgs &a) {
    ChunkWalk w;
    w.remap = a.xcd_remap != 0;
    w.per = (a.nchunks + kXcds - 1) / kXcds;
    w.end = w.remap ? w.per * kXcds : a.nchunks;
    w.t = blockIdx.x;
    w.step = gridDim.x;
    return w;
}

__device__ __forceinline__ uint32_t walk_chunk(const ChunkWalk &w, uint32_t nchunks) {
    uint32_t c = w.remap ? (w.t % kXcds) * w.per + w.t / kXcds : w.t;
    return c < nchunks ? c : 0xffffffffu;
}

// ---------------------------------------------------------------------------
// small helpers

// Slot (within its chunk) of a thread's k-th particle.  Lanes of a wave are
// 4*PPT slots apart, so in cell-sorted storage they mostly touch different
// cells and their LDS adds rarely hit the same address in one instruction.
template <int PPT>
__device__ __forceinline__ uint32_t slot_of(int k) {
    return (threadIdx.x & 63) * (4 * PPT) + (threadIdx.x >> 6) * PPT + k;
}

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

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan over the block; *total receives the block sum.
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

__device__ __forceinline__ void load_vC(const float *__restrict__ st, uint32_t stride,
                                        uint32_t i, float v[3], float C[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = st[(VX + a) * stride + i];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = st[(C00 + 3 * c + r) * stride + i];
}

// Linear cell index of integer cell coordinates, i32 wrap-around like WGSL
// (mls-mpm/p2g_1.wgsl:53-56).
__device__ __forceinline__ int32_t cell_linear(int32_t x, int32_t y, int32_t z, int32_t Y,
                                               int32_t Z) {
    return (int32_t)((uint32_t)x * (uint32_t)Y * (uint32_t)Z + (uint32_t)y * (uint32_t)Z +
                     (uint32_t)z);
}

// ---------------------------------------------------------------------------
// tile windows: each block registers the 4^3-cell tiles its particles occupy
// in a small LDS table; tile slot s owns a fixed 6^3-cell window (the tile
// plus a one-cell halo) at LDS cells [s*216, (s+1)*216).  A particle's
// 27-node stencil always lies inside its own tile's window.

constexpr int kTileEdge = 1 << kTileBits;            // 4
constexpr int kWinEdge = kTileEdge + 2;              // 6
constexpr int kWinCells = kWinEdge * kWinEdge * kWinEdge;  // 216
constexpr int kMaxSlots = 16;
constexpr uint32_t kNoTile = 0xffffffffu;
constexpr int kTileCoordLimit = 512;                 // packed tile coords: 10 bits each

__device__ __forceinline__ uint32_t tile_key(int tx, int ty, int tz) {
    return ((uint32_t)(tx + kTileCoordLimit) << 20) | ((uint32_t)(ty + kTileCoordLimit) << 10) |
           (uint32_t)(tz + kTileCoordLimit);
}

__device__ __forceinline__ void tile_origin(uint32_t key, int o[3]) {
    // first cell of the slot's window: tile origin minus the halo
    o[0] = ((int)(key >> 20) - kTileCoordLimit) * kTileEdge - 1;
    o[1] = ((int)((key >> 10) & 1023u) - kTileCoordLimit) * kTileEdge - 1;
    o[2] = ((int)(key & 1023u) - kTileCoordLimit) * kTileEdge - 1;
}

// Registers the tiles of the particles still to do (bit k of `todo`) in the
// block's slot table; slot[k] = the tile's slot, or -1 when the table is full
// (the particle waits for the next pass).  Returns the number of slots in use
// (>= 1 whenever any thread has work).  Called by every thread of the block.
template <int PPT>
__device__ __forceinline__ int block_tiles(const int ic[PPT][3], uint32_t todo, int max_slots,
                                           int slot[PPT], uint32_t *s_tiles) {
    if (threadIdx.x < kMaxSlots) s_tiles[threadIdx.x] = kNoTile;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        slot[k] = -1;
        if (!(todo & (1u << k))) continue;
        const uint32_t key = tile_key(ic[k][0] >> kTileBits, ic[k][1] >> kTileBits,
                                      ic[k][2] >> kTileBits);
        // Slots fill as a prefix: a thread tries slot s only after seeing
        // slots 0..s-1 taken, and a taken slot is never released.
        for (int sI = 0; sI < max_slots; ++sI) {
            uint32_t cur = s_tiles[sI];
            if (cur == kNoTile) cur = atomicCAS(&s_tiles[sI], kNoTile, key);
            if (cur == kNoTile || cur == key) {
                slot[k] = sI;
                break;
            }
        }
    }
    __syncthreads();
    int used = 0;
    for (int sI = 0; sI < max_slots; ++sI) used += s_tiles[sI] != kNoTile;
    return used;
}

// LDS cell of node (gx,gy,gz) = l0 + (gx*6 + gy)*6 + gz for a particle whose
// base cell is `ic` in tile slot `sl`.
__device__ __forceinline__ int tile_l0(int sl, const int ic[3]) {
    const int m = kTileEdge - 1;
    return sl * kWinCells + ((ic[0] & m) * kWinEdge + (ic[1] & m)) * kWinEdge + (ic[2] & m);
}

// Cell coordinates of LDS cell l of the slot table.
__device__ __forceinline__ void slot_cell(const uint32_t *s_tiles, int l, int &x, int &y,
                                          int &z) {
    const int sl = l / kWinCells, r = l - sl * kWinCells;
    int o[3];
    tile_origin(s_tiles[sl], o);
    x = o[0] + r / (kWinEdge * kWinEdge);
    y = o[1] + (r / kWinEdge) % kWinEdge;
    z = o[2] + r % kWinEdge;
}

// Loads a chunk's positions, integer base cells and "tame" flags.
template <int PPT>
__device__ __forceinline__ void load_chunk(const float *__restrict__ st, uint32_t stride,
                                           const Chunk &ch, float x[PPT][3], int ic[PPT][3],
                                           bool valid[PPT], bool tame[PPT]) {
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const uint32_t j = slot_of<PPT>(k);
        const uint32_t i = ch.begin + j;
        valid[k] = j < ch.count;
        tame[k] = false;
        if (valid[k]) {
            bool t = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                x[k][a] = st[(PX + a) * stride + i];
                float ci = floorf(x[k][a]);
                t = t && (ci >= -kTameLimit) && (ci < kTameLimit);
                ic[k][a] = t ? (int)ci : 0;
            }
            tame[k] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// binning sort

__global__ __launch_bounds__(kBlock) void k_bin_count(BatchArgs a, BinBuffers b) {
    for (ChunkWalk cw = walk_begin(a); cw.t < cw.end; cw.t += cw.step) {
        const uint32_t c = walk_chunk(cw, a.nchunks);
        if (c == 0xffffffffu) continue;
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        for (uint32_t j = threadIdx.x; j < ch.count; j += kBlock) {
            const uint32_t i = ch.begin + j;
            int cc[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const float f = floorf(a.state[(PX + ax) * a.stride + i]);
                const int dim = sd->kdim[ax];
                cc[ax] = (f >= 0.0f) ? ((f < (float)dim) ? (int)f : dim - 1) : 0;  // NaN -> 0
            }
            const uint32_t tile = a.tile_rank[sd->rank_off +
                                              ((uint32_t)(cc[0] >> kTileBits) * sd->knt[1] +
                                               (uint32_t)(cc[1] >> kTileBits)) *
                                                  sd->knt[2] +
                                              (uint32_t)(cc[2] >> kTileBits)];
            constexpr int m = (1 << kTileBits) - 1;
            const uint32_t local = ((uint32_t)(cc[0] & m) << (2 * kTileBits)) |
                                   ((uint32_t)(cc[1] & m) << kTileBits) | (uint32_t)(cc[2] & m);
            const uint32_t k = sd->koff + ((tile << (3 * kTileBits)) | local);
            b.key[i] = k;
            b.rank[i] = atomicAdd(&b.count[k], 1u);
        }
    }
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

// Keys of scene s lie in [koff_s, koff_s + nkeys_s) and scenes are packed in
// slot order, so the scattered slot stays inside the scene's slot range.
__global__ __launch_bounds__(kBlock) void k_bin_scatter(BatchArgs a, BinBuffers b) {
    for (ChunkWalk cw = walk_begin(a); cw.t < cw.end; cw.t += cw.step) {
        const uint32_t c = walk_chunk(cw, a.nchunks);
        if (c == 0xffffffffu) continue;
        const Chunk ch = a.chunks[c];
        for (uint32_t j = threadIdx.x; j < ch.count; j += kBlock) {
            const uint32_t i = ch.begin + j;
            const uint32_t d = b.offset[b.key[i]] + b.rank[i];
#pragma unroll
            for (int k = 0; k < NCOMP; ++k) b.state_out[k * a.stride + d] = a.state[k * a.stride + i];
            b.id_out[d] = a.id[i];
        }
    }
}

// ---------------------------------------------------------------------------
// sinks: where a particle's 27 node contributions go / come from

struct WinSink1 {  // p2g_1 into the LDS windows: [mass | vx | vy | vz] x vol
    int *lds;
    int vol, l0;
    __device__ __forceinline__ void add4(const Node3 &, int gx, int gy, int gz, int32_t m,
                                         int32_t x, int32_t y, int32_t z) const {
        const int l = l0 + (gx * kWinEdge + gy) * kWinEdge + gz;
        atomicAdd(&lds[l], m);
        atomicAdd(&lds[vol + l], x);
        atomicAdd(&lds[2 * vol + l], y);
        atomicAdd(&lds[3 * vol + l], z);
    }
};

struct GridSink1 {  // p2g_1 straight to the grid (reference-shaped atomics)
    int32_t *grid;
    uint32_t G;
    int32_t Y, Z;
    __device__ __forceinline__ void add4(const Node3 &n, int gx, int gy, int gz, int32_t m,
                                         int32_t x, int32_t y, int32_t z) const {
        const int32_t idx = node3_index(n, gx, gy, gz, Y, Z);
        if (!in_grid(idx, G)) return;
        int32_t *cell = grid + 4 * (size_t)idx;
        atomicAdd(cell + 3, m);
        atomicAdd(cell + 0, x);
        atomicAdd(cell + 1, y);
        atomicAdd(cell + 2, z);
    }
};

struct WinSink2 {  // p2g_2: decoded masses [0, vol), momentum ints [vol, 4 vol)
    int *lds;
    int vol, l0;
    __device__ __forceinline__ float mass(const Node3 &, int gx, int gy, int gz) const {
        return reinterpret_cast<const float *>(lds)[l0 + (gx * kWinEdge + gy) * kWinEdge + gz];
    }
    __device__ __forceinline__ void add3(const Node3 &, int gx, int gy, int gz, int32_t x,
                                         int32_t y, int32_t z) const {
        const int l = l0 + (gx * kWinEdge + gy) * kWinEdge + gz;
        atomicAdd(&lds[vol + l], x);
        atomicAdd(&lds[2 * vol + l], y);
        atomicAdd(&lds[3 * vol + l], z);
    }
};

struct GridSink2 {
    int32_t *grid;
    uint32_t G;
    int32_t Y, Z;
    float mult;
    __device__ __forceinline__ float mass(const Node3 &n, int gx, int gy, int gz) const {
        const int32_t idx = node3_index(n, gx, gy, gz, Y, Z);
        return decode(in_grid(idx, G) ? grid[4 * (size_t)idx + 3] : 0, mult);
    }
    __device__ __forceinline__ void add3(const Node3 &n, int gx, int gy, int gz, int32_t x,
                                         int32_t y, int32_t z) const {
        const int32_t idx = node3_index(n, gx, gy, gz, Y, Z);
        if (!in_grid(idx, G)) return;
        int32_t *cell = grid + 4 * (size_t)idx;
        atomicAdd(cell + 0, x);
        atomicAdd(cell + 1, y);
        atomicAdd(cell + 2, z);
    }
};

struct WinSinkG {  // g2p: updated, decoded velocities staged in LDS
    const float4 *lvel;
    int l0;
    __device__ __forceinline__ void vel(const Node3 &, int gx, int gy, int gz, float cv[3]) const {
        const float4 o = lvel[l0 + (gx * kWinEdge + gy) * kWinEdge + gz];
        cv[0] = o.x;
        cv[1] = o.y;
        cv[2] = o.z;
    }
};

struct GridSinkG {  // g2p from the raw accumulators, updateGrid applied per node
    const int4 *grid;
    uint32_t G;
    int32_t Y, Z;
    int32_t wall_hi[3];
    Params prm;
    __device__ __forceinline__ void vel(const Node3 &n, int gx, int gy, int gz, float cv[3]) const {
        const int32_t idx = node3_index(n, gx, gy, gz, Y, Z);
        int4 c = make_int4(0, 0, 0, 0);
        if (in_grid(idx, G)) c = update_cell_id(grid[idx], idx, Y, Z, wall_hi, prm);
        cv[0] = decode(c.x, prm.mult);
        cv[1] = decode(c.y, prm.mult);
        cv[2] = decode(c.z, prm.mult);
    }
};

// ---------------------------------------------------------------------------
// p2g_1: mass and APIC momentum (mls-mpm/p2g_1.wgsl:24-65)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g1_win(BatchArgs a) {
    extern __shared__ int lds[];  // [mass | vx | vy | vz] x vol
    __shared__ uint32_t s_tiles[kMaxSlots];
    for (ChunkWalk cw = walk_begin(a); cw.t < cw.end; cw.t += cw.step) {
        const uint32_t c = walk_chunk(cw, a.nchunks);
        if (c == 0xffffffffu) continue;
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        const int32_t Y = sd->uni.Y, Z = sd->uni.Z;
        const uint32_t G = sd->G;
        const float mult = sd->prm.mult;
        int32_t *grid = (int32_t *)(a.grid + sd->goff);
        float x[PPT][3], v[PPT][3], C[PPT][3][3];
        int ic[PPT][3];
        bool valid[PPT], tame[PPT];
        load_chunk<PPT>(a.state, a.stride, ch, x, ic, valid, tame);
        uint32_t todo = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            load_vC(a.state, a.stride, ch.begin + slot_of<PPT>(k), v[k], C[k]);
            if (tame[k]) todo |= 1u << k;
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            if (valid[k] && !tame[k]) {
                Node3 n;
                node3_init(x[k], n);
                p2g1_particle(n, v[k], C[k], mult, GridSink1{grid, G, Y, Z});
            }
        // one pass per group of at most max_slots tiles
        while (__syncthreads_or(todo != 0)) {
            int slot[PPT];
            const int vol = kWinCells * block_tiles<PPT>(ic, todo, a.max_slots, slot, s_tiles);
            for (int l = threadIdx.x; l < 4 * vol; l += kBlock) lds[l] = 0;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (slot[k] < 0) continue;
                Node3 n;
                node3_init(x[k], n);
                p2g1_particle(n, v[k], C[k], mult, WinSink1{lds, vol, tile_l0(slot[k], ic[k])});
                todo &= ~(1u << k);
            }
            __syncthreads();
            for (int l = threadIdx.x; l < vol; l += kBlock) {
                int cx, cy, cz;
                slot_cell(s_tiles, l, cx, cy, cz);
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
            // the next pass (or chunk) reuses LDS and the slot table; the
            // barrier at the loop head orders them
        }
    }
}

// ---------------------------------------------------------------------------
// p2g_2: density, pressure, viscous stress momentum (mls-mpm/p2g_2.wgsl:30-94)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g2_win(BatchArgs a) {
    extern __shared__ int lds[];  // [decoded mass (f32) | vx | vy | vz] x vol
    __shared__ uint32_t s_tiles[kMaxSlots];
    float *dmass = reinterpret_cast<float *>(lds);
    for (ChunkWalk cw = walk_begin(a); cw.t < cw.end; cw.t += cw.step) {
        const uint32_t c = walk_chunk(cw, a.nchunks);
        if (c == 0xffffffffu) continue;
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        const int32_t Y = sd->uni.Y, Z = sd->uni.Z;
        const uint32_t G = sd->G;
        const Params prm = sd->prm;
        int32_t *grid = (int32_t *)(a.grid + sd->goff);
        float x[PPT][3], C[PPT][3][3];
        int ic[PPT][3];
        bool valid[PPT], tame[PPT];
        load_chunk<PPT>(a.state, a.stride, ch, x, ic, valid, tame);
        uint32_t todo = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            const uint32_t i = ch.begin + slot_of<PPT>(k);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int r = 0; r < 3; ++r) C[k][cc][r] = a.state[(C00 + 3 * cc + r) * a.stride + i];
            if (tame[k]) todo |= 1u << k;
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k)
            if (valid[k] && !tame[k]) {
                Node3 n;
                node3_init(x[k], n);
                p2g2_particle(n, C[k], prm, GridSink2{grid, G, Y, Z, prm.mult});
            }
        while (__syncthreads_or(todo != 0)) {
            int slot[PPT];
            const int vol = kWinCells * block_tiles<PPT>(ic, todo, a.max_slots, slot, s_tiles);
            for (int l = threadIdx.x; l < vol; l += kBlock) {
                int cx, cy, cz;
                slot_cell(s_tiles, l, cx, cy, cz);
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
                if (slot[k] < 0) continue;
                Node3 n;
                node3_init(x[k], n);
                p2g2_particle(n, C[k], prm, WinSink2{lds, vol, tile_l0(slot[k], ic[k])});
                todo &= ~(1u << k);
            }
            __syncthreads();
            for (int l = threadIdx.x; l < vol; l += kBlock) {
                int cx, cy, cz;
                slot_cell(s_tiles, l, cx, cy, cz);
                const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
                if (!in_grid(gi, G)) continue;
                int32_t *cell = grid + 4 * (size_t)gi;
                const int32_t vx = lds[vol + l], vy = lds[2 * vol + l], vz = lds[3 * vol + l];
                if (vx) atomicAdd(cell + 0, vx);
                if (vy) atomicAdd(cell + 1, vy);
                if (vz) atomicAdd(cell + 2, vz);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// g2p with updateGrid applied on window load and copyPosition fused
// (mls-mpm/updateGrid.wgsl:23-46, g2p.wgsl:26-92, copyPosition.wgsl:15-21)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_g2p_win(BatchArgs a) {
    extern __shared__ float4 lvel[];  // decoded updated velocity x vol
    __shared__ uint32_t s_tiles[kMaxSlots];
    for (ChunkWalk cw = walk_begin(a); cw.t < cw.end; cw.t += cw.step) {
        const uint32_t c = walk_chunk(cw, a.nchunks);
        if (c == 0xffffffffu) continue;
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        const int32_t Y = sd->uni.Y, Z = sd->uni.Z;
        const int32_t wall_hi[3] = {sd->uni.wall_hi[0], sd->uni.wall_hi[1], sd->uni.wall_hi[2]};
        const uint32_t G = sd->G;
        const Params prm = sd->prm;
        const int4 *grid = a.grid + sd->goff;
        const float rb[3] = {sd->uni.real_box[0], sd->uni.real_box[1], sd->uni.real_box[2]};
        float x[PPT][3];
        int ic[PPT][3];
        bool valid[PPT], tame[PPT];
        load_chunk<PPT>(a.state, a.stride, ch, x, ic, valid, tame);
        uint32_t todo = 0;
        // gather, then the particle's tail (C, advection, walls) and stores
        auto finish = [&](int k, float v[3], float B[3][3]) {
            const uint32_t i = ch.begin + slot_of<PPT>(k);
            float xo[3], Cn[3][3];
            g2p_finish(x[k], rb, prm.dt, v, B, xo, Cn);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int r = 0; r < 3; ++r) a.state[(C00 + 3 * cc + r) * a.stride + i] = Cn[cc][r];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                a.state[(PX + ax) * a.stride + i] = xo[ax];
                a.state[(VX + ax) * a.stride + i] = v[ax];
            }
            const size_t o = a.id[i];
            a.posvel[2 * o] = make_float4(xo[0], xo[1], xo[2], 0.0f);
            a.posvel[2 * o + 1] = make_float4(v[0], v[1], v[2], 0.0f);
        };
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            if (tame[k]) {
                todo |= 1u << k;
                continue;
            }
            Node3 n;
            node3_init(x[k], n);
            float v[3], B[3][3];
            g2p_gather(n, GridSinkG{grid, G, Y, Z, {wall_hi[0], wall_hi[1], wall_hi[2]}, prm}, v,
                       B);
            finish(k, v, B);
        }
        while (__syncthreads_or(todo != 0)) {
            int slot[PPT];
            const int vol = kWinCells * block_tiles<PPT>(ic, todo, a.max_slots, slot, s_tiles);
            for (int l = threadIdx.x; l < vol; l += kBlock) {
                int cx, cy, cz;
                slot_cell(s_tiles, l, cx, cy, cz);
                const int32_t gi = cell_linear(cx, cy, cz, Y, Z);
                float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                // A cell inside [0,X) x [0,Y) x [0,Z) whose index did not wrap
                // decomposes to its own coordinates; others like updateGrid.wgsl:37-39.
                const bool nice = cx >= 0 && (uint32_t)cy < (uint32_t)Y &&
                                  (uint32_t)cz < (uint32_t)Z &&
                                  (int64_t)cx * Y * Z + (int64_t)cy * Z + cz < (int64_t)G;
                if (in_grid(gi, G)) {
                    const int4 raw = grid[gi];
                    const int4 cv = nice ? update_cell_xyz(raw, cx, cy, cz, wall_hi, prm)
                                         : update_cell_id(raw, gi, Y, Z, wall_hi, prm);
                    o = make_float4(decode(cv.x, prm.mult), decode(cv.y, prm.mult),
                                    decode(cv.z, prm.mult), 0.0f);
                }
                lvel[l] = o;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (slot[k] < 0) continue;
                Node3 n;
                node3_init(x[k], n);
                float v[3], B[3][3];
                g2p_gather(n, WinSinkG{lvel, tile_l0(slot[k], ic[k])}, v, B);
                finish(k, v, B);
                todo &= ~(1u << k);
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_clear(int4 *__restrict__ grid, uint32_t cells) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cells; i += stride)
        grid[i] = make_int4(0, 0, 0, 0);
}

__global__ void k_set_scene_uniforms(SceneDesc *d, Uniforms u) { d->uni = u; }

inline unsigned blocks_for(uint64_t count, uint64_t per) {
    return (unsigned)((count + per - 1) / per);
}

inline unsigned walk_blocks(const BatchArgs &a) {
    unsigned nb = a.nchunks;
    if (a.grid_blocks > 0 && (unsigned)a.grid_blocks < nb) nb = (unsigned)a.grid_blocks;
    return nb ? nb : 1;
}

template <int PPT>
hipError_t launch_win(int which, const BatchArgs &a, hipStream_t st) {
    if (!a.nchunks) return hipSuccess;
    const size_t lds = 16 * (size_t)kWinCells * (size_t)a.max_slots;
    const unsigned nb = walk_blocks(a);
    switch (which) {
    case 1: k_p2g1_win<PPT><<<nb, kBlock, lds, st>>>(a); break;
    case 2: k_p2g2_win<PPT><<<nb, kBlock, lds, st>>>(a); break;
    case 3: k_g2p_win<PPT><<<nb, kBlock, lds, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_win_ppt(int which, const BatchArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_win<1>(which, a, st);
    case 2: return launch_win<2>(which, a, st);
    case 4: return launch_win<4>(which, a, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

void scene_key_geometry(SceneDesc &d) {
    uint64_t nk = 1;
    for (int a = 0; a < 3; ++a) {
        int32_t dim = (int32_t)std::ceil(d.uni.init_box[a]);
        d.kdim[a] = dim < 1 ? 1 : dim;
        d.knt[a] = (uint32_t)((d.kdim[a] + (1 << kTileBits) - 1) >> kTileBits);
        nk *= d.knt[a];
    }
    d.nkeys = (uint32_t)(nk << (3 * kTileBits));
}

// 3D Hilbert index of (x, y, z) in a 2^bits cube: J. Skilling, "Programming
// the Hilbert curve", AIP Conf. Proc. 707, 381 (2004) -- axes to transposed
// form, then the transposed bits interleaved most significant first.
static uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    if (bits <= 0) return 0;
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t h = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

void append_hilbert_ranks(const SceneDesc &d, std::vector<uint32_t> &ranks) {
    const uint32_t n0 = d.knt[0], n1 = d.knt[1], n2 = d.knt[2];
    const uint32_t nt = n0 * n1 * n2;
    int bits = 0;
    while ((1u << bits) < std::max(n0, std::max(n1, n2))) ++bits;
    std::vector<std::pair<uint64_t, uint32_t>> order(nt);
    for (uint32_t x = 0; x < n0; ++x)
        for (uint32_t y = 0; y < n1; ++y)
            for (uint32_t z = 0; z < n2; ++z) {
                const uint32_t t = (x * n1 + y) * n2 + z;
                order[t] = {hilbert3(x, y, z, bits), t};
            }
    std::sort(order.begin(), order.end());
    const size_t base = ranks.size();
    ranks.resize(base + nt);
    for (uint32_t r = 0; r < nt; ++r) ranks[base + order[r].second] = r;
}

hipError_t launch_bin(const BatchArgs &a, const BinBuffers &b, hipStream_t st) {
    if (!a.nchunks) return hipSuccess;
    const unsigned nb = walk_blocks(a);
    k_bin_count<<<nb, kBlock, 0, st>>>(a, b);
    const unsigned nsb = blocks_for(b.nkeys, kScanTile);
    k_scan_reduce<<<nsb, kBlock, 0, st>>>(b.count, b.nkeys, b.block_sums);
    k_scan_top<<<1, kBlock, 0, st>>>(b.block_sums, nsb);
    k_scan_down<<<nsb, kBlock, 0, st>>>(b.count, b.offset, b.nkeys, b.block_sums);
    k_bin_scatter<<<nb, kBlock, 0, st>>>(a, b);
    return hipGetLastError();
}

hipError_t launch_clear(int4 *grid, uint32_t cells, hipStream_t st) {
    if (!cells) return hipSuccess;
    unsigned nb = blocks_for(cells, kBlock);
    if (nb > 8192) nb = 8192;
    k_clear<<<nb, kBlock, 0, st>>>(grid, cells);
    return hipGetLastError();
}

hipError_t launch_p2g1(const BatchArgs &a, hipStream_t st) { return launch_win_ppt(1, a, st); }
hipError_t launch_p2g2(const BatchArgs &a, hipStream_t st) { return launch_win_ppt(2, a, st); }
hipError_t launch_g2p(const BatchArgs &a, hipStream_t st) { return launch_win_ppt(3, a, st); }

hipError_t launch_set_scene_uniforms(SceneDesc *d, const Uniforms &u, hipStream_t st) {
    k_set_scene_uniforms<<<1, 1, 0, st>>>(d, u);
    return hipGetLastError();
}

}  // namespace mpm
