// mpm_fused_kernels.hip -- the production substep for gfx950.
//
//   bin -> p2g_1 -> reduce_1 -> p2g_2 -> reduce_2 -> g2p
//
// Particles live in SoA storage permuted into bin order (tile-major cell keys,
// tiles along a Hilbert curve); id[] maps a storage slot to the caller's
// particle index, so readout order never changes.  The binning scan also
// emits "tile chunks": runs of at most 256*ppt particles of ONE 4^3-cell
// tile.  A P2G block accumulates its chunk into the tile's 6^3-cell LDS
// window (tile + one-cell halo) with i32 fixed-point LDS adds, then stores the
// whole window with coalesced plain stores as a partial.  The reduce kernels
// give every grid cell exactly one writer: one wave per tile sums, for each of
// its 64 cells, the partials of the chunks of the 27 surrounding tiles whose
// windows cover it, then writes the cell, so there is no clear pass and no
// global atomic.  reduce_2 also applies updateGrid and stores the decoded
// velocities that g2p stages in LDS for its 27-node gathers.
//
// Integer adds commute, and every float expression is evaluated per particle
// exactly as in the reference-shaped stage kernels, so results are identical
// to them.  Particles whose stencil leaves [0, X) x [0, Y) x [0, Z) (where the
// i32 cell index wraps onto other cells), non-finite positions, and every
// particle of a scene whose init box is not integral are "wild": they are
// binned to a separate key tile and add into a separate grid-sized
// accumulator with global atomics; the reduce kernels fold it into the cells
// and clear it again, only for scenes that have wild particles.
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
constexpr unsigned kDefaultWalkBlocks = 2048;

// ---------------------------------------------------------------------------
// helpers

// Slot (within its chunk) of a thread's k-th particle.  Lanes of a wave are
// 4*PPT slots apart, so in cell-sorted storage they mostly touch different
// cells and their LDS adds rarely hit the same address in one instruction.
template <int PPT>
__device__ __forceinline__ uint32_t slot_of(int k) {
    return (threadIdx.x & 63) * (4 * PPT) + (threadIdx.x >> 6) * PPT + k;
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

// Linear cell index of integer cell coordinates, i32 wrap-around like WGSL
// (mls-mpm/p2g_1.wgsl:53-56).
__device__ __forceinline__ uint32_t cell_linear(int32_t x, int32_t y, int32_t z, int32_t Y,
                                                int32_t Z) {
    return (uint32_t)x * (uint32_t)Y * (uint32_t)Z + (uint32_t)y * (uint32_t)Z + (uint32_t)z;
}

__device__ __forceinline__ int4 add4(int4 a, int4 b) {
    return make_int4((int32_t)((uint32_t)a.x + (uint32_t)b.x),
                     (int32_t)((uint32_t)a.y + (uint32_t)b.y),
                     (int32_t)((uint32_t)a.z + (uint32_t)b.z),
                     (int32_t)((uint32_t)a.w + (uint32_t)b.w));
}

// Tile coordinates of an x-major tile index.
__device__ __forceinline__ void tile_coords(uint32_t xm, const SceneDesc *sd, int t[3]) {
    const uint32_t n12 = sd->knt[1] * sd->knt[2];
    t[0] = (int)(xm / n12);
    const uint32_t r = xm - (uint32_t)t[0] * n12;
    t[1] = (int)(r / sd->knt[2]);
    t[2] = (int)(r - (uint32_t)t[1] * sd->knt[2]);
}

// Scene owning element `v` of a per-scene offset field (binary search; the
// scenes are packed in increasing offset order).
template <typename F>
__device__ __forceinline__ uint32_t find_scene(const SceneDesc *sc, uint32_t nscenes, uint32_t v,
                                               F off) {
    uint32_t lo = 0, hi = nscenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off(sc[mid]) <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Persistent walk over the device-side tile-chunk list.  With xcd_remap the
// blocks of one XCD (block id mod 8) take a contiguous range of chunks, so
// neighbouring tiles' windows meet in the same L2.
struct TileWalk {
    uint32_t c, end, step;
};

__device__ __forceinline__ TileWalk tile_walk(const BatchArgs &a) {
    const uint32_t n = *a.ntchunks;
    TileWalk w;
    if (a.xcd_remap && (gridDim.x % kXcds) == 0) {
        const uint32_t per = (n + kXcds - 1) / kXcds, x = blockIdx.x % kXcds;
        w.c = x * per + blockIdx.x / kXcds;
        w.end = min(n, (x + 1) * per);
        w.step = gridDim.x / kXcds;
    } else {
        w.c = blockIdx.x;
        w.end = n;
        w.step = gridDim.x;
    }
    return w;
}

__device__ __forceinline__ bool in_dims(int x, int y, int z, const SceneDesc *sd) {
    return (uint32_t)x < (uint32_t)sd->kdim[0] && (uint32_t)y < (uint32_t)sd->kdim[1] &&
           (uint32_t)z < (uint32_t)sd->kdim[2];
}

// Grid cell of window cell l (0..215) of the tile with coordinates t, or -1
// outside the scene's cells.
__device__ __forceinline__ int64_t win_cell_index(int l, const int t[3], const SceneDesc *sd) {
    const int x = t[0] * kTileEdge - 1 + l / (kWinEdge * kWinEdge);
    const int y = t[1] * kTileEdge - 1 + (l / kWinEdge) % kWinEdge;
    const int z = t[2] * kTileEdge - 1 + l % kWinEdge;
    if (!in_dims(x, y, z, sd)) return -1;
    return (int64_t)sd->goff + cell_linear(x, y, z, sd->uni.Y, sd->uni.Z);
}

// LDS cell of node (0,0,0) of a particle with base cell `x` in its tile's
// window (the window starts one cell before the tile).
__device__ __forceinline__ int win_l0(const float x[3]) {
    const int m = kTileEdge - 1;
    const int cx = (int)floorf(x[0]), cy = (int)floorf(x[1]), cz = (int)floorf(x[2]);
    return ((cx & m) * kWinEdge + (cy & m)) * kWinEdge + (cz & m);
}

__device__ __forceinline__ void load_x(const float *__restrict__ st, uint32_t stride, uint32_t i,
                                       float x[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = st[(PX + a) * stride + i];
}

__device__ __forceinline__ void load_v(const float *__restrict__ st, uint32_t stride, uint32_t i,
                                       float v[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = st[(VX + a) * stride + i];
}

__device__ __forceinline__ void load_C(const float *__restrict__ st, uint32_t stride, uint32_t i,
                                       float C[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[c][r] = st[(C00 + 3 * c + r) * stride + i];
}

// ---------------------------------------------------------------------------
// sinks: where a particle's 27 node contributions go / come from

struct WinSink1 {  // p2g_1 into the LDS window: [mass | vx | vy | vz] x 216
    int *lds;
    int l0;
    __device__ __forceinline__ void add4(const Node3 &, int gx, int gy, int gz, int32_t m,
                                         int32_t x, int32_t y, int32_t z) const {
        const int l = l0 + (gx * kWinEdge + gy) * kWinEdge + gz;
        atomicAdd(&lds[l], m);
        atomicAdd(&lds[kWinCells + l], x);
        atomicAdd(&lds[2 * kWinCells + l], y);
        atomicAdd(&lds[3 * kWinCells + l], z);
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

struct WinSink2 {  // p2g_2: decoded masses [0, 216), momentum ints [216, 864)
    int *lds;
    int l0;
    __device__ __forceinline__ float mass(const Node3 &, int gx, int gy, int gz) const {
        return reinterpret_cast<const float *>(lds)[l0 + (gx * kWinEdge + gy) * kWinEdge + gz];
    }
    __device__ __forceinline__ void add3(const Node3 &, int gx, int gy, int gz, int32_t x,
                                         int32_t y, int32_t z) const {
        const int l = l0 + (gx * kWinEdge + gy) * kWinEdge + gz;
        atomicAdd(&lds[kWinCells + l], x);
        atomicAdd(&lds[2 * kWinCells + l], y);
        atomicAdd(&lds[3 * kWinCells + l], z);
    }
};

struct GridSink2 {  // masses from the completed grid, momentum to the wild accumulator
    const int32_t *grid;
    int32_t *acc;
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
        int32_t *cell = acc + 4 * (size_t)idx;
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

struct GridSinkG {  // g2p straight from the decoded velocities
    const float4 *cells;
    uint32_t G;
    int32_t Y, Z;
    __device__ __forceinline__ void vel(const Node3 &n, int gx, int gy, int gz, float cv[3]) const {
        const int32_t idx = node3_index(n, gx, gy, gz, Y, Z);
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in_grid(idx, G)) o = cells[idx];
        cv[0] = o.x;
        cv[1] = o.y;
        cv[2] = o.z;
    }
};

// ---------------------------------------------------------------------------
// binning: key per particle, counting sort, tile-chunk list

__global__ __launch_bounds__(kBlock) void k_bin_count(BatchArgs a, BinBuffers b) {
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        for (uint32_t j = threadIdx.x; j < ch.count; j += kBlock) {
            const uint32_t i = ch.begin + j;
            float x[3];
            load_x(a.state, a.stride, i, x);
            // tame: whole 3x3x3 stencil inside [0, kdim) of an integral box
            // (NaN fails every comparison and is wild)
            bool tame = sd->fast != 0;
            int cc[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const float f = floorf(x[ax]);
                tame = tame && f >= 1.0f && f <= (float)(sd->kdim[ax] - 2);
                cc[ax] = tame ? (int)f : 0;
            }
            uint32_t k;
            if (tame) {
                const uint32_t tile =
                    a.tile_rank[sd->rank_off +
                                ((uint32_t)(cc[0] >> kTileBits) * sd->knt[1] +
                                 (uint32_t)(cc[1] >> kTileBits)) *
                                    sd->knt[2] +
                                (uint32_t)(cc[2] >> kTileBits)];
                constexpr int m = kTileEdge - 1;
                const uint32_t local = ((uint32_t)(cc[0] & m) << (2 * kTileBits)) |
                                       ((uint32_t)(cc[1] & m) << kTileBits) | (uint32_t)(cc[2] & m);
                k = sd->koff + ((tile << (3 * kTileBits)) | local);
            } else {
                k = sd->koff + (sd->ntiles << (3 * kTileBits));  // the scene's wild tile
            }
            b.key[i] = k;
            b.rank[i] = atomicAdd(&b.count[k], 1u);
        }
    }
}

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096 keys = 64 key tiles

// Sum of v over the 4 consecutive lanes that share one 64-key tile.
__device__ __forceinline__ uint32_t quad_sum(uint32_t v) {
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    return v;
}

// Per block: particles and tile chunks (4 threads cover one 64-key tile).
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *__restrict__ count,
                                                        uint32_t nkeys, uint32_t chunk_cap,
                                                        uint2 *__restrict__ block_sums) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < nkeys) s += count[base + j];
    const uint32_t t = quad_sum(s);
    const uint32_t nch = (threadIdx.x & 3) == 0 ? (t + chunk_cap - 1) / chunk_cap : 0u;
    uint32_t tot_s, tot_c;
    (void)block_excl_scan(s, &tot_s);
    (void)block_excl_scan(nch, &tot_c);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = make_uint2(tot_s, tot_c);
}

__global__ __launch_bounds__(kBlock) void k_scan_top(uint2 *__restrict__ block_sums,
                                                     uint32_t nblocks,
                                                     uint32_t *__restrict__ ntchunks) {
    uint32_t carry_s = 0, carry_c = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const uint2 v = i < nblocks ? block_sums[i] : make_uint2(0u, 0u);
        uint32_t ts, tc;
        const uint32_t es = block_excl_scan(v.x, &ts);
        const uint32_t ec = block_excl_scan(v.y, &tc);
        if (i < nblocks) block_sums[i] = make_uint2(carry_s + es, carry_c + ec);
        carry_s += ts;
        carry_c += tc;
    }
    if (threadIdx.x == 0) *ntchunks = carry_c;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(BatchArgs a, BinBuffers b,
                                                      uint32_t chunk_cap) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < b.nkeys) ? b.count[base + j] : 0;
        s += v[j];
    }
    const uint32_t t = quad_sum(s);
    const bool leader = (threadIdx.x & 3) == 0 && base < b.nkeys;
    const uint32_t nch = leader ? (t + chunk_cap - 1) / chunk_cap : 0u;
    const uint2 bs = b.block_sums[blockIdx.x];
    uint32_t tot;
    uint32_t run = bs.x + block_excl_scan(s, &tot);
    const uint32_t cbase = bs.y + block_excl_scan(nch, &tot);
    const uint32_t tile_start = run;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < b.nkeys) {
            b.offset[base + j] = run;
            run += v[j];
            b.count[base + j] = 0;  // ready for the next binning pass
        }
    if (leader) {
        const uint32_t ktile = base / kTileCells;  // key tile across all scenes
        b.tile_chunk[ktile] = make_uint2(cbase, nch);
        if (nch) {
            const uint32_t sc = find_scene(a.scenes, a.nscenes, ktile,
                                           [](const SceneDesc &d) { return d.toff; });
            const uint32_t xm = a.tile_xmaj[ktile];
            for (uint32_t j = 0; j < nch; ++j)
                b.tchunks[cbase + j] = make_uint4(sc, xm, tile_start + j * chunk_cap,
                                                  min(chunk_cap, t - j * chunk_cap));
        }
    }
}

// Keys of scene s lie in [koff_s, koff_s + nkeys_s) and scenes are packed in
// slot order, so the scattered slot stays inside the scene's slot range.
__global__ __launch_bounds__(kBlock) void k_bin_scatter(BatchArgs a, BinBuffers b) {
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
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
// p2g_1: mass and APIC momentum (mls-mpm/p2g_1.wgsl:24-65) into tile partials

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g1(BatchArgs a) {
    __shared__ int lds[4 * kWinCells];  // [mass | vx | vy | vz]
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        const SceneDesc *sd = a.scenes + ch.x;
        const float mult = sd->prm.mult;
        if (ch.y == kWildTile) {
            const GridSink1 sink{(int32_t *)(a.wild + sd->goff), sd->G, sd->uni.Y, sd->uni.Z};
            for (uint32_t j = threadIdx.x; j < ch.w; j += kBlock) {
                float x[3], v[3], C[3][3];
                load_x(a.state, a.stride, ch.z + j, x);
                load_v(a.state, a.stride, ch.z + j, v);
                load_C(a.state, a.stride, ch.z + j, C);
                Node3 n;
                node3_init(x, n);
                p2g1_particle(n, v, C, mult, sink);
            }
            continue;
        }
        float x[PPT][3], v[PPT][3], C[PPT][3][3];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = slot_of<PPT>(k);
            valid[k] = j < ch.w;
            if (valid[k]) {
                load_x(a.state, a.stride, ch.z + j, x[k]);
                load_v(a.state, a.stride, ch.z + j, v[k]);
                load_C(a.state, a.stride, ch.z + j, C[k]);
            }
        }
        for (int l = threadIdx.x; l < 4 * kWinCells; l += kBlock) lds[l] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            Node3 n;
            node3_init(x[k], n);
            p2g1_particle(n, v[k], C[k], mult, WinSink1{lds, win_l0(x[k])});
        }
        __syncthreads();
        int4 *out = a.partial + (size_t)w.c * kWinCells;
        for (int l = threadIdx.x; l < kWinCells; l += kBlock)
            out[l] = make_int4(lds[kWinCells + l], lds[2 * kWinCells + l], lds[3 * kWinCells + l],
                               lds[l]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// reduce: one wave per spatial tile (tiles in Hilbert order).  Lanes 0..26
// fetch the chunk ranges of the 27 neighbour tiles; then each lane owns one
// of the tile's 64 cells, which lies in the window of its own tile plus, per
// axis, the neighbour below (local coordinate 0) or above (3): at most 8
// windows.  The partial loads of all those windows are issued together, one
// round per chunk of the fullest neighbour.

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) v = max(v, (uint32_t)__shfl_xor(v, o, kWave));
    return v;
}

// Spatial tile of this block: with xcd_remap the blocks of one XCD take a
// contiguous range of tiles, so neighbouring partials meet in one L2.  The
// grid is a multiple of 8 blocks; tiles >= stiles are padding.
__device__ __forceinline__ uint32_t reduce_tile(const BatchArgs &a) {
    if (a.xcd_remap) {
        const uint32_t per = gridDim.x / kXcds;
        return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    }
    return blockIdx.x;
}

struct CellSum {
    int4 acc;         // sum of the partials covering the lane's cell
    uint32_t rounds;  // chunks of the fullest neighbour (wave-uniform)
};

__device__ __forceinline__ CellSum sum_partials(const BatchArgs &a, uint32_t b, const int l[3],
                                                bool cell_ok) {
    const int lane = threadIdx.x;
    uint32_t nfirst = 0, ncnt = 0;
    if (lane < 27) {
        const uint32_t kt = a.nbr[(size_t)b * 27 + lane];
        if (kt != kNoTile) {
            const uint2 tc = a.tile_chunk[kt];
            nfirst = tc.x;
            ncnt = tc.y;
        }
    }
    int ext[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) ext[ax] = l[ax] == 0 ? -1 : (l[ax] == kTileEdge - 1 ? 1 : 0);
    uint32_t base[8], cnt[8], mc = 0;
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
        const int q = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1);
        const uint32_t f = __shfl(nfirst, q, kWave), c = __shfl(ncnt, q, kWave);
        const int widx = ((l[0] + 1 - kTileEdge * d[0]) * kWinEdge + (l[1] + 1 - kTileEdge * d[1])) *
                             kWinEdge +
                         (l[2] + 1 - kTileEdge * d[2]);
        base[s] = f * (uint32_t)kWinCells + (uint32_t)widx;
        cnt[s] = valid ? c : 0u;
        mc = max(mc, cnt[s]);
    }
    CellSum r;
    r.acc = make_int4(0, 0, 0, 0);
    r.rounds = wave_max(mc);
    for (uint32_t j = 0; j < r.rounds; ++j) {
        int4 v[8];
#pragma unroll
        for (int s = 0; s < 8; ++s)
            v[s] = j < cnt[s] ? a.partial[(size_t)base[s] + (size_t)j * kWinCells]
                              : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 8; ++s) r.acc = add4(r.acc, v[s]);
    }
    return r;
}

// Folds the wild accumulator of cell gi into v and clears it.
__device__ __forceinline__ void take_wild(const BatchArgs &a, uint64_t gi, int4 &v, bool mass) {
    const int4 w = a.wild[gi];
    v = add4(v, make_int4(w.x, w.y, w.z, mass ? w.w : 0));
    a.wild[gi] = make_int4(0, 0, 0, 0);
}

template <int PHASE>
__global__ __launch_bounds__(kWave) void k_reduce(BatchArgs a) {
    const uint32_t b = reduce_tile(a);
    if (b >= a.stiles) return;
    const uint4 ti = a.tile_info[b];
    const SceneDesc *sd = a.scenes + ti.w;
    const bool has_wild = a.tile_chunk[sd->toff + sd->ntiles].y != 0;
    const Params prm = sd->prm;
    const int32_t wall_hi[3] = {sd->uni.wall_hi[0], sd->uni.wall_hi[1], sd->uni.wall_hi[2]};
    if (!sd->fast) {  // every particle is wild: cells in linear order
        const uint32_t li = (b - sd->soff) * kTileCells + threadIdx.x;
        if (li >= sd->G) return;
        const uint64_t gi = (uint64_t)sd->goff + li;
        if (PHASE == 1) {
            int4 g = make_int4(0, 0, 0, 0);
            if (has_wild) take_wild(a, gi, g, true);
            a.grid[gi] = g;
        } else {
            int4 g = a.grid[gi];
            if (has_wild) take_wild(a, gi, g, false);
            a.grid[gi] = g;
            const int4 u = update_cell_id(g, (int32_t)li, sd->uni.Y, sd->uni.Z, wall_hi, prm);
            a.vel[gi] = make_float4(decode(u.x, prm.mult), decode(u.y, prm.mult),
                                    decode(u.z, prm.mult), 0.0f);
        }
        return;
    }
    const int l[3] = {(int)threadIdx.x >> 4, ((int)threadIdx.x >> 2) & 3, (int)threadIdx.x & 3};
    const int cx = (int)ti.x * kTileEdge + l[0], cy = (int)ti.y * kTileEdge + l[1],
              cz = (int)ti.z * kTileEdge + l[2];
    const bool cell_ok = in_dims(cx, cy, cz, sd);
    const uint64_t gi = (uint64_t)sd->goff + cell_linear(cx, cy, cz, sd->uni.Y, sd->uni.Z);
    if (PHASE == 1) {
        CellSum cs = sum_partials(a, b, l, cell_ok);
        if (!cell_ok) return;
        if (has_wild) take_wild(a, gi, cs.acc, true);
        a.grid[gi] = cs.acc;
    } else {
        int4 g = make_int4(0, 0, 0, 0);
        if (cell_ok) g = a.grid[gi];
        const CellSum cs = sum_partials(a, b, l, cell_ok);
        if (!cell_ok) return;
        if (cs.rounds == 0 && !has_wild) {  // no particle nearby: the cell stays empty
            a.vel[gi] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return;
        }
        g = add4(g, make_int4(cs.acc.x, cs.acc.y, cs.acc.z, 0));
        if (has_wild) take_wild(a, gi, g, false);
        a.grid[gi] = g;
        const int4 u = update_cell_xyz(g, cx, cy, cz, wall_hi, prm);
        a.vel[gi] = make_float4(decode(u.x, prm.mult), decode(u.y, prm.mult),
                                decode(u.z, prm.mult), 0.0f);
    }
}

// ---------------------------------------------------------------------------
// p2g_2: density, pressure, viscous stress momentum (mls-mpm/p2g_2.wgsl:30-94)

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_p2g2(BatchArgs a) {
    __shared__ int lds[4 * kWinCells];  // [decoded mass (f32) | vx | vy | vz]
    float *dm = reinterpret_cast<float *>(lds);
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        const SceneDesc *sd = a.scenes + ch.x;
        const Params prm = sd->prm;
        if (ch.y == kWildTile) {
            const GridSink2 sink{(const int32_t *)(a.grid + sd->goff),
                                 (int32_t *)(a.wild + sd->goff), sd->G, sd->uni.Y, sd->uni.Z,
                                 prm.mult};
            for (uint32_t j = threadIdx.x; j < ch.w; j += kBlock) {
                float x[3], C[3][3];
                load_x(a.state, a.stride, ch.z + j, x);
                load_C(a.state, a.stride, ch.z + j, C);
                Node3 n;
                node3_init(x, n);
                p2g2_particle(n, C, prm, sink);
            }
            continue;
        }
        float x[PPT][3], C[PPT][3][3];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = slot_of<PPT>(k);
            valid[k] = j < ch.w;
            if (valid[k]) {
                load_x(a.state, a.stride, ch.z + j, x[k]);
                load_C(a.state, a.stride, ch.z + j, C[k]);
            }
        }
        int t[3];
        tile_coords(ch.y, sd, t);
        for (int l = threadIdx.x; l < kWinCells; l += kBlock) {
            const int64_t gi = win_cell_index(l, t, sd);
            dm[l] = gi >= 0 ? decode(a.grid[gi].w, prm.mult) : 0.0f;
            lds[kWinCells + l] = 0;
            lds[2 * kWinCells + l] = 0;
            lds[3 * kWinCells + l] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            Node3 n;
            node3_init(x[k], n);
            p2g2_particle(n, C[k], prm, WinSink2{lds, win_l0(x[k])});
        }
        __syncthreads();
        int4 *out = a.partial + (size_t)w.c * kWinCells;
        for (int l = threadIdx.x; l < kWinCells; l += kBlock)
            out[l] = make_int4(lds[kWinCells + l], lds[2 * kWinCells + l], lds[3 * kWinCells + l], 0);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// g2p with copyPosition fused (mls-mpm/g2p.wgsl:26-92, copyPosition.wgsl:15-21)

__device__ __forceinline__ void g2p_store(const BatchArgs &a, uint32_t i, const float xo[3],
                                          const float v[3], const float Cn[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) a.state[(C00 + 3 * c + r) * a.stride + i] = Cn[c][r];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        a.state[(PX + ax) * a.stride + i] = xo[ax];
        a.state[(VX + ax) * a.stride + i] = v[ax];
    }
    const size_t o = a.id[i];
    a.posvel[2 * o] = make_float4(xo[0], xo[1], xo[2], 0.0f);
    a.posvel[2 * o + 1] = make_float4(v[0], v[1], v[2], 0.0f);
}

template <int PPT>
__global__ __launch_bounds__(kBlock) void k_g2p(BatchArgs a) {
    __shared__ float4 lvel[kWinCells];
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        const SceneDesc *sd = a.scenes + ch.x;
        const float dt = sd->prm.dt;
        const float rb[3] = {sd->uni.real_box[0], sd->uni.real_box[1], sd->uni.real_box[2]};
        if (ch.y == kWildTile) {
            const GridSinkG sink{a.vel + sd->goff, sd->G, sd->uni.Y, sd->uni.Z};
            for (uint32_t j = threadIdx.x; j < ch.w; j += kBlock) {
                const uint32_t i = ch.z + j;
                float x[3], v[3], B[3][3], xo[3], Cn[3][3];
                load_x(a.state, a.stride, i, x);
                Node3 n;
                node3_init(x, n);
                g2p_gather(n, sink, v, B);
                g2p_finish(x, rb, dt, v, B, xo, Cn);
                g2p_store(a, i, xo, v, Cn);
            }
            continue;
        }
        float x[PPT][3];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = slot_of<PPT>(k);
            valid[k] = j < ch.w;
            if (valid[k]) load_x(a.state, a.stride, ch.z + j, x[k]);
        }
        int t[3];
        tile_coords(ch.y, sd, t);
        for (int l = threadIdx.x; l < kWinCells; l += kBlock) {
            const int64_t gi = win_cell_index(l, t, sd);
            lvel[l] = gi >= 0 ? a.vel[gi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            float v[3], B[3][3], xo[3], Cn[3][3];
            Node3 n;
            node3_init(x[k], n);
            g2p_gather(n, WinSinkG{lvel, win_l0(x[k])}, v, B);
            g2p_finish(x[k], rb, dt, v, B, xo, Cn);
            g2p_store(a, ch.z + slot_of<PPT>(k), xo, v, Cn);
        }
        __syncthreads();
    }
}

__global__ void k_set_scene_uniforms(SceneDesc *d, Uniforms u) { d->uni = u; }

inline unsigned blocks_for(uint64_t count, uint64_t per) {
    return (unsigned)((count + per - 1) / per);
}

inline unsigned walk_blocks(const BatchArgs &a) {
    const unsigned nb = a.grid_blocks > 0 ? (unsigned)a.grid_blocks : kDefaultWalkBlocks;
    return ((nb + kXcds - 1) / kXcds) * kXcds;
}

template <int PPT>
hipError_t launch_ppt(int which, const BatchArgs &a, hipStream_t st) {
    const unsigned nb = walk_blocks(a);
    switch (which) {
    case 1: k_p2g1<PPT><<<nb, kBlock, 0, st>>>(a); break;
    case 2: k_p2g2<PPT><<<nb, kBlock, 0, st>>>(a); break;
    case 3: k_g2p<PPT><<<nb, kBlock, 0, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_by_ppt(int which, const BatchArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_ppt<1>(which, a, st);
    case 2: return launch_ppt<2>(which, a, st);
    case 4: return launch_ppt<4>(which, a, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

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

void append_hilbert_ranks(const SceneDesc &d, std::vector<uint32_t> &ranks,
                          std::vector<uint32_t> &xmaj) {
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
    for (uint32_t r = 0; r < nt; ++r) {
        ranks[base + order[r].second] = r;
        xmaj.push_back(order[r].second);
    }
    xmaj.push_back(kWildTile);
}

void append_tile_tables(const SceneDesc &d, uint32_t s, const uint32_t *ranks,
                        const uint32_t *xmaj, std::vector<uint32_t> &nbr,
                        std::vector<uint4> &info) {
    const uint32_t n1 = d.knt[1], n2 = d.knt[2];
    for (uint32_t r = 0; r < d.ntiles; ++r) {
        const uint32_t xm = xmaj[r];
        const uint32_t t[3] = {xm / (n1 * n2), (xm / n2) % n1, xm % n2};
        info.push_back(make_uint4(t[0], t[1], t[2], s));
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int64_t nx = (int64_t)t[0] + dx, ny = (int64_t)t[1] + dy,
                                  nz = (int64_t)t[2] + dz;
                    if (nx < 0 || ny < 0 || nz < 0 || nx >= d.knt[0] || ny >= n1 || nz >= n2) {
                        nbr.push_back(kNoTile);
                        continue;
                    }
                    nbr.push_back(d.toff + ranks[((uint32_t)nx * n1 + (uint32_t)ny) * n2 +
                                                 (uint32_t)nz]);
                }
    }
}

void scene_key_geometry(SceneDesc &d) {
    uint64_t nt = 1;
    bool integral = true;
    for (int a = 0; a < 3; ++a) {
        const float b = d.uni.init_box[a];
        const int32_t dim = (int32_t)std::ceil(b);
        d.kdim[a] = dim < 1 ? 1 : dim;
        d.knt[a] = (uint32_t)((d.kdim[a] + kTileEdge - 1) >> kTileBits);
        nt *= d.knt[a];
        integral = integral && std::floor(b) == b;
    }
    d.ntiles = (uint32_t)nt;
    d.nkeys = (uint32_t)((nt + 1) << (3 * kTileBits));
    d.fast = integral ? 1u : 0u;
}

hipError_t launch_bin(const BatchArgs &a, const BinBuffers &b, hipStream_t st) {
    const uint32_t cap = (uint32_t)kBlock * (uint32_t)a.ppt;
    const unsigned nbin = std::max(1u, std::min(a.nchunks, 4096u));
    if (a.nchunks) k_bin_count<<<nbin, kBlock, 0, st>>>(a, b);
    const unsigned nsb = blocks_for(b.nkeys, kScanTile);
    k_scan_reduce<<<nsb, kBlock, 0, st>>>(b.count, b.nkeys, cap, b.block_sums);
    k_scan_top<<<1, kBlock, 0, st>>>(b.block_sums, nsb, b.ntchunks);
    k_scan_down<<<nsb, kBlock, 0, st>>>(a, b, cap);
    if (a.nchunks) k_bin_scatter<<<nbin, kBlock, 0, st>>>(a, b);
    return hipGetLastError();
}

hipError_t launch_p2g1(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(1, a, st); }
hipError_t launch_p2g2(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(2, a, st); }
hipError_t launch_g2p(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(3, a, st); }

hipError_t launch_reduce1(const BatchArgs &a, hipStream_t st) {
    if (a.stiles) k_reduce<1><<<(a.stiles + kXcds - 1) / kXcds * kXcds, kWave, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_reduce2(const BatchArgs &a, hipStream_t st) {
    if (a.stiles) k_reduce<2><<<(a.stiles + kXcds - 1) / kXcds * kXcds, kWave, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_set_scene_uniforms(SceneDesc *d, const Uniforms &u, hipStream_t st) {
    k_set_scene_uniforms<<<1, 1, 0, st>>>(d, u);
    return hipGetLastError();
}

}  // namespace mpm
