// mpm_index.h -- index arithmetic of the production substep.
//
// Pure functions of their arguments, shared by the gfx950 kernels
// (mpm_fused_kernels.hip, which pass blockIdx / gridDim) and by the host-side
// emulation that checks every index they form stays inside its buffer
// (tests/emu/chunk_index_emu.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_fused.h"

namespace mpm {

// Reduce launches: one lane per cell of a tile.  A 4^3 tile is one wave, and
// a block packs kTileWaves waves, i.e. consecutive tiles (fewer, larger
// blocks dispatch faster); an 8^3 tile is a block of 8 waves.
constexpr uint32_t kTileWaves = 4;

template <int TB>
struct ReduceGeo {
    static constexpr uint32_t waves_per_tile = (uint32_t)TileGeo<TB>::cells / 64u;  // 1 | 8
    static constexpr uint32_t tiles_per_block = TB == 2 ? kTileWaves : 1u;
    static constexpr uint32_t threads = tiles_per_block * waves_per_tile * 64u;   // 256 | 512
};

// Runs of consecutive work items dealt round robin over the XCDs (see
// tile_walk_of).
constexpr uint32_t kXcdRun = 16;

constexpr uint32_t kDirs = 27;              // near directions (dx+1)*9 + (dy+1)*3 + dz+1
constexpr uint32_t kDirWild = kDirs;        // became wild: block-level global atomic
constexpr uint32_t kDirFar = kDirs + 1;     // moved further: per-particle global atomic

// Blocks of a reduce launch, padded to a multiple of 8 runs (tiles >= stiles
// are padding).
template <int TB = 2>
__host__ __device__ __forceinline__ uint32_t tile_blocks(uint32_t stiles) {
    constexpr uint32_t per = ReduceGeo<TB>::tiles_per_block;
    const uint32_t nb = (stiles + per - 1) / per;
    return (nb + kXcds * kXcdRun - 1) / (kXcds * kXcdRun) * (kXcds * kXcdRun);
}

// Spatial tile of wave `wave` of block `block` of a reduce launch.
// The dispatcher puts block b on XCD b mod 8; blocks take Hilbert-consecutive
// tiles in runs of kXcdRun blocks per XCD turn, so neighbouring partial
// windows mostly meet in one L2 while every XCD gets a share of the fluid's
// dense and sparse regions (with xcd_remap: one contiguous range of tiles per
// XCD).  Both are bijections of the padded block range.
template <int TB = 2>
__host__ __device__ __forceinline__ uint32_t reduce_tile_of(uint32_t block, uint32_t wave,
                                                           uint32_t stiles, int xcd_remap) {
    const uint32_t x = block % kXcds, j = block / kXcds;
    uint32_t blk;
    if (xcd_remap)
        blk = x * (tile_blocks<TB>(stiles) / kXcds) + j;
    else
        blk = ((j / kXcdRun) * kXcds + x) * kXcdRun + j % kXcdRun;
    constexpr uint32_t per = ReduceGeo<TB>::tiles_per_block;
    return blk * per + (per > 1 ? wave : 0u);
}

__host__ __device__ __forceinline__ uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Persistent walk of block `block` of `grid` over the n tile chunks of the
// current binning (n = min(chunk count, capacity)).  By default the chunks
// are dealt to the XCDs in runs of kXcdRun consecutive (Hilbert-ordered, so
// neighbouring) chunks: neighbouring windows share an L2 within a run while
// every XCD gets an even mix of the fluid's dense body and sparse surface.
// With xcd_remap each XCD takes one contiguous range of chunks holding an
// equal share of the particles, as the scan split them (xcd_first).
struct TileWalk {
    uint32_t c, end, step;
};

// First chunk of block `block` in the default walk (runs of kXcdRun chunks
// dealt round robin over the XCDs; a bijection of [0, grid) when grid is a
// multiple of kXcds * kXcdRun).  It does not depend on the chunk count, so a
// block can load its first chunk's descriptor beside the count.
__host__ __device__ __forceinline__ uint32_t walk_first_of(uint32_t block, uint32_t grid) {
    if ((grid % (kXcds * kXcdRun)) == 0) {
        const uint32_t x = block % kXcds, j = block / kXcds;
        return ((j / kXcdRun) * kXcds + x) * kXcdRun + j % kXcdRun;
    }
    return block;
}

__host__ __device__ __forceinline__ TileWalk tile_walk_of(uint32_t block, uint32_t grid,
                                                         uint32_t n, const uint32_t *xcd_first,
                                                         int xcd_remap) {
    TileWalk w;
    if (xcd_remap && (grid % kXcds) == 0) {
        const uint32_t x = block % kXcds;
        w.c = min_u32(xcd_first[x], n) + block / kXcds;
        w.end = min_u32(xcd_first[x + 1], n);
        w.step = grid / kXcds;
    } else {
        w.c = walk_first_of(block, grid);
        w.end = n;
        w.step = grid;
    }
    return w;
}

// Descriptor slot a block reads for chunk c before it knows whether c is
// below the chunk count: c itself, or the last slot of the cap descriptors
// the walk may use (slot 0 when cap is 0; the buffer always holds one).  The
// value read is used only for c < count <= cap.
__host__ __device__ __forceinline__ uint32_t spec_chunk_of(uint32_t c, uint32_t cap) {
    return c < cap ? c : (cap ? cap - 1u : 0u);
}

// Linear cell index of integer cell coordinates, i32 wrap-around like WGSL
// (mls-mpm/p2g_1.wgsl:53-56).
__host__ __device__ __forceinline__ uint32_t cell_linear(int32_t x, int32_t y, int32_t z,
                                                         int32_t Y, int32_t Z) {
    return (uint32_t)x * (uint32_t)Y * (uint32_t)Z + (uint32_t)y * (uint32_t)Z + (uint32_t)z;
}

// Tile coordinates of an x-major tile index.
__host__ __device__ __forceinline__ void tile_coords(uint32_t xm, const SceneDesc *sd, int t[3]) {
    const uint32_t n12 = sd->knt[1] * sd->knt[2];
    t[0] = (int)(xm / n12);
    const uint32_t r = xm - (uint32_t)t[0] * n12;
    t[1] = (int)(r / sd->knt[2]);
    t[2] = (int)(r - (uint32_t)t[1] * sd->knt[2]);
}

__host__ __device__ __forceinline__ bool in_dims(int x, int y, int z, const SceneDesc *sd) {
    return (uint32_t)x < (uint32_t)sd->kdim[0] && (uint32_t)y < (uint32_t)sd->kdim[1] &&
           (uint32_t)z < (uint32_t)sd->kdim[2];
}

// Grid cell of window cell l (0..win_cells-1) of the tile with coordinates
// t, or -1 outside the scene's cells.
template <int TB = 2>
__host__ __device__ __forceinline__ int64_t win_cell_index(int l, const int t[3],
                                                          const SceneDesc *sd) {
    constexpr int E = TileGeo<TB>::edge, W = TileGeo<TB>::win_edge;
    const int x = t[0] * E - 1 + l / (W * W);
    const int y = t[1] * E - 1 + (l / W) % W;
    const int z = t[2] * E - 1 + l % W;
    if (!in_dims(x, y, z, sd)) return -1;
    return (int64_t)sd->goff + cell_linear(x, y, z, sd->uni.Y, sd->uni.Z);
}

// LDS cell of node (0,0,0) of a particle with base cell `x` in its tile's
// window (the window starts one cell before the tile).
template <int TB = 2>
__host__ __device__ __forceinline__ int win_l0(const float x[3]) {
    constexpr int m = TileGeo<TB>::edge - 1, W = TileGeo<TB>::win_edge;
    const int cx = (int)floorf(x[0]), cy = (int)floorf(x[1]), cz = (int)floorf(x[2]);
    return ((cx & m) * W + (cy & m)) * W + (cz & m);
}

// Bin key of a particle at x: its tile's key, or the scene's wild key when
// the 3x3x3 stencil is not inside [0, kdim) of an integral box (NaN fails
// every comparison).  *dir receives the direction from tile `from` (kDirFar
// when further away or when `from` is not a tile, kDirWild when wild).
template <int TB = 2>
__host__ __device__ __forceinline__ uint32_t bin_key_of(const uint32_t *tile_rank,
                                                       const SceneDesc *sd, const float x[3],
                                                       const int from[3], uint32_t *dir) {
    bool tame = sd->fast != 0;
    int t[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float f = floorf(x[ax]);
        tame = tame && f >= 1.0f && f <= (float)(sd->kdim[ax] - 2);
        t[ax] = tame ? (int32_t)f >> TB : 0;
    }
    if (!tame) {
        *dir = kDirWild;
        return sd->koff + sd->ntiles;
    }
    const int dx = t[0] - from[0], dy = t[1] - from[1], dz = t[2] - from[2];
    const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
    *dir = near ? (uint32_t)((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)) : kDirFar;
    return sd->koff + tile_rank[sd->rank_off +
                                ((uint32_t)t[0] * sd->knt[1] + (uint32_t)t[1]) * sd->knt[2] +
                                (uint32_t)t[2]];
}

__host__ __device__ __forceinline__ uint32_t chunks_of(uint32_t n, uint32_t cap) {
    return (n + cap - 1) / cap;
}

// The covering slot s (0..7) of tile-local cell l[3] in a per-tile reduce:
// per axis bit s>>ax selects the tile itself or its neighbour below (local
// coordinate 0) or above (edge - 1).  *q receives the neighbour (27-entry
// order) and the return value the cell's index in that neighbour's window;
// *valid is false when the slot needs a neighbour the cell does not touch.
template <int TB = 2>
__host__ __device__ __forceinline__ int cover_slot(const int l[3], int s, int *q, bool *valid) {
    constexpr int E = TileGeo<TB>::edge, W = TileGeo<TB>::win_edge;
    bool ok = true;
    int d[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int ext = l[ax] == 0 ? -1 : (l[ax] == E - 1 ? 1 : 0);
        const bool bit = (s >> ax) & 1;
        d[ax] = bit ? ext : 0;
        ok = ok && !(bit && ext == 0);
    }
    *q = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1);
    *valid = ok;
    return ((l[0] + 1 - E * d[0]) * W + (l[1] + 1 - E * d[1])) * W + (l[2] + 1 - E * d[2]);
}

// Level order of a 4^3-tile chunk (the P2G kernels' lane assignment).  With
// cnt[c] particles in tile cell c, levels 0..kLevelCap-1 are level-major: level
// k holds the k-th particle of every cell with more than k, cells in index
// order (mask[k]: those cells, base[k]: the level's first position).  A
// crowded cell's particles beyond level kLevelCap-1 follow cell-major after
// all the levels, from exoff[c].
constexpr uint32_t kLevelCap = 32;

// Position of the particle at level `lev` of tile cell `cell` (0..63).
__host__ __device__ __forceinline__ uint32_t level_position(uint32_t lev, uint32_t cell,
                                                           const uint32_t *base,
                                                           const uint64_t *mask,
                                                           const uint32_t *exoff) {
    if (lev >= kLevelCap) return exoff[cell] + (lev - kLevelCap);
    const uint64_t below = (1ull << cell) - 1ull;
    return base[lev] + (uint32_t)__builtin_popcountll(mask[lev] & below);
}

// Slot where range x (1..kXcds-1) of the XCD split of `total` particles
// starts: ceil(x * total / kXcds).
__host__ __device__ __forceinline__ uint32_t xcd_split_slot(uint32_t x, uint32_t total) {
    return (uint32_t)(((uint64_t)x * total + kXcds - 1) / kXcds);
}

}  // namespace mpm
