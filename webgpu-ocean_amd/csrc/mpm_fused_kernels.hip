// mpm_fused_kernels.hip -- the production substep for gfx950.
//
//   bin -> p2g_1 -> reduce_1 -> p2g_2 -> reduce_2 -> g2p
//
// Particles live in SoA storage in the bin order of the previous substep
// (tiles along a Hilbert curve); id[] maps a storage slot to the caller's
// particle index, so readout order never changes.  The binning computes a
// permutation into this substep's bin order without moving the state: the
// chunk kernels read through it, and g2p writes its output state in bin
// order into the other buffer.  The binning scan also emits "tile chunks":
// runs of at most 256*ppt particles of ONE tile (4^3 or 8^3 cells, the
// batch's tile_bits; kernels are instantiated for both).  A P2G block
// accumulates its chunk into the tile's LDS window (tile + one-cell halo:
// 6^3 or 10^3 cells) with i32 fixed-point LDS adds, then stores the whole
// window with coalesced plain stores as a partial.  The reduce kernels give
// every grid cell exactly one writer: one lane per cell sums the partials of
// the chunks of the 27 surrounding tiles whose
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
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <utility>
#include <vector>

#include "mpm_fused.h"
#include "mpm_index.h"
#include "mpm_particle.h"

#pragma clang fp contract(off)

namespace mpm {
namespace {

constexpr int kWave = 64;
constexpr unsigned kDefaultWalkBlocks = 4096;

// ---------------------------------------------------------------------------
// helpers

// A chunk of `count` particles is ceil(count / 64) wave tasks; task t of the
// block's waves is k * 4 + wave for the thread's k-th particle.
__device__ __forceinline__ uint32_t task_of(int k) { return (uint32_t)k * 4 + (threadIdx.x >> 6); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) v = max(v, (uint32_t)__shfl_xor(v, o, kWave));
    return v;
}

template <int TB>
__device__ __forceinline__ uint32_t reduce_tile(const BatchArgs &a) {
    // (the wave index is wave-uniform: say so, so that the tile is too)
    return reduce_tile_of<TB>(blockIdx.x, __builtin_amdgcn_readfirstlane(threadIdx.x / kWave),
                              a.stiles, a.xcd_remap);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) v += (uint32_t)__shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ int4 add4(int4 a, int4 b) {
    return make_int4((int32_t)((uint32_t)a.x + (uint32_t)b.x),
                     (int32_t)((uint32_t)a.y + (uint32_t)b.y),
                     (int32_t)((uint32_t)a.z + (uint32_t)b.z),
                     (int32_t)((uint32_t)a.w + (uint32_t)b.w));
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

// Scene of bin key k (no lookup for a batch of one).
__device__ __forceinline__ uint32_t key_scene_of(const BatchArgs &a, uint32_t k) {
    return a.nscenes == 1 ? 0u : a.key_scene[k];
}

__device__ __forceinline__ TileWalk tile_walk(const BatchArgs &a) {
    return tile_walk_of(blockIdx.x, gridDim.x, min(*a.ntchunks, a.tchunk_cap), a.xcd_first,
                        a.xcd_remap);
}

// A chunk kernel block's walk with the descriptor of its current chunk.  In
// the default walk the first chunk's index does not depend on the chunk
// count, so the descriptor, the count and the wild-chunk count (and, for a
// batch of one scene, the scene's descriptor) are loaded together and waited
// for once; each iteration loads the next descriptor before its own particle
// loads.
struct ChunkWalk {
    TileWalk w;
    TileChunk ch;    // descriptor of chunk w.c (meaningful while w.c < w.end)
    uint32_t wild;   // tile chunks of the last key (a single scene's wild key)

    template <bool kSingle>
    __device__ __forceinline__ ChunkWalk(const BatchArgs &a, SceneDesc &scene,
                                         std::integral_constant<bool, kSingle>) {
        if constexpr (kSingle) scene = load_const(a.scenes);
        const uint32_t grid = a.walk_grid;  // = gridDim.x
        w.c = walk_first_of(blockIdx.x, grid);
        ch = load_const(a.tchunks + spec_chunk_of(w.c, a.tchunk_cap));
        const uint2 head = load_const(reinterpret_cast<const uint2 *>(a.ntchunks));
        // Wait for these loads together, here: left to itself the compiler
        // issues each at its first use, the descriptor's after the count's
        // wait (an empty asm that needs the values, not a load).
        if constexpr (kSingle)
            asm volatile("" ::"s"(head.x), "s"(head.y), "s"(ch.x), "s"(ch.y), "s"(ch.z),
                         "s"(ch.w), "s"(scene.goff), "s"(scene.uni.Y), "s"(scene.uni.Z),
                         "s"(scene.knt[1]), "s"(scene.knt[2]), "s"(scene.kdim[0]),
                         "s"(scene.kdim[1]), "s"(scene.kdim[2]), "s"(scene.prm.mult));
        else
            asm volatile("" ::"s"(head.x), "s"(head.y), "s"(ch.x), "s"(ch.y), "s"(ch.z),
                         "s"(ch.w));
        wild = head.y;
        w.end = min(head.x, a.tchunk_cap);
        w.step = grid;
        if (a.xcd_remap && (grid % kXcds) == 0) {  // XCD ranges: another round of loads
            w = tile_walk_of(blockIdx.x, grid, w.end, a.xcd_first, 1);
            ch = load_const(a.tchunks + spec_chunk_of(w.c, a.tchunk_cap));
        }
    }
    // Returns the current chunk's descriptor and starts loading the next one.
    __device__ __forceinline__ TileChunk take(const BatchArgs &a) {
        const TileChunk cur = ch;
        const uint32_t nc = w.c + w.step;
        if (nc < w.end) ch = load_const(a.tchunks + nc);
        return cur;
    }
};

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

// LDS replicas of the P2G windows, used for a "crowded" chunk (a tile cell
// with more than kLevelCap particles; such a cell's particles past the level
// order's levels fill whole waves of one cell).  Lane L then adds into
// replica L mod count, so the lanes of a wave that all add to one node meet
// `count` addresses, 64/count lanes each, instead of one address 64 times
// over (same-address LDS atomics of one instruction serialise); the flush
// sums the replicas.  Other chunks use replica 0 only and clear and flush
// nothing else.  Each replica is offset by 8 more banks than a whole number
// of windows.  (8^3 tiles: one replica, the window is already 16 KB.)
template <int TB>
struct WinRep {
    static constexpr int count = TB == 2 ? 4 : 1;
    static constexpr int WC = TileGeo<TB>::win_cells;
    static constexpr int stride1 = 4 * WC + 8;  // p2g_1: [mass | vx | vy | vz] per replica
    static constexpr int stride2 = 3 * WC + 8;  // p2g_2: [vx | vy | vz] per replica
    __device__ __forceinline__ static int of_lane(bool crowded) {
        return crowded ? (int)(threadIdx.x & (uint32_t)(count - 1)) : 0;
    }
};

// p2g_1 marks a crowded chunk in bit 31 of the first entry of its level order
// for p2g_2 (state slots are below 2^31).
constexpr uint32_t kCrowdedBit = 0x80000000u;

template <int TB>
struct WinSink1 {  // p2g_1 into the LDS window: [mass | vx | vy | vz] x win_cells
    static constexpr int W = TileGeo<TB>::win_edge, WC = TileGeo<TB>::win_cells;
    int *lds;
    int l0;
    __device__ __forceinline__ void add4(const Node3 &, const AxisNode &, const AxisNode &, int gx,
                                         int gy, int gz, int32_t m, int32_t x, int32_t y,
                                         int32_t z) const {
        const int l = l0 + (gx * W + gy) * W + gz;
        atomicAdd(&lds[l], m);
        atomicAdd(&lds[WC + l], x);
        atomicAdd(&lds[2 * WC + l], y);
        atomicAdd(&lds[3 * WC + l], z);
    }
};

struct GridSink1 {  // p2g_1 straight to the grid (reference-shaped atomics)
    int32_t *grid;
    uint32_t G;
    int32_t Y, Z;
    __device__ __forceinline__ void add4(const Node3 &n, const AxisNode &ax, const AxisNode &ay,
                                         int, int, int gz, int32_t m, int32_t x, int32_t y,
                                         int32_t z) const {
        const int32_t idx = node_index(n, ax, ay, gz, Y, Z);
        if (!in_grid(idx, G)) return;
        int32_t *cell = grid + 4 * (size_t)idx;
        atomicAdd(cell + 3, m);
        atomicAdd(cell + 0, x);
        atomicAdd(cell + 1, y);
        atomicAdd(cell + 2, z);
    }
};

template <int TB>
struct WinSink2 {  // p2g_2: decoded masses dm[WC]; momentum ints [vx | vy | vz] x WC
    static constexpr int W = TileGeo<TB>::win_edge, WC = TileGeo<TB>::win_cells;
    const float *dm;
    int *mom;
    int l0;
    __device__ __forceinline__ float mass(const Node3 &, const AxisNode &, const AxisNode &, int gx,
                                          int gy, int gz) const {
        return dm[l0 + (gx * W + gy) * W + gz];
    }
    __device__ __forceinline__ void add3(const Node3 &, const AxisNode &, const AxisNode &, int gx,
                                         int gy, int gz, int32_t x, int32_t y, int32_t z) const {
        const int l = l0 + (gx * W + gy) * W + gz;
        atomicAdd(&mom[l], x);
        atomicAdd(&mom[WC + l], y);
        atomicAdd(&mom[2 * WC + l], z);
    }
};

struct GridSink2 {  // masses from the completed grid, momentum to the wild accumulator
    const int32_t *grid;
    int32_t *acc;
    uint32_t G;
    int32_t Y, Z;
    float mult;
    __device__ __forceinline__ float mass(const Node3 &n, const AxisNode &ax, const AxisNode &ay,
                                          int, int, int gz) const {
        const int32_t idx = node_index(n, ax, ay, gz, Y, Z);
        return decode(in_grid(idx, G) ? grid[4 * (size_t)idx + 3] : 0, mult);
    }
    __device__ __forceinline__ void add3(const Node3 &n, const AxisNode &ax, const AxisNode &ay,
                                         int, int, int gz, int32_t x, int32_t y,
                                         int32_t z) const {
        const int32_t idx = node_index(n, ax, ay, gz, Y, Z);
        if (!in_grid(idx, G)) return;
        int32_t *cell = acc + 4 * (size_t)idx;
        atomicAdd(cell + 0, x);
        atomicAdd(cell + 1, y);
        atomicAdd(cell + 2, z);
    }
};

template <int TB>
struct WinSinkG {  // g2p: updated, decoded velocities staged in LDS
    static constexpr int W = TileGeo<TB>::win_edge;
    const float4 *lvel;
    int l0;
    __device__ __forceinline__ void vel(const Node3 &, const AxisNode &, const AxisNode &, int gx,
                                        int gy, int gz, float cv[3]) const {
        const float4 o = lvel[l0 + (gx * W + gy) * W + gz];
        cv[0] = o.x;
        cv[1] = o.y;
        cv[2] = o.z;
    }
};

struct GridSinkG {  // g2p straight from the decoded velocities
    const float4 *cells;
    uint32_t G;
    int32_t Y, Z;
    __device__ __forceinline__ void vel(const Node3 &n, const AxisNode &ax, const AxisNode &ay, int,
                                        int, int gz, float cv[3]) const {
        const int32_t idx = node_index(n, ax, ay, gz, Y, Z);
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in_grid(idx, G)) o = cells[idx];
        cv[0] = o.x;
        cv[1] = o.y;
        cv[2] = o.z;
    }
};

// ---------------------------------------------------------------------------
// binning: one key per (scene, tile) in Hilbert order plus one wild key per
// scene; counting sort; tile-chunk list.
//
// The counts come from the previous substep's g2p: a g2p block counts, in
// LDS, how many of its chunk's particles moved to each of the 27 tiles around
// the chunk's tile ("mig", 27 per chunk) and gives each particle its rank
// within its (chunk, direction).  k_bin_gather turns the counts into offsets
// within the destination tiles: it visits every key and sums, in a fixed
// order, the counts of the chunks that can move into it (no atomics).
// Particles that jump further than one tile, or become wild, take a rank
// from a global atomic on their key instead (rank "final").  After the state
// was replaced from outside, k_bin_count ranks every particle that way.

constexpr uint32_t kRankFinal = 0xffffffffu;

template <int TB>
__device__ __forceinline__ uint32_t bin_key(const BatchArgs &a, const SceneDesc *sd,
                                            const float x[3], const int from[3], uint32_t *dir) {
    return bin_key_of<TB>(a.tile_rank, sd, x, from, dir);
}

// Counts one particle per active lane into count[key] and returns the lane's
// rank among its key's particles.  Every lane of the wave must call this.
// Runs of consecutive active lanes with equal keys (the common case: storage
// is already in bin order) share one atomic.
__device__ __forceinline__ uint32_t count_rank(uint32_t *count, uint32_t key, bool active) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t prev_key = __shfl_up(key, 1, kWave);
    const int prev_active = __shfl_up(active ? 1 : 0, 1, kWave);
    const bool head = lane == 0 || !active || !prev_active || prev_key != key;
    const uint64_t heads = __ballot(head);
    const uint64_t le = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);  // lanes <= this
    const uint64_t above = heads & ~le;
    const uint32_t end = above ? (uint32_t)__ffsll((unsigned long long)above) - 1 : (uint32_t)kWave;
    uint32_t base = 0;
    if (head && active) base = atomicAdd(&count[key], end - (uint32_t)lane);
    const int my_head = 63 - __clzll((unsigned long long)(heads & le));  // lane 0 is a head
    return __shfl(base, my_head, kWave) + (uint32_t)(lane - my_head);
}


__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = 0;
}

template <int TB>
__global__ __launch_bounds__(kBlock) void k_bin_count(BatchArgs a) {
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        const Chunk ch = a.chunks[c];
        const SceneDesc *sd = a.scenes + ch.scene;
        const uint32_t j = threadIdx.x;  // chunks hold <= kBlock slots
        const bool active = j < ch.count;
        uint32_t key = 0;
        if (active) {
            float x[3];
            const int nowhere[3] = {-4, -4, -4};
            uint32_t dir;
            load_x(a.state, a.stride, ch.begin + j, x);
            key = bin_key<TB>(a, sd, x, nowhere, &dir);
        }
        const uint32_t r = count_rank(a.count, key, active);
        if (active) {
            a.key[ch.begin + j] = key;
            a.rank[ch.begin + j] = r;
            a.src[ch.begin + j] = kRankFinal;
        }
    }
}

// Exclusive scan over a block of NT threads; *total receives the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t s_w[NT / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) {
        const uint32_t s = s_w[i];
        if (i < w) wbase += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - v;
}

// The scan's prefix comes from totals of (particles, tile chunks) per
// super-segment of 256 keys, double-buffered by substep parity: a scan
// clears the other parity's, which the next k_bin_gather fills with atomics,
// plus the counts of the super-segment's keys before the block's.
constexpr uint32_t kSegKeys = 1u << kSegBits;  // keys per k_bin_gather block
constexpr uint32_t kSegPerSup = 1u << (kSupBits - kSegBits);

// One block of 4 waves per 64 keys.  Wave 0 computes the key offsets, the
// keys' tile-chunk ranges and the tile chunks themselves; meanwhile waves
// 1..3 check, for each empty tile key, whether a particle is within reach
// (a third of its 27 neighbours' counts each).  After one barrier wave 0
// writes the reach flags and clears the next pass's totals.
constexpr int kScanWaves = 4;
constexpr int kReachWaves = kScanWaves - 1;
constexpr int kReachPer = (kDirs + kReachWaves - 1) / kReachWaves;

__global__ __launch_bounds__(kScanWaves * kWave) void k_scan_keys(BatchArgs a, BinBuffers b,
                                                                  uint32_t cap) {
    __shared__ uint32_t s_any[kReachWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1), part = threadIdx.x / kWave;
    const uint32_t k0 = blockIdx.x * kWave, k = k0 + (uint32_t)lane;
    const bool in = k < b.nkeys;
    const uint32_t c = in ? b.count[k] : 0u;
    const uint32_t xm = in ? a.tile_xmaj[k] : 0u;
    if (part != 0) {
        // Every lane loads its key's neighbour keys and their counts, whether
        // or not it needs them: two rounds of loads with no branch around
        // them (the wild key and lanes past the last key load the entries of
        // a valid tile; missing neighbours read key 0).
        const uint32_t kk = min(k, b.nkeys - 1);
        const SceneDesc sd = load_const(a.scenes + (a.nscenes == 1 ? 0u : a.key_scene[kk]));
        const uint32_t t = sd.soff + min(kk - sd.koff, sd.ntiles - 1);  // spatial tile
        const uint32_t *nb = a.nbr + (size_t)t * kDirs;
        uint32_t nk[kReachPer], cn[kReachPer];
#pragma unroll
        for (int j = 0; j < kReachPer; ++j)
            nk[j] = nb[min((uint32_t)((part - 1) + kReachWaves * j), kDirs - 1)];
#pragma unroll
        for (int j = 0; j < kReachPer; ++j) cn[j] = b.count[nk[j] != kNoTile ? nk[j] : 0u];
        bool any = false;
        if (in && xm != kWildTile && c == 0 && sd.fast) {
#pragma unroll
            for (int j = 0; j < kReachPer; ++j) {
                const uint32_t d = (uint32_t)((part - 1) + kReachWaves * j);
                any = any || (d < kDirs && nk[j] != kNoTile && cn[j] != 0);
            }
        }
        s_any[part - 1][lane] = any ? 1u : 0u;
        __syncthreads();
        return;
    }
    const uint32_t sp = k0 >> kSupBits;  // super-segment of the block
    uint32_t ps = 0, pc = 0;  // totals of all earlier keys
    for (uint32_t i0 = 0; i0 < sp; i0 += kWave) {
        const uint32_t i = i0 + (uint32_t)lane;
        if (i < sp) {
            const uint2 v = b.sup[i];
            ps += v.x;
            pc += v.y;
        }
    }
    for (uint32_t i0 = sp << kSupBits; i0 < k0; i0 += kWave) {  // <= 3 rounds
        const uint32_t i = i0 + (uint32_t)lane;
        if (i < k0) {
            const uint32_t v = b.count[i];
            ps += v;
            pc += chunks_of(v, cap);
        }
    }
    // a scene with wild particles (or a non-integral box) needs every tile visited
    bool wild = false;
    if (in && xm != kWildTile) {
        const SceneDesc *sd = a.scenes + key_scene_of(a, k);
        wild = !sd->fast || b.count[sd->koff + sd->ntiles] != 0;
    }
    ps = wave_sum(ps);
    pc = wave_sum(pc);
    const uint32_t nch = chunks_of(c, cap);
    const uint32_t off = ps + wave_incl_scan(c, lane) - c;
    const uint32_t cb = pc + wave_incl_scan(nch, lane) - nch;
    if (in) {
        b.offset[k] = off;
        b.tile_chunk[k] = make_uint2(cb, nch);
        if (nch) {
            const uint32_t sc = key_scene_of(a, k);
            for (uint32_t i = 0; i < nch; ++i) {
                const uint32_t z = off + i * cap, zn = min(cap, c - i * cap);
                // (the slot bound only guards the buffers against inconsistent totals)
                if (cb + i < b.tchunk_cap && z + zn <= a.total_slots)
                    b.tchunks[cb + i] = make_uint4(sc, xm, z, zn);
                // XCD ranges: slot s_x = ceil(x * total / kXcds) starts range x;
                // the chunk holding slot s_x - 1 is the last of range x - 1
                // (chunks tile the slots [0, total) in order).
#pragma unroll
                for (int x = 1; x < kXcds; ++x) {
                    const uint32_t sx = xcd_split_slot((uint32_t)x, a.total_slots);
                    if (z < sx && sx <= z + zn) b.xcd_first[x] = cb + i + 1;
                }
            }
        }
        if (k == b.nkeys - 1) {
            b.ntchunks[0] = cb + nch;
            b.ntchunks[1] = nch;  // the last key is the last scene's wild key
            b.xcd_first[0] = 0;
            b.xcd_first[kXcds] = cb + nch;
        }
    }
    // clear the next binning's totals (its k_bin_gather adds to them)
    if (lane == 0 && (k0 & ((1u << kSupBits) - 1)) == 0) b.sup_next[sp] = make_uint2(0u, 0u);
    __syncthreads();
    // Every tile key writes its reach flag.  Counts stay readable until
    // reduce_2 clears them.
    if (in && xm != kWildTile) {
        const SceneDesc *sd = a.scenes + key_scene_of(a, k);
        bool any = c != 0 || wild;
#pragma unroll
        for (int w = 0; w < kReachWaves; ++w) any = any || s_any[w][lane] != 0;
        a.near[sd->soff + (k - sd->koff)] = any ? 1 : 0;
    }
}

// Writes the permutation of the binning: perm[d] = slot of the particle that
// takes bin-order slot d.  Keys of scene s lie in [koff_s, koff_s + nkeys_s)
// and scenes are packed in slot order, so d stays inside the scene's slot
// range (the check only guards the buffers against inconsistent ranks).
__global__ __launch_bounds__(kBlock) void k_bin_scatter(BatchArgs a, BinBuffers b) {
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        // a batch of one scene: static chunk c is slots [256 c, 256 c + 256) of
        // its n = total_slots (Batch::upload_chunks), known without a load
        const Chunk ch =
            a.nscenes == 1
                ? Chunk{0u, c * kBlock, min((uint32_t)kBlock, a.total_slots - c * kBlock), 0u}
                : load_const(a.chunks + c);
        const SceneDesc *sd = a.scenes + ch.scene;
        const uint32_t lo = sd->poff, hi = sd->poff + sd->n;
        for (uint32_t j = threadIdx.x; j < ch.count; j += kBlock) {
            const uint32_t i = ch.begin + j;
            const uint32_t src = a.src[i];
            uint32_t r = a.rank[i];
            if (src != kRankFinal) r += a.mig[src];
            const uint32_t d = b.offset[a.key[i]] + r;
            if (d < lo || d >= hi) continue;
            b.perm_out[d] = i;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_gather_binned(BatchArgs a, float *dst, uint32_t *dst_id) {
    for (uint32_t d = blockIdx.x * kBlock + threadIdx.x; d < a.total_slots; d += gridDim.x * kBlock) {
        const uint32_t i = a.perm[d];
#pragma unroll
        for (int k = 0; k < NCOMP; ++k) dst[k * a.stride + d] = a.state[k * a.stride + i];
        dst_id[d] = a.id[i];
    }
}

// Tile key k (rank `rank` in scene sd), whose count so far is `total` (the
// particles that reached it through global atomics): the chunks of the 27
// tiles around it (previous binning) that sent particles into it, in a fixed
// order; their counts in mig become running offsets within the tile.  Each
// (chunk, direction) entry belongs to exactly one destination tile.  Every
// lane of the wave calls this with the same key; returns the key's total,
// which lane 0 also stores.
// s: lane q < 27 holds the key of neighbour q of key k's tile (a.nbr), the
// source tile T + d of direction q; other lanes hold kNoTile.
__device__ __forceinline__ uint32_t gather_key(const BatchArgs &a, const BinBuffers &bb,
                                               uint32_t k, uint32_t s, uint32_t total) {
    const int lane = threadIdx.x & (kWave - 1);
    // (no branch around the load: a missing source reads key 0 and counts nothing)
    const uint2 tc = a.tile_chunk[s != kNoTile ? s : 0u];
    const uint32_t first = s != kNoTile ? tc.x : 0u;
    const uint32_t cnt = s != kNoTile && tc.x <= a.tchunk_cap && tc.y <= a.tchunk_cap - tc.x
                             ? tc.y : 0u;  // (guard)
    const uint32_t dir = lane < (int)kDirs ? kDirs - 1 - lane : 0u;  // from the source, T lies at -d
    const uint32_t rounds = wave_max(cnt);
    for (uint32_t j = 0; j < rounds; ++j) {
        const bool has = j < cnt;
        const size_t e = (size_t)(first + j) * kDirs + dir;
        const uint32_t v = has ? a.mig[e] : 0u;
        const uint32_t inc = wave_incl_scan(v, lane);
        if (has) a.mig[e] = total + inc - v;
        total += __shfl(inc, kWave - 1, kWave);
    }
    if (lane == 0) bb.count[k] = total;
    return total;
}

// Adds a block's key totals to their super-segment's.
__device__ __forceinline__ void add_segment_totals(const BinBuffers &bb, const uint32_t *s_cnt,
                                                   const uint32_t *s_nch, uint32_t sup) {
    uint32_t ts = 0, tc = 0;
#pragma unroll
    for (int i = 0; i < (int)kSegKeys; ++i) {
        ts += s_cnt[i];
        tc += s_nch[i];
    }
    if (ts) {
        atomicAdd(&bb.sup[sup].x, ts);
        atomicAdd(&bb.sup[sup].y, tc);
    }
}

// Dense key spaces: one block of 16 waves per segment of 16 keys, one wave
// per key (gather_key when a particle came near the tile last substep).
// With `final_counts` (after k_bin_count) or for wild keys the count is
// already complete.  The block then adds its keys' totals to their
// super-segment's.
__global__ __launch_bounds__(kSegKeys * kWave) void k_bin_gather(BatchArgs a, BinBuffers bb,
                                                                 uint32_t cap, int final_counts) {
    __shared__ uint32_t s_cnt[kSegKeys], s_nch[kSegKeys];
    const uint32_t nseg = (bb.nkeys + kSegKeys - 1) / kSegKeys;
    uint32_t sg = blockIdx.x;
    if (a.xcd_remap) {  // contiguous segments per XCD: neighbouring tiles share an L2
        const uint32_t per = gridDim.x / kXcds;
        sg = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    }
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave),
              lane = threadIdx.x & (kWave - 1);
    const uint32_t k = sg * kSegKeys + (uint32_t)w;
    uint32_t total = 0;
    if (sg < nseg && k < bb.nkeys) {
        // the key's count, its tile's reach flag and its 27 source keys are
        // loaded in one round (a batch of one scene needs no scene lookup
        // first; lanes past 26 and the wild key load a valid neighbour entry)
        const SceneDesc sd = load_const(a.scenes + (a.nscenes == 1 ? 0u : a.key_scene[k]));
        const uint32_t rank = k - sd.koff;
        const uint32_t b = sd.soff + min(rank, sd.ntiles - 1);  // spatial tile of key k
        const uint32_t cnt_raw = bb.count[k];
        const uint32_t near_raw = a.near_prev[b];
        const uint32_t sl = a.nbr[(size_t)b * kDirs + min((uint32_t)lane, kDirs - 1)];
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::"v"(cnt_raw), "v"(near_raw), "v"(sl));
        total = cnt_raw;
        // a tile no particle came near last substep has no source chunk
        const bool sources = rank < sd.ntiles && near_raw != 0;
        if (!final_counts && sd.fast && rank < sd.ntiles && sources)
            total = gather_key(a, bb, k, lane < (int)kDirs ? sl : kNoTile, total);
    }
    if (lane == 0) {
        s_cnt[w] = total;
        s_nch[w] = chunks_of(total, cap);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sg < nseg) add_segment_totals(bb, s_cnt, s_nch, sg / kSegPerSup);
}

// Sparse key spaces (keys far outnumber particle chunks, as in many small
// scenes): most keys have no source chunk, and one wave per key would spend
// the launch on waves that do nothing.  Here a block of 16 waves takes one
// super-segment (256 keys), 16 keys per wave: the wave's lanes first read
// their keys' counts and near flags in parallel, then the wave gathers only
// the keys that need it, one after another.  Same results as k_bin_gather.
constexpr uint32_t kSparseKeysPerWave = 1u << (kSupBits - kSegBits);

__global__ __launch_bounds__(kSegKeys * kWave) void k_bin_gather_sparse(BatchArgs a, BinBuffers bb,
                                                                        uint32_t cap,
                                                                        int final_counts) {
    __shared__ uint32_t s_cnt[kSegKeys], s_nch[kSegKeys];
    const uint32_t nsup = (bb.nkeys + (1u << kSupBits) - 1) >> kSupBits;
    uint32_t sp = blockIdx.x;  // super-segment of this block
    if (a.xcd_remap) {
        const uint32_t per = gridDim.x / kXcds;
        sp = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    }
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const uint32_t k = (sp * kSegKeys + (uint32_t)w) * kSparseKeysPerWave + (uint32_t)lane;
    const bool mine = sp < nsup && lane < (int)kSparseKeysPerWave && k < bb.nkeys;
    uint32_t total = 0, rank = 0, sc = 0;
    bool todo = false;
    if (mine) {
        sc = key_scene_of(a, k);
        const SceneDesc *sd = a.scenes + sc;
        rank = k - sd->koff;
        total = bb.count[k];
        todo = !final_counts && sd->fast && rank < sd->ntiles && a.near_prev[sd->soff + rank] != 0;
    }
    for (uint64_t pending = __ballot(todo); pending; pending &= pending - 1) {  // wave-uniform
        const int i = __ffsll((unsigned long long)pending) - 1;
        const uint32_t kk = __shfl(k, i, kWave), rr = __shfl(rank, i, kWave);
        const uint32_t ss = __shfl(sc, i, kWave), t0 = __shfl(total, i, kWave);
        const SceneDesc *sd = a.scenes + ss;
        const uint32_t b = sd->soff + rr;  // spatial tile of key kk
        const uint32_t sk = lane < (int)kDirs ? a.nbr[(size_t)b * kDirs + lane] : kNoTile;
        const uint32_t t = gather_key(a, bb, kk, sk, t0);
        if (lane == i) total = t;
    }
    const uint32_t wt = wave_sum(mine ? total : 0u);
    const uint32_t wc = wave_sum(mine ? chunks_of(total, cap) : 0u);
    if (lane == 0) {
        s_cnt[w] = wt;
        s_nch[w] = wc;
    }
    __syncthreads();
    if (threadIdx.x == 0 && sp < nsup) add_segment_totals(bb, s_cnt, s_nch, sp);
}

// ---------------------------------------------------------------------------
// Level order of a window chunk for the P2G adds.  Level k holds the k-th
// particle of every cell of the tile that has more than k (cells in index
// order), so the 64 lanes of a wave, which take 64 consecutive positions,
// mostly hold 64 different cells: their ds_add_u32 to one stencil node then
// rarely meet on one LDS address (same-address atomics of one instruction
// serialise).  The position of the k-th particle of cell c is
// base[k] + #{cells c' < c holding more than k} for the first kLevelCap levels
// (a crowded cell's particles past them follow cell-major: level_position;
// the window replicas bound the cost of such a cell's adds).  Integer adds
// commute, so the order changes no result.  With 8^3 tiles (512 cells, too many for the
// 64-bit level masks) the chunk's particles keep their bin order: a wave's
// 64 particles then spread over 8 times as many cells.

template <int PPT, int TB>
struct LevelOrder {
    uint32_t cnt[TileGeo<TB>::cells];    // particles per cell of the tile
    uint32_t base[kLevelCap];            // first position of each level
    uint64_t mask[kLevelCap];            // cells present in each level
    uint32_t exoff[TileGeo<TB>::cells];  // first position of each cell's particles past
                                         // the last level (level_position)
    uint32_t crowded;                    // a cell holds more than kLevelCap particles
    uint16_t perm[kBlock * PPT];   // position -> slot within the chunk
    uint32_t slot[kBlock * PPT];   // slot within the chunk -> state slot (a.perm)
    float x[3][kBlock * PPT];      // position of each slot within the chunk (read once)
};

// Fills o.perm and o.crowded for window chunk ch and clears the LDS window
// win: replica 0 (nwin ints), and for a crowded chunk the other replicas
// (nextra ints after it); the barriers here publish it.  Every thread of the
// block calls it.  The state slots of all the thread's particles are loaded
// first, then all their positions: two rounds of loads in flight, not two per
// particle.
template <int PPT, int TB>
__device__ __forceinline__ void level_order(const BatchArgs &a, const TileChunk &ch,
                                            LevelOrder<PPT, TB> &o, int *win, int nwin,
                                            int nextra) {
    // (slots past the chunk's end load its last particle: no branch around
    // the loads; nothing is stored for them)
    uint32_t src[PPT];
    float xr[PPT][3];
#pragma unroll
    for (int k = 0; k < PPT; ++k)
        src[k] = a.perm[ch.z + min(threadIdx.x + (uint32_t)k * kBlock, ch.w - 1)];
#pragma unroll
    for (int k = 0; k < PPT; ++k) load_x(a.state, a.stride, src[k], xr[k]);
    for (int l = threadIdx.x; l < nwin; l += kBlock) win[l] = 0;
    if constexpr (TB == 3) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = threadIdx.x + (uint32_t)k * kBlock;
            if (j < ch.w) {
                o.slot[j] = src[k];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) o.x[ax][j] = xr[k][ax];
                o.perm[j] = (uint16_t)j;
            }
        }
        if (threadIdx.x == 0) o.crowded = 0;
        __syncthreads();
    } else {
    const int lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x < kTileCells) o.cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t cell[PPT], lev[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const uint32_t j = threadIdx.x + (uint32_t)k * kBlock;
        cell[k] = lev[k] = 0;
        if (j < ch.w) {
            o.slot[j] = src[k];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) o.x[ax][j] = xr[k][ax];
            const int m = kTileEdge - 1;
            const int cx = (int)floorf(xr[k][0]) & m;
            const int cy = (int)floorf(xr[k][1]) & m;
            const int cz = (int)floorf(xr[k][2]) & m;
            cell[k] = (uint32_t)((cx * kTileEdge + cy) * kTileEdge + cz);
            lev[k] = atomicAdd(&o.cnt[cell[k]], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kWave) {  // one wave: lane = cell
        const uint32_t c = o.cnt[lane];
        const uint32_t cmax = wave_max(c), levels = min(cmax, kLevelCap);
        uint32_t run = 0;  // (wave-uniform)
        for (uint32_t k = 0; k < levels; ++k) {
            const uint64_t m = __ballot(c > k);
            if (lane == 0) {
                o.mask[k] = m;
                o.base[k] = run;
            }
            run += (uint32_t)__popcll(m);
        }
        if (cmax > kLevelCap) {  // (wave-uniform) cell-major past the levels
            const uint32_t ex = c > kLevelCap ? c - kLevelCap : 0u;
            o.exoff[lane] = run + wave_incl_scan(ex, lane) - ex;
        }
        if (lane == 0) o.crowded = cmax > kLevelCap ? 1u : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const uint32_t j = threadIdx.x + (uint32_t)k * kBlock;
        if (j < ch.w) o.perm[level_position(lev[k], cell[k], o.base, o.mask, o.exoff)] = (uint16_t)j;
    }
    if (o.crowded)  // (block-uniform)
        for (int l = threadIdx.x; l < nextra; l += kBlock) win[nwin + l] = 0;
    __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// p2g_1: mass and APIC momentum (mls-mpm/p2g_1.wgsl:24-65) into tile partials

// Register budget of the chunk kernels at the default 2 particles per thread:
// 80 VGPRs (6 waves per SIMD).  At this budget k_g2p spills only in its
// wild-chunk loop; measured 6% (p2g_1) and 8% (g2p) faster than the
// compiler's own ~92-VGPR choice, and p2g_2 (85 VGPRs on its own, i.e. 5
// waves per SIMD) about 5% faster.
template <int PPT>
constexpr int kChunkWaves = PPT == 2 ? 6 : 1;  // amdgpu_waves_per_eu minimum (1: no constraint)

// The walk of a p2g_1 block.  kSingle: a batch of one scene, whose
// descriptor is then known before the first chunk's.
template <int PPT, int TB, bool kSingle>
__device__ __forceinline__ void p2g1_walk(const BatchArgs &a, int *lds, LevelOrder<PPT, TB> &ord) {
    constexpr int WC = TileGeo<TB>::win_cells;
    bool wild_seen = false;  // block-uniform: a wild chunk among this block's
    SceneDesc scene;
    for (ChunkWalk cw(a, scene, std::integral_constant<bool, kSingle>{}); cw.w.c < cw.w.end;
         cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(a);
        if (ch.y == kWildTile) {  // second loop
            wild_seen = true;
            continue;
        }
        if constexpr (!kSingle) scene = load_const(a.scenes + ch.x);
        const float mult = scene.prm.mult;
        level_order<PPT, TB>(a, ch, ord, lds, WinRep<TB>::stride1,
                             (WinRep<TB>::count - 1) * WinRep<TB>::stride1);
        const bool crowded = ord.crowded != 0;
        // positions past the chunk's end load its last particle (no branch
        // around the loads; those lanes store and add nothing)
        float x[PPT][3], v[PPT][3], C[PPT][3][3];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t pos = threadIdx.x + (uint32_t)k * kBlock;
            valid[k] = pos < ch.w;
            const uint32_t j = ord.perm[min(pos, ch.w - 1)];
            const uint32_t i = ord.slot[j];
            if (valid[k])  // p2g_2 takes the same order (and the crowded mark)
                a.order[ch.z + pos] = i | (pos == 0 && crowded ? kCrowdedBit : 0u);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) x[k][ax] = ord.x[ax][j];
            load_v(a.state, a.stride, i, v[k]);
            load_C(a.state, a.stride, i, C[k]);
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            Node3 n;
            node3_init(x[k], n);
            p2g1_particle(n, v[k], C[k], mult,
                          WinSink1<TB>{lds + WinRep<TB>::of_lane(crowded) * WinRep<TB>::stride1,
                                       win_l0<TB>(x[k])});
        }
        __syncthreads();
        int4 *out = a.partial + (size_t)cw.w.c * WC;
        auto flush1 = [&](auto nrep) {
            for (int l = threadIdx.x; l < WC; l += kBlock) {
                int4 s = make_int4(0, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < decltype(nrep)::value; ++r) {
                    const int *w = lds + r * WinRep<TB>::stride1;
                    s = add4(s, make_int4(w[WC + l], w[2 * WC + l], w[3 * WC + l], w[l]));
                }
                out[l] = s;
            }
        };
        if (crowded)  // (block-uniform)
            flush1(std::integral_constant<int, WinRep<TB>::count>{});
        else
            flush1(std::integral_constant<int, 1>{});
        __syncthreads();
    }
    // Wild chunks (rare): global atomics into the wild accumulator.  A loop of
    // their own keeps their registers out of the window loop's budget.
    if (!wild_seen) return;
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const SceneDesc *sd = a.scenes + ch.x;
        const GridSink1 sink{(int32_t *)(a.wild + sd->goff), sd->G, sd->uni.Y, sd->uni.Z};
        for (uint32_t j = threadIdx.x; j < ch.w; j += kBlock) {
            float x[3], v[3], C[3][3];
            const uint32_t i = a.perm[ch.z + j];
            load_x(a.state, a.stride, i, x);
            load_v(a.state, a.stride, i, v);
            load_C(a.state, a.stride, i, C);
            Node3 n;
            node3_init(x, n);
            p2g1_particle<true>(n, v, C, sd->prm.mult, sink);
        }
    }
}

template <int PPT, int TB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kChunkWaves<PPT>)))
void k_p2g1(BatchArgs a) {
    __shared__ int lds[WinRep<TB>::count * WinRep<TB>::stride1];  // [mass | vx | vy | vz] x replica
    __shared__ LevelOrder<PPT, TB> ord;
    if (a.nscenes == 1)
        p2g1_walk<PPT, TB, true>(a, lds, ord);
    else
        p2g1_walk<PPT, TB, false>(a, lds, ord);
}

// ---------------------------------------------------------------------------
// reduce: one wave per spatial tile (tiles in Hilbert order).  Lanes 0..26
// fetch the chunk ranges of the 27 neighbour tiles; then each lane owns one
// of the tile's 64 cells, which lies in the window of its own tile plus, per
// axis, the neighbour below (local coordinate 0) or above (3): at most 8
// windows.  The partial loads of all those windows are issued together, one
// round per chunk of the fullest neighbour.

// The partial windows covering one lane's cell: up to 8 (neighbour, window
// index) slots and the chunk count of each.
struct CellPlan {
    uint32_t base[8], cnt[8];
    uint32_t rounds;  // chunks of the fullest neighbour (wave-uniform)
};

// kt: lane q < 27 holds the key of neighbour q of the tile (a.nbr).
template <int TB>
__device__ __forceinline__ CellPlan plan_partials(const BatchArgs &a, uint32_t kt, const int l[3],
                                                  bool cell_ok) {
    // (no branch around the load: a missing neighbour, or a lane past 26,
    // whose kt is kNoTile, reads key 0 and takes no chunks)
    const uint2 tc = a.tile_chunk[kt != kNoTile ? kt : 0u];
    const bool has = kt != kNoTile;
    const uint32_t nfirst = has ? tc.x : 0u;
    const uint32_t ncnt =
        has && tc.x <= a.tchunk_cap && tc.y <= a.tchunk_cap - tc.x ? tc.y : 0u;  // guard
    CellPlan p;
    uint32_t mc = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        int q;
        bool valid;
        const int widx = cover_slot<TB>(l, s, &q, &valid);
        valid = valid && cell_ok;
        const uint32_t f = __shfl(nfirst, q, kWave), c = __shfl(ncnt, q, kWave);
        p.base[s] = f * (uint32_t)TileGeo<TB>::win_cells + (uint32_t)widx;
        p.cnt[s] = valid ? c : 0u;
        mc = max(mc, p.cnt[s]);
    }
    p.rounds = wave_max(mc);
    return p;
}

// Sum of the planned partials: all slots' loads of one round in flight at once.
template <int TB>
__device__ __forceinline__ int4 sum_partials(const BatchArgs &a, const CellPlan &p) {
    int4 acc = make_int4(0, 0, 0, 0);
    for (uint32_t j = 0; j < p.rounds; ++j) {
        int4 v[8];
#pragma unroll
        for (int s = 0; s < 8; ++s)
            v[s] = j < p.cnt[s] ? a.partial[(size_t)p.base[s] + (size_t)j * TileGeo<TB>::win_cells]
                                : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = add4(acc, v[s]);
    }
    return acc;
}

// Folds the wild accumulator of cell gi into v and clears it.
__device__ __forceinline__ void take_wild(const BatchArgs &a, uint64_t gi, int4 &v, bool mass) {
    const int4 w = a.wild[gi];
    v = add4(v, make_int4(w.x, w.y, w.z, mass ? w.w : 0));
    a.wild[gi] = make_int4(0, 0, 0, 0);
}

// The reduce of one tile.  kSingle: a batch of one scene, whose descriptor is
// then known before the tile's.
template <int PHASE, int TB, bool kSingle>
__device__ __forceinline__ void reduce_run(const BatchArgs &a) {
    constexpr int E = TileGeo<TB>::edge;
    constexpr uint32_t kCells = (uint32_t)TileGeo<TB>::cells;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t cell = TB == 2 ? lane : threadIdx.x;  // cell of the tile this lane owns
    // the kernel arguments the first loads need, loaded as one batch
    asm volatile("" ::"s"(a.scenes), "s"(a.ntchunks), "s"(a.tile_info), "s"(a.dirty),
                 "s"(a.near), "s"(a.nbr), "s"(a.stiles), "s"(a.xcd_remap));
    const uint32_t b = reduce_tile<TB>(a);
    if (b >= a.stiles) return;  // (8^3: the whole block)
    // Everything the tile's first decisions need is loaded together: its
    // flags and coordinates, its neighbours' keys, and for a batch of one
    // scene the scene and whether it has wild particles.
    // (lanes past 26 load neighbour 26 again: no branch around the load)
    const uint32_t ktl = a.nbr[(size_t)b * 27 + min(lane, 26u)];
    const uint32_t kt = lane < 27 ? ktl : kNoTile;
    SceneDesc scene;
    uint2 head = make_uint2(0u, 0u);
    if constexpr (kSingle) {
        scene = load_const(a.scenes);
        head = load_const(reinterpret_cast<const uint2 *>(a.ntchunks));
    }
    const uint4 ti = load_const(a.tile_info + b);
    const uint32_t dirty_raw = a.dirty[b], near_raw = a.near[b];
    // (an empty asm that needs the values: their loads complete here
    // together; the scheduling barrier keeps every use below all the loads)
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (kSingle)
        asm volatile("" ::"s"(ti.x), "s"(ti.y), "s"(ti.z), "s"(head.y), "s"(scene.goff),
                     "s"(scene.uni.Y), "s"(scene.uni.Z), "s"(scene.kdim[0]), "s"(scene.kdim[1]),
                     "s"(scene.kdim[2]), "v"(ktl), "v"(dirty_raw), "v"(near_raw));
    else
        asm volatile("" ::"s"(ti.w), "v"(ktl), "v"(dirty_raw), "v"(near_raw));
    const bool dirty = dirty_raw != 0, near = near_raw != 0;
    // no particle near, nothing left to clear, no wild particle in the scene:
    // the tile's cells are empty and already written as such
    if (!near && !dirty) return;  // (its count is zero too)
    if constexpr (!kSingle) scene = load_const(a.scenes + ti.w);
    const SceneDesc *sd = &scene;
    if (PHASE == 2 && cell == 0) {  // ready for the next counting pass
        a.count[sd->koff + (b - sd->soff)] = 0;
        if (b == sd->soff) a.count[sd->koff + sd->ntiles] = 0;  // the scene's wild key
    }
    const bool has_wild =
        kSingle ? head.y != 0 : a.tile_chunk[sd->toff + sd->ntiles].y != 0;
    const Params prm = sd->prm;
    const int32_t wall_hi[3] = {sd->uni.wall_hi[0], sd->uni.wall_hi[1], sd->uni.wall_hi[2]};
    if (!sd->fast) {  // every particle is wild: cells in linear order
        const uint32_t li = (b - sd->soff) * kCells + cell;
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
    const int l[3] = {(int)(cell >> (2 * TB)), (int)(cell >> TB) & (E - 1), (int)cell & (E - 1)};
    const int cx = (int)ti.x * E + l[0], cy = (int)ti.y * E + l[1], cz = (int)ti.z * E + l[2];
    const bool cell_ok = in_dims(cx, cy, cz, sd);
    const uint64_t gi = (uint64_t)sd->goff + cell_linear(cx, cy, cz, sd->uni.Y, sd->uni.Z);
    const CellPlan plan = plan_partials<TB>(a, kt, l, cell_ok);
    // No particle near the tile: its cells are empty.  If they were already
    // empty after the last substep (not dirty), nothing needs writing.
    bool empty = plan.rounds == 0 && !has_wild;
    if constexpr (TB == 3) empty = !__syncthreads_or(empty ? 0 : 1);
    if (PHASE == 1) {
        if (empty && !dirty) return;
        int4 acc = sum_partials<TB>(a, plan);
        if (!cell_ok) return;
        if (has_wild) take_wild(a, gi, acc, true);
        a.grid[gi] = acc;
    } else {
        if (empty) {
            if (dirty && cell_ok) a.vel[gi] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (cell == 0 && dirty) a.dirty[b] = 0;
            return;
        }
        int4 g = make_int4(0, 0, 0, 0);
        if (cell_ok) g = a.grid[gi];
        const int4 acc = sum_partials<TB>(a, plan);
        if (cell == 0 && !dirty) a.dirty[b] = 1;
        if (!cell_ok) return;
        g = add4(g, make_int4(acc.x, acc.y, acc.z, 0));
        if (has_wild) take_wild(a, gi, g, false);
        a.grid[gi] = g;
        const int4 u = update_cell_xyz(g, cx, cy, cz, wall_hi, prm);
        a.vel[gi] = make_float4(decode(u.x, prm.mult), decode(u.y, prm.mult),
                                decode(u.z, prm.mult), 0.0f);
    }
}

// One lane per cell: a 4^3 tile is one wave (a block holds 4 tiles); an 8^3
// tile is a whole block of 8 waves, whose "anything to write" decision is
// taken for the tile as a whole (__syncthreads_or), so its waves agree on
// the tile's dirty flag.
template <int PHASE, int TB>
__global__ __launch_bounds__(ReduceGeo<TB>::threads) void k_reduce(BatchArgs a) {
    if (a.nscenes == 1)
        reduce_run<PHASE, TB, true>(a);
    else
        reduce_run<PHASE, TB, false>(a);
}

// ---------------------------------------------------------------------------
// p2g_2: density, pressure, viscous stress momentum (mls-mpm/p2g_2.wgsl:30-94)

template <int PPT, int TB, bool kSingle>
__device__ __forceinline__ void p2g2_walk(const BatchArgs &a, int *lds, uint32_t &s_crowded) {
    constexpr int WC = TileGeo<TB>::win_cells;
    using Rep = WinRep<TB>;
    // [decoded mass (f32)] then per replica [vx | vy | vz]
    float *dm = reinterpret_cast<float *>(lds);
    int *mom = lds + WC;
    bool wild_seen = false;  // block-uniform: a wild chunk among this block's
    SceneDesc scene;
    const SceneDesc *sd = &scene;
    for (ChunkWalk cw(a, scene, std::integral_constant<bool, kSingle>{}); cw.w.c < cw.w.end;
         cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(a);
        if (ch.y == kWildTile) {  // second loop
            wild_seen = true;
            continue;
        }
        if constexpr (!kSingle) scene = load_const(a.scenes + ch.x);
        const Params prm = sd->prm;
        // the level order p2g_1 recorded: all slots, then all particle loads
        // (positions past the chunk's end load its last particle)
        uint32_t oi[PPT];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t pos = threadIdx.x + (uint32_t)k * kBlock;
            valid[k] = pos < ch.w;
            oi[k] = a.order[ch.z + min(pos, ch.w - 1)];
        }
        if (threadIdx.x == 0) s_crowded = oi[0] >> 31;  // (position 0 carries the mark)
#pragma unroll
        for (int k = 0; k < PPT; ++k) oi[k] &= ~kCrowdedBit;
        float x[PPT][3], C[PPT][3][3];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            load_x(a.state, a.stride, oi[k], x[k]);
            load_C(a.state, a.stride, oi[k], C[k]);
        }
        int t[3];
        tile_coords(ch.y, sd, t);
        for (int l = threadIdx.x; l < WC; l += kBlock) {
            const int64_t gi = win_cell_index<TB>(l, t, sd);
            dm[l] = gi >= 0 ? decode(a.grid[gi].w, prm.mult) : 0.0f;
        }
        for (int l = threadIdx.x; l < Rep::stride2; l += kBlock) mom[l] = 0;
        __syncthreads();
        const bool crowded = s_crowded != 0;
        if (crowded) {  // (block-uniform) the other replicas too
            for (int l = threadIdx.x; l < (Rep::count - 1) * Rep::stride2; l += kBlock)
                mom[Rep::stride2 + l] = 0;
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            Node3 n;
            node3_init(x[k], n);
            p2g2_particle(n, C[k], prm,
                          WinSink2<TB>{dm, mom + Rep::of_lane(crowded) * Rep::stride2,
                                       win_l0<TB>(x[k])});
        }
        __syncthreads();
        int4 *out = a.partial + (size_t)cw.w.c * WC;
        auto flush2 = [&](auto nrep) {
            for (int l = threadIdx.x; l < WC; l += kBlock) {
                int4 s = make_int4(0, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < decltype(nrep)::value; ++r) {
                    const int *w = mom + r * Rep::stride2;
                    s = add4(s, make_int4(w[l], w[WC + l], w[2 * WC + l], 0));
                }
                out[l] = s;
            }
        };
        if (crowded)  // (block-uniform)
            flush2(std::integral_constant<int, Rep::count>{});
        else
            flush2(std::integral_constant<int, 1>{});
        __syncthreads();
    }
    // Wild chunks (rare), in a loop of their own like k_p2g1's.
    if (!wild_seen) return;
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const SceneDesc *sd = a.scenes + ch.x;
        const Params prm = sd->prm;
        const GridSink2 sink{(const int32_t *)(a.grid + sd->goff), (int32_t *)(a.wild + sd->goff),
                             sd->G, sd->uni.Y, sd->uni.Z, prm.mult};
        for (uint32_t j = threadIdx.x; j < ch.w; j += kBlock) {
            float x[3], C[3][3];
            const uint32_t i = a.perm[ch.z + j];
            load_x(a.state, a.stride, i, x);
            load_C(a.state, a.stride, i, C);
            Node3 n;
            node3_init(x, n);
            p2g2_particle<true>(n, C, prm, sink);
        }
    }
}

template <int PPT, int TB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kChunkWaves<PPT>)))
void k_p2g2(BatchArgs a) {
    __shared__ int lds[TileGeo<TB>::win_cells + WinRep<TB>::count * WinRep<TB>::stride2];
    __shared__ uint32_t s_crowded;  // the current chunk is crowded (p2g_1's mark)
    if (a.nscenes == 1)
        p2g2_walk<PPT, TB, true>(a, lds, s_crowded);
    else
        p2g2_walk<PPT, TB, false>(a, lds, s_crowded);
}

// ---------------------------------------------------------------------------
// g2p with copyPosition fused (mls-mpm/g2p.wgsl:26-92, copyPosition.wgsl:15-21)

// Writes particle o (its caller index, a.id of its input slot) updated, at
// bin-order slot i of the output state.
__device__ __forceinline__ void g2p_store(const BatchArgs &a, uint32_t i, uint32_t o,
                                          const float xo[3], const float v[3],
                                          const float Cn[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) a.state_out[(C00 + 3 * c + r) * a.stride + i] = Cn[c][r];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        a.state_out[(PX + ax) * a.stride + i] = xo[ax];
        a.state_out[(VX + ax) * a.stride + i] = v[ax];
    }
    a.id_out[i] = o;
    a.posvel[2 * o] = make_float4(xo[0], xo[1], xo[2], 0.0f);
    a.posvel[2 * o + 1] = make_float4(v[0], v[1], v[2], 0.0f);
}

// Ranks slot i for the next binning (every lane of the wave calls this).
// Near moves rank within (chunk c, direction) through the block's LDS
// counters; wild ones too, finished after the chunk (wild_rank); far moves
// take a final rank from a global atomic on their key.
__device__ __forceinline__ void rank_next(const BatchArgs &a, uint32_t *lcnt, uint32_t c,
                                          uint32_t i, uint32_t key, uint32_t dir, bool active) {
    const bool local = active && dir <= kDirWild;
    const uint32_t lr = count_rank(lcnt, dir, local);
    uint32_t fr = 0;
    if (__ballot(active && dir == kDirFar)) {  // rare: a particle left its tile's neighbourhood
        const bool far = active && dir == kDirFar;
        fr = count_rank(a.count, key, far);
    }
    if (active) {
        a.key[i] = key;
        a.rank[i] = local ? lr : fr;
        a.src[i] = dir < kDirs ? c * kDirs + dir : kRankFinal;
    }
}

// After a chunk: its near counts go to mig for k_bin_gather, and one lane
// reserves the wild particles' ranks on the scene's wild key.
// lcnt[kDirWild + 1] carries the wild base between the two barriers.
__device__ __forceinline__ void finish_chunk_counts(const BatchArgs &a, const SceneDesc *sd,
                                                    uint32_t *lcnt, uint32_t c) {
    __syncthreads();
    if (threadIdx.x < kDirs) a.mig[(size_t)c * kDirs + threadIdx.x] = lcnt[threadIdx.x];
    if (threadIdx.x == kDirWild) {
        const uint32_t nw = lcnt[kDirWild];
        lcnt[kDirWild + 1] = nw ? atomicAdd(&a.count[sd->koff + sd->ntiles], nw) : 0u;
    }
    __syncthreads();
}

template <int PPT, int TB, bool kSingle>
__device__ __forceinline__ void g2p_walk(const BatchArgs &a, float4 *lvel, uint32_t *lcnt) {
    constexpr int WC = TileGeo<TB>::win_cells;
    SceneDesc scene;
    const SceneDesc *sd = &scene;
    ChunkWalk cw(a, scene, std::integral_constant<bool, kSingle>{});
    // whether the binning has wild chunks at all (a batch: assume it has)
    const bool any_wild = !kSingle || cw.wild != 0;
    for (; cw.w.c < cw.w.end; cw.w.c += cw.w.step) {
        const TileChunk ch = cw.take(a);
        if (ch.y == kWildTile) continue;  // second loop
        if constexpr (!kSingle) scene = load_const(a.scenes + ch.x);
        const float dt = sd->prm.dt;
        const float rb[3] = {sd->uni.real_box[0], sd->uni.real_box[1], sd->uni.real_box[2]};
        if (threadIdx.x < kDirWild + 1) lcnt[threadIdx.x] = 0;
        // contiguous slots per task: coalesced state stores.  All state slots
        // first, then all positions and caller indices (slots past the
        // chunk's end load its last particle and store nothing).
        uint32_t src[PPT];
        bool valid[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = task_of(k) * 64 + (threadIdx.x & 63);
            valid[k] = j < ch.w;
            src[k] = a.perm[ch.z + min(j, ch.w - 1)];
        }
        float x[PPT][3];
        uint32_t oid[PPT];  // caller index of each particle
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            load_x(a.state, a.stride, src[k], x[k]);
            oid[k] = a.id[src[k]];
        }
        int t[3];
        tile_coords(ch.y, sd, t);
        for (int l = threadIdx.x; l < WC; l += kBlock) {
            const int64_t gi = win_cell_index<TB>(l, t, sd);
            lvel[l] = gi >= 0 ? a.vel[gi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        __syncthreads();
        bool wild[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t i = ch.z + task_of(k) * 64 + (threadIdx.x & 63);
            uint32_t key = 0, dir = kDirFar;
            if (valid[k]) {
                float v[3], B[3][3], xo[3], Cn[3][3];
                Node3 n;
                node3_init(x[k], n);
                g2p_gather(n, WinSinkG<TB>{lvel, win_l0<TB>(x[k])}, v, B);
                g2p_finish(x[k], rb, dt, v, B, xo, Cn);
                g2p_store(a, i, oid[k], xo, v, Cn);
                key = bin_key<TB>(a, sd, xo, t, &dir);
            }
            wild[k] = valid[k] && dir == kDirWild;
            // (a wave whose k-th task holds no particle skips the ranking together)
            if (a.count_next && __ballot(valid[k])) rank_next(a, lcnt, cw.w.c, i, key, dir, valid[k]);
        }
        if (a.count_next) {
            finish_chunk_counts(a, sd, lcnt, cw.w.c);
#pragma unroll
            for (int k = 0; k < PPT; ++k)
                if (wild[k]) {
                    const uint32_t i = ch.z + task_of(k) * 64 + (threadIdx.x & 63);
                    a.rank[i] += lcnt[kDirWild + 1];
                }
        }
        __syncthreads();
    }
    // Wild chunks (rare), in a loop of their own like k_p2g1's.  Every move
    // from a wild chunk is "far" (no tile to count from): final ranks from
    // global atomics, no LDS counts.
    if (!any_wild) return;
    for (TileWalk w = tile_walk(a); w.c < w.end; w.c += w.step) {
        const TileChunk ch = a.tchunks[w.c];
        if (ch.y != kWildTile) continue;
        const SceneDesc *sd = a.scenes + ch.x;
        const float dt = sd->prm.dt;
        const float rb[3] = {sd->uni.real_box[0], sd->uni.real_box[1], sd->uni.real_box[2]};
        const GridSinkG sink{a.vel + sd->goff, sd->G, sd->uni.Y, sd->uni.Z};
        const int nowhere[3] = {-4, -4, -4};
        for (uint32_t j0 = 0; j0 < ch.w; j0 += kBlock) {  // uniform trip count
            const uint32_t i = ch.z + j0 + threadIdx.x;
            const bool active = j0 + threadIdx.x < ch.w;
            uint32_t key = 0, dir = kDirFar;
            if (active) {
                float x[3], v[3], B[3][3], xo[3], Cn[3][3];
                const uint32_t si = a.perm[i];
                load_x(a.state, a.stride, si, x);
                Node3 n;
                node3_init(x, n);
                g2p_gather<true>(n, sink, v, B);
                g2p_finish(x, rb, dt, v, B, xo, Cn);
                g2p_store(a, i, a.id[si], xo, v, Cn);
                key = bin_key<TB>(a, sd, xo, nowhere, &dir);
                dir = kDirFar;
            }
            if (a.count_next) rank_next(a, lcnt, w.c, i, key, dir, active);
        }
    }
}

template <int PPT, int TB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kChunkWaves<PPT>)))
void k_g2p(BatchArgs a) {
    __shared__ float4 lvel[TileGeo<TB>::win_cells];
    __shared__ uint32_t lcnt[kDirWild + 2];  // per-direction counts, wild base
    if (a.nscenes == 1)
        g2p_walk<PPT, TB, true>(a, lvel, lcnt);
    else
        g2p_walk<PPT, TB, false>(a, lvel, lcnt);
}

__global__ void k_set_scene_uniforms(SceneDesc *d, Uniforms u) { d->uni = u; }

// Events armed for the next chunk/tile kernel launch (time_stages): the
// launch goes through hipExtLaunchKernel, whose events take the kernel's own
// start and end timestamps, so a slot's time excludes the dispatch gaps that
// events recorded between launches would add.
thread_local hipEvent_t t_timer_start = nullptr, t_timer_stop = nullptr;

template <typename K, typename... Args>
void launch_timed(K kernel, unsigned grid, unsigned block, hipStream_t st, Args... args) {
    if (t_timer_start) {
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, t_timer_start, t_timer_stop,
                              0, args...);
        t_timer_start = t_timer_stop = nullptr;
    } else {
        kernel<<<grid, block, 0, st>>>(args...);
    }
}

inline unsigned blocks_for(uint64_t count, uint64_t per) {
    return (unsigned)((count + per - 1) / per);
}

inline unsigned walk_blocks(const BatchArgs &a) {
    const unsigned nb = a.grid_blocks > 0 ? (unsigned)a.grid_blocks : kDefaultWalkBlocks;
    return ((nb + kXcds - 1) / kXcds) * kXcds;
}

template <int PPT, int TB>
hipError_t launch_ppt(int which, const BatchArgs &args, hipStream_t st) {
    const unsigned nb = walk_blocks(args);
    BatchArgs a = args;
    a.walk_grid = nb;
    switch (which) {
    case 1: launch_timed(k_p2g1<PPT, TB>, nb, kBlock, st, a); break;
    case 2: launch_timed(k_p2g2<PPT, TB>, nb, kBlock, st, a); break;
    case 3: launch_timed(k_g2p<PPT, TB>, nb, kBlock, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int TB>
hipError_t launch_by_ppt_tb(int which, const BatchArgs &a, hipStream_t st) {
    switch (a.ppt) {
    case 1: return launch_ppt<1, TB>(which, a, st);
    case 2: return launch_ppt<2, TB>(which, a, st);
    case 4: return launch_ppt<4, TB>(which, a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_by_ppt(int which, const BatchArgs &a, hipStream_t st) {
    switch (a.tile_bits) {
    case 2: return launch_by_ppt_tb<2>(which, a, st);
    case 3: return launch_by_ppt_tb<3>(which, a, st);
    default: return hipErrorInvalidValue;
    }
}

template <int PHASE, int TB>
hipError_t launch_reduce_tb(const BatchArgs &a, hipStream_t st) {
    if (a.stiles)
        launch_timed(k_reduce<PHASE, TB>, tile_blocks<TB>(a.stiles), ReduceGeo<TB>::threads, st, a);
    return hipGetLastError();
}

template <int PHASE>
hipError_t launch_reduce(const BatchArgs &a, hipStream_t st) {
    switch (a.tile_bits) {
    case 2: return launch_reduce_tb<PHASE, 2>(a, st);
    case 3: return launch_reduce_tb<PHASE, 3>(a, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_bin_gather(const BatchArgs &a, const BinBuffers &b, bool count,
                             hipStream_t st) {
    const uint32_t cap = (uint32_t)kBlock * (uint32_t)a.ppt;
    const unsigned nseg = blocks_for(b.nkeys, kSegKeys);
    if (count) {
        const unsigned nbin = std::max(1u, std::min(a.nchunks, 4096u));
        const unsigned zn = std::max(1u, std::min(blocks_for(b.nkeys, kBlock), 1024u));
        k_zero_u32<<<zn, kBlock, 0, st>>>(b.count, b.nkeys);
        k_zero_u32<<<1, kBlock, 0, st>>>(reinterpret_cast<uint32_t *>(b.sup), 2 * b.nsup);
        if (a.nchunks) {
            if (a.tile_bits == 3)
                k_bin_count<3><<<nbin, kBlock, 0, st>>>(a);
            else
                k_bin_count<2><<<nbin, kBlock, 0, st>>>(a);
        }
    }
    // many more keys than particle chunks: the sparse variant (measured on the
    // 64-scene mix); otherwise one wave per key
    const bool sparse = (uint64_t)b.nkeys > 16ull * ((uint64_t)a.total_slots / kBlock + 1);
    if (sparse) {
        const unsigned nsup = blocks_for(b.nkeys, 1u << kSupBits);
        launch_timed(k_bin_gather_sparse, (nsup + kXcds - 1) / kXcds * kXcds, kSegKeys * kWave, st,
                     a, b, cap, count ? 1 : 0);
    } else {
        launch_timed(k_bin_gather, (nseg + kXcds - 1) / kXcds * kXcds, kSegKeys * kWave, st, a, b,
                     cap, count ? 1 : 0);
    }
    return hipGetLastError();
}

hipError_t launch_bin_scan(const BatchArgs &a, const BinBuffers &b, hipStream_t st) {
    const uint32_t cap = (uint32_t)kBlock * (uint32_t)a.ppt;
    launch_timed(k_scan_keys, blocks_for(b.nkeys, kWave), kScanWaves * kWave, st, a, b, cap);
    return hipGetLastError();
}

hipError_t launch_bin_scatter(const BatchArgs &a, const BinBuffers &b, hipStream_t st) {
    const unsigned nbin = std::max(1u, std::min(a.nchunks, 4096u));
    if (a.nchunks) launch_timed(k_bin_scatter, nbin, kBlock, st, a, b);
    return hipGetLastError();
}

hipError_t launch_gather_binned(const BatchArgs &a, float *dst_state, uint32_t *dst_id,
                                hipStream_t st) {
    if (a.total_slots)
        k_gather_binned<<<std::min(blocks_for(a.total_slots, kBlock), 4096u), kBlock, 0, st>>>(
            a, dst_state, dst_id);
    return hipGetLastError();
}

hipError_t launch_bin(const BatchArgs &a, const BinBuffers &b, bool count, hipStream_t st) {
    hipError_t e = launch_bin_gather(a, b, count, st);
    if (e == hipSuccess) e = launch_bin_scan(a, b, st);
    if (e == hipSuccess) e = launch_bin_scatter(a, b, st);
    return e;
}

hipError_t launch_p2g1(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(1, a, st); }
hipError_t launch_p2g2(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(2, a, st); }
hipError_t launch_g2p(const BatchArgs &a, hipStream_t st) { return launch_by_ppt(3, a, st); }

hipError_t launch_reduce1(const BatchArgs &a, hipStream_t st) { return launch_reduce<1>(a, st); }

hipError_t launch_reduce2(const BatchArgs &a, hipStream_t st) { return launch_reduce<2>(a, st); }

void arm_kernel_timer(hipEvent_t start, hipEvent_t stop) {
    t_timer_start = start;
    t_timer_stop = stop;
}

bool disarm_kernel_timer() {
    const bool was = t_timer_start != nullptr;
    t_timer_start = t_timer_stop = nullptr;
    return was;
}

hipError_t launch_set_scene_uniforms(SceneDesc *d, const Uniforms &u, hipStream_t st) {
    k_set_scene_uniforms<<<1, 1, 0, st>>>(d, u);
    return hipGetLastError();
}

}  // namespace mpm
