// mpm_fused.h -- launch interface of the production substep kernels.
//
// The kernels serve a *batch* of independent scenes (one scene is a batch of
// one): particle slots, grid cells and bin keys of all scenes are packed into
// shared buffers, and every stage is one launch over all scenes.
//
// Substep: bin (counting sort of particles by tile key; the scan also emits
// the list of tile chunks) -> p2g_1 (LDS tile windows -> partial windows) ->
// reduce_1 (per tile: sum the partials covering it; writes every cell) ->
// p2g_2 -> reduce_2 (+ updateGrid -> decoded velocities) -> g2p (+ the
// counts of the next binning).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "mpm_device.h"

namespace mpm {

// Tiles are (2^TB)^3 cells, TB = 2 (4^3, the default) or 3 (8^3), chosen per
// batch (Tuning::tile_bits, autotuned), ordered along a 3D Hilbert curve per
// scene; bin key = koff + tile rank.  Each scene has one extra key (rank =
// ntiles) collecting the particles whose stencil leaves the grid ("wild"
// particles).  A tile chunk's window is the tile plus one cell on every side.
template <int TB>
struct TileGeo {
    static_assert(TB == 2 || TB == 3, "tiles of 4^3 or 8^3 cells");
    static constexpr int bits = TB;
    static constexpr int edge = 1 << TB;                          // 4 | 8
    static constexpr int cells = edge * edge * edge;              // 64 | 512
    static constexpr int win_edge = edge + 2;                     // 6 | 10
    static constexpr int win_cells = win_edge * win_edge * win_edge;  // 216 | 1000
};
constexpr int kMaxWinCells = TileGeo<3>::win_cells;  // partial window stride bound

// The 4^3 tile: the default, and the one the adjoint's reverse kernels use.
constexpr int kTileBits = 2;
constexpr int kBlock = 256;
constexpr int kTileEdge = TileGeo<kTileBits>::edge;
constexpr int kTileCells = TileGeo<kTileBits>::cells;
constexpr int kWinEdge = TileGeo<kTileBits>::win_edge;
constexpr int kWinCells = TileGeo<kTileBits>::win_cells;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X, each with its own L2
constexpr uint32_t kSegBits = 4;   // 16 keys per block of the count-path totals kernel
constexpr uint32_t kSupBits = 8;   // binning scan totals per 256 keys (super-segment; g2p
                                   // adds to them with atomics, so they must not be too few)
constexpr uint32_t kWildTile = 0xffffffffu;
// Words of the walk head buffer: chunk count, chunks of the last key, then
// the kXcds + 1 XCD range boundaries (BatchArgs::ntchunks, ::xcd_first).
constexpr uint32_t kWalkHeadWords = 2 + kXcds + 1;
constexpr uint32_t kNoTile = 0xffffffffu;  // neighbour outside the grid

// Per-scene constants, device resident.
struct SceneDesc {
    Uniforms uni;        // boxes and derived strides / wall limits
    Params prm;          // WGSL override constants of this scene
    uint32_t poff, n;    // particle slots [poff, poff + n)
    uint32_t goff, G;    // grid cells [goff, goff + G)
    uint32_t koff;       // first bin key of the scene
    uint32_t toff;       // = koff (keys are tiles)
    uint32_t soff;       // first spatial tile in the reduce launch
    uint32_t ntiles;     // spatial tiles knt0*knt1*knt2 (the wild tile is extra)
    int32_t kdim[3];     // ceil(init_box): cells per axis
    uint32_t knt[3];     // tiles per axis
    uint32_t nkeys;      // ntiles + 1
    uint32_t rank_off;   // first entry of the scene's x-major tile -> rank table
    uint32_t fast;       // 1: integer box (cells map to distinct indices), 0: all wild
    uint32_t tile_bits;  // tiles of (2^tile_bits)^3 cells (the batch's)
};

// A run of consecutive particle slots of one scene (static partition used by
// the binning kernels).
struct Chunk {
    uint32_t scene, begin, count, pad;
};

// A run of at most 256*ppt particles of one tile (emitted by the scan):
// x = scene, y = x-major tile index or kWildTile, z = first slot, w = count.
using TileChunk = uint4;

struct BatchArgs {
    const SceneDesc *scenes;
    uint32_t nscenes;
    const uint32_t *tile_rank;   // rank_off + x-major tile -> Hilbert rank
    const uint32_t *tile_xmaj;   // key (koff + rank) -> x-major tile (kWildTile: wild)
    const Chunk *chunks;         // static slot chunks
    uint32_t nchunks;
    const TileChunk *tchunks;    // tile chunks of the current binning
    const uint32_t *ntchunks;    // [0]: their number, [1]: chunks of the last key (a
                                 // single scene's wild key), both written by the scan
    const uint32_t *xcd_first;   // kXcds + 1 chunk indices: XCD x walks chunks
                                 // [xcd_first[x], xcd_first[x+1]) (equal particle shares)
    uint32_t total_slots;        // particles of all scenes
    uint32_t tchunk_cap;         // capacity of tchunks, partial and mig (in chunks)
    const uint2 *tile_chunk;     // key -> (first tile chunk, count)
    uint32_t stiles;             // spatial tiles of all scenes (reduce launch)
    const uint32_t *nbr;         // 27 per spatial tile: key of each neighbour or kNoTile
    const uint4 *tile_info;      // per spatial tile: (tile x, y, z, scene)
    uint8_t *dirty;              // per spatial tile: cells or velocities may be non-zero
    uint8_t *near;               // per spatial tile: particles within reach (this binning),
                                 // or a scene with wild particles / a non-integral box
    uint8_t *near_prev;          // the same for the previous binning
    const uint32_t *key_scene;   // per key: its scene
    float *state;                // SoA particle state, NCOMP x stride, in the order the
                                 // previous substep left (read through perm)
    uint32_t *id;                // slot of state -> global caller index (poff + local index)
    const uint32_t *perm;        // slot of this substep's bin order -> slot of state
    float *state_out;            // g2p's output state, in this substep's bin order
    uint32_t *id_out;            // its caller indices
    uint32_t stride;
    int4 *grid;                  // all scenes' cells: raw P2G accumulators (vx,vy,vz,mass)
    float4 *vel;                 // decoded updated velocities (written by reduce_2)
    int4 *wild;                  // global-atomic accumulators of wild particles;
                                 // all zero between substeps
    int4 *partial;               // TileGeo<tile_bits>::win_cells entries per tile chunk
    float4 *posvel;              // 2 float4 per global caller index
    uint32_t *key, *rank;        // per slot: bin key and rank within the key (or within
                                 // its source chunk and direction, see src)
    uint32_t *src;               // per slot: chunk * 27 + direction, or final-rank marker
    uint32_t *mig;               // 27 per tile chunk: particles moving each direction,
                                 // then their offset within the destination tile
    uint32_t *order;             // per bin-order position of a window chunk: the state slot
                                 // of the particle p2g_1 processed there (its level order)
    uint32_t *count;             // per key: particles (reduce_2 clears it for g2p's counting)
    int count_next;              // g2p also counts the keys of its output
    int ppt;                     // particles per thread: chunk <= kBlock * ppt
    int grid_blocks;             // persistent blocks of the chunk kernels (0: default)
    uint32_t walk_grid;          // blocks of the chunk kernel being launched (= gridDim.x;
                                 // set by its launcher, a kernel argument loaded with the rest)
    int xcd_remap;               // contiguous chunk ranges per XCD
    int tile_bits;               // 2: 4^3 tiles, 3: 8^3 tiles
};

struct BinBuffers {
    uint32_t *count, *offset;  // per key; reduce_2 clears count
    uint2 *sup;                // (particles, chunks) totals per 256 keys, this substep
    uint2 *sup_next;           // the next substep's (the scan clears them; k_bin_gather adds)
    uint32_t nsup;
    uint32_t nkeys;            // all scenes
    uint32_t *perm_out;        // bin-order slot -> slot of the state (the scatter writes it)
    TileChunk *tchunks;
    uint32_t tchunk_cap;
    uint32_t *ntchunks;        // [0] count, [1] chunks of the last key
    uint32_t *xcd_first;       // kXcds + 1 entries, written by the scan
    uint2 *tile_chunk;
};

// Derives key/tile geometry of one scene from its init box, for tiles of
// (2^tile_bits)^3 cells.
void scene_key_geometry(SceneDesc &d, int tile_bits);

// Appends the scene's x-major tile -> Hilbert-rank table (ntiles entries) and
// its inverse (ntiles + 1 entries, the last = kWildTile).
void append_hilbert_ranks(const SceneDesc &d, std::vector<uint32_t> &ranks,
                          std::vector<uint32_t> &xmaj);

// Appends the reduce tables of scene `s` (spatial tiles in Hilbert order):
// 27 neighbour key tiles per tile and the tile's coordinates.  `ranks` and
// `xmaj` are the scene's own tables from append_hilbert_ranks.
void append_tile_tables(const SceneDesc &d, uint32_t s, const uint32_t *ranks,
                        const uint32_t *xmaj, std::vector<uint32_t> &nbr,
                        std::vector<uint4> &info);

// Counting sort of the slots of a.state by bin key into the permutation
// b.perm_out (the state itself stays where it is; the chunk kernels read it
// through the permutation and g2p writes its output in bin order).  With
// `count`, the keys, counts and key-range totals are taken first (after
// clearing count); otherwise the previous g2p left them.
hipError_t launch_bin(const BatchArgs &a, const BinBuffers &b, bool count, hipStream_t st);
// The three parts of launch_bin: (with `count`, keys and ranks from scratch,
// then) the per-key gather of the counts (skipped when g2p published them
// with atomics); key scan and tile chunks; the permutation.
hipError_t launch_bin_gather(const BatchArgs &a, const BinBuffers &b, bool count,
                             hipStream_t st);
hipError_t launch_bin_scan(const BatchArgs &a, const BinBuffers &b, hipStream_t st);
hipError_t launch_bin_scatter(const BatchArgs &a, const BinBuffers &b, hipStream_t st);
// Copies the state into bin order through a.perm (dst_state: NCOMP x a.stride,
// dst_id: a.stride), for the adjoint's tape.
hipError_t launch_gather_binned(const BatchArgs &a, float *dst_state, uint32_t *dst_id,
                                hipStream_t st);
hipError_t launch_p2g1(const BatchArgs &a, hipStream_t st);
hipError_t launch_reduce1(const BatchArgs &a, hipStream_t st);
hipError_t launch_p2g2(const BatchArgs &a, hipStream_t st);
hipError_t launch_reduce2(const BatchArgs &a, hipStream_t st);
hipError_t launch_g2p(const BatchArgs &a, hipStream_t st);
// Kernel timing for time_stages: the next chunk/tile kernel launched on this
// thread records `start`/`stop` at its own start and end (hipExtLaunchKernel).
// disarm_kernel_timer() returns whether the events were still unused.
void arm_kernel_timer(hipEvent_t start, hipEvent_t stop);
bool disarm_kernel_timer();
// Writes uniforms of one scene (stream-ordered, like queue.writeBuffer).
hipError_t launch_set_scene_uniforms(SceneDesc *d, const Uniforms &u, hipStream_t st);

}  // namespace mpm
