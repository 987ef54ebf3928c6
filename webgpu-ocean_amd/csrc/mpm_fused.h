// mpm_fused.h -- launch interface of the production substep kernels.
//
// The kernels serve a *batch* of independent scenes (one scene is a batch of
// one): particle slots, grid cells and bin keys of all scenes are packed into
// single buffers, and every stage is ONE launch over a table of particle
// chunks, each chunk lying inside one scene.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "mpm_device.h"

namespace mpm {

// Bin keys are tile-major: tiles of (2^kTileBits)^3 cells ordered along a 3D
// Hilbert curve (consecutive tiles share a face, so a run of particles in key
// order covers a compact box), cells x-major inside a tile.
constexpr int kTileBits = 2;
constexpr int kBlock = 256;

// Per-scene constants, device resident.
struct SceneDesc {
    Uniforms uni;        // boxes and derived strides / wall limits
    Params prm;          // WGSL override constants of this scene
    uint32_t poff, n;    // particle slots [poff, poff + n)
    uint32_t goff, G;    // grid cells [goff, goff + G)
    uint32_t koff;       // first bin key of the scene
    int32_t kdim[3];     // ceil(init_box): cells per axis used for keys
    uint32_t knt[3];     // key tiles per axis
    uint32_t nkeys;      // knt0*knt1*knt2 * 2^(3*kTileBits)
    uint32_t rank_off;   // first entry of the scene's tile -> Hilbert rank table
};

// A run of consecutive particle slots of one scene, handled by one block.
struct Chunk {
    uint32_t scene, begin, count, pad;
};

struct BatchArgs {
    const SceneDesc *scenes;
    const uint32_t *tile_rank;  // per scene: x-major tile index -> Hilbert rank
    const Chunk *chunks;
    uint32_t nchunks;
    float *state;        // SoA particle state, NCOMP x stride, bin order
    uint32_t *id;        // slot -> global caller index (poff + local index)
    uint32_t stride;
    int4 *grid;          // all scenes' cells
    float4 *posvel;      // 2 float4 per global caller index
    int ppt;             // particles per thread: chunk = kBlock * ppt
    int max_slots;       // tile windows per block (1..16), 216 cells x 16 B each
    int grid_blocks;     // persistent blocks per launch (0: one per chunk)
    int xcd_remap;       // give each XCD a contiguous range of chunks
};

struct BinBuffers {
    uint32_t *key, *rank;      // per slot
    uint32_t *count, *offset;  // per key; count is zero on entry and exit
    uint32_t *block_sums;      // ceil(nkeys / 4096)
    uint32_t nkeys;            // all scenes
    float *state_out;
    uint32_t *id_out;
};

// Derives key geometry of one scene from its init box.
void scene_key_geometry(SceneDesc &d);

// Appends the scene's tile -> Hilbert-rank table (knt0*knt1*knt2 entries).
void append_hilbert_ranks(const SceneDesc &d, std::vector<uint32_t> &ranks);

hipError_t launch_bin(const BatchArgs &a, const BinBuffers &b, hipStream_t st);
hipError_t launch_clear(int4 *grid, uint32_t cells, hipStream_t st);
hipError_t launch_p2g1(const BatchArgs &a, hipStream_t st);
hipError_t launch_p2g2(const BatchArgs &a, hipStream_t st);
hipError_t launch_g2p(const BatchArgs &a, hipStream_t st);
// Writes uniforms of one scene (stream-ordered, like queue.writeBuffer).
hipError_t launch_set_scene_uniforms(SceneDesc *d, const Uniforms &u, hipStream_t st);

}  // namespace mpm
