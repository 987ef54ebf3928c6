// mpm_batch.h -- device state and production substep of a batch of
// independent MLS-MPM scenes (a single simulator is a batch of one).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "mpm_fused.h"
#include "mpm_internal.h"

namespace mpm {

struct Tuning {
    int path = 0;          // 0: binned tile-partial kernels, 1: reference-shaped stage kernels
    int ppt = 2;           // particles per thread: tile chunks of <= 256 * ppt particles
    int grid_blocks = 0;   // persistent blocks of the chunk kernels (0: 4096)
    int xcd_remap = 0;     // 1: contiguous chunk ranges per XCD (0: runs of 16, round robin)
    int tile_bits = 2;     // P2G tiles of (2^tile_bits)^3 cells: 2 (4^3) or 3 (8^3)
};

// Uniforms of a scene from its init and real boxes.
Uniforms derive_uniforms(const float init_box[3], const float real_box[3]);

struct SceneInit {
    const void *particles;  // n x 80-B reference records (host)
    uint32_t n;
    float init_box[3];
    Params prm;
};

// Differentiable rollout (mpm_backward): loss = |com - target|^2 +
// ke_weight * mean |v|^2 after `substeps` substeps.
struct AdjointConfig {
    uint32_t substeps;
    uint32_t checkpoint_every;
    double target[3];
    double ke_weight;
};

struct AdjointResult {
    double loss = 0.0;
    float *grad_v0 = nullptr;  // n x 3 (host, caller order) or null
    float *grad_x0 = nullptr;  // n x 3 (host, caller order) or null
    float grad_stiffness = 0.0f;
    float grad_viscosity = 0.0f;
};

class Batch {
public:
    Batch() = default;
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    ~Batch() { release(); }

    // (Re)creates the batch: packs every scene's particles, grid and keys.
    // On failure everything is released and ready() is false.
    hipError_t init(const SceneInit *scenes, uint32_t count, hipStream_t st);
    void release();

    // False until init succeeds, and again after a failed init.
    bool ready() const { return !h_scenes_.empty(); }
    uint32_t scenes() const { return (uint32_t)h_scenes_.size(); }
    const SceneDesc &scene(uint32_t s) const { return h_scenes_[s]; }

    hipError_t set_box(uint32_t s, const float real_box[3], hipStream_t st);
    hipError_t substep(hipStream_t st);
    hipError_t time_stages(hipStream_t st, unsigned substeps, float *ms, int *calls,
                           const char **names, int capacity, int *count);

    hipError_t read_posvel(uint32_t s, void *dst, uint32_t n, hipStream_t st);
    hipError_t read_particles(uint32_t s, void *dst, uint32_t n, hipStream_t st);
    hipError_t write_particles(uint32_t s, const void *src, uint32_t n, hipStream_t st);
    hipError_t read_grid(uint32_t s, int32_t *dst, uint64_t ncells, hipStream_t st);
    // Decoded updated grid velocities (vx, vy, vz, 0) of the last production
    // substep, G x 4 floats.
    hipError_t read_vel(uint32_t s, float *dst, uint64_t ncells, hipStream_t st);
    hipError_t write_grid(uint32_t s, const int32_t *src, uint64_t ncells, hipStream_t st);
    void *posvel_ptr(uint32_t s) const { return posvel_ + 2 * (size_t)h_scenes_[s].poff; }

    // View of scene 0 for the reference-shaped stage kernels.
    DeviceScene stage_view() const;
    // Call after changing particles or grid cells outside substep() (stage
    // kernels, copies): the next substep re-counts the keys and rewrites
    // every cell.
    hipError_t external_change(hipStream_t st);

    const Tuning &tuning() const { return tun_; }
    // A new tile size rebuilds the key geometry and tile tables (after
    // waiting for `st`); the buffers are sized for both tile sizes.
    hipError_t set_tuning(const Tuning &t, hipStream_t st);
    // Times candidate tunings on a snapshot of the current state (restored
    // afterwards) and keeps the fastest.
    hipError_t autotune(hipStream_t st, unsigned substeps_per_trial, Tuning *chosen);

    // Forward rollout from the current state with checkpoints, then the
    // adjoint back to the current state (single-scene batches).  Leaves the
    // batch at the rollout's final state.
    hipError_t backward(const AdjointConfig &cfg, hipStream_t st, AdjointResult &res);

private:
    hipError_t init_impl(const SceneInit *scenes, uint32_t count, hipStream_t st);
    hipError_t ensure_capacity(uint32_t total_n, uint64_t cells, uint64_t keys,
                               uint64_t tchunks, uint64_t partial_cells);
    // Key geometry of every scene for tiles of (2^tb)^3 cells (koff, soff,
    // ntiles, nkeys); returns the keys and spatial tiles of all scenes.
    void layout_keys(int tb, uint64_t *keys, uint64_t *tiles);
    // Makes (2^tb)^3 tiles current: key geometry, Hilbert ranks, neighbour
    // and tile tables, scene records; every tile marked dirty.
    hipError_t apply_tile_geometry(int tb);
    hipError_t upload_chunks();
    BatchArgs args() const;
    BinBuffers bin_buffers() const;
    uint32_t tchunk_cap() const;
    // Per-slot events of time_stages (slots in launch order).
    struct SlotTimes {
        static constexpr int kMax = 8;
        hipEvent_t start[kMax], stop[kMax];
        bool used[kMax];
    };
    hipError_t launch_substep(hipStream_t st, SlotTimes *t);
    // One path-0 substep that also records its binned state, tile chunks and
    // raw grid at `rec` (layout private to the adjoint).
    hipError_t substep_taped(hipStream_t st, char *rec, const void *layout);
    void drop_graphs();
    void after_substep(hipStream_t st);

    std::vector<SceneDesc> h_scenes_;
    std::vector<Chunk> h_chunks_;
    SceneDesc *d_scenes_ = nullptr;
    Chunk *d_chunks_ = nullptr;
    uint32_t *d_tile_rank_ = nullptr;  // per scene: x-major tile -> Hilbert rank
    uint32_t *d_tile_xmaj_ = nullptr;  // key tile -> x-major tile (kWildTile: wild)
    uint32_t *d_nbr_ = nullptr;        // 27 neighbour key tiles per spatial tile
    uint4 *d_tile_info_ = nullptr;     // (x, y, z, scene) per spatial tile
    uint8_t *d_dirty_ = nullptr;       // per spatial tile: cells may be non-zero
    uint8_t *d_near_ = nullptr;        // per spatial tile, two substep parities
    uint32_t *d_key_scene_ = nullptr;  // per key
    size_t key_scene_cap_ = 0;
    size_t tile_rank_cap_ = 0, tile_xmaj_cap_ = 0, tile_tab_cap_ = 0;
    size_t chunk_cap_ = 0;

    uint32_t total_n_ = 0, stride_ = 0;
    uint64_t total_cells_ = 0, total_keys_ = 0;
    uint32_t total_tiles_ = 0;  // spatial tiles of all scenes
    uint32_t cap_n_ = 0;
    uint64_t cap_cells_ = 0, cap_keys_ = 0, cap_tchunks_ = 0, cap_partial_ = 0;
    uint32_t scene_cap_ = 0;

    float *state_ = nullptr, *spare_state_ = nullptr;
    uint32_t *id_ = nullptr, *spare_id_ = nullptr;
    int4 *grid_ = nullptr;
    float4 *posvel_ = nullptr;
    float *staging_ = nullptr;  // 80-B records of every particle
    float4 *vel_ = nullptr;     // decoded updated grid velocities
    int4 *wild_ = nullptr;      // wild-particle accumulators, zero between substeps
    int4 *partial_ = nullptr;   // one window (216 or 1000 cells) per tile chunk
    uint32_t *key_ = nullptr, *rank_ = nullptr, *src_ = nullptr, *count_ = nullptr,
             *offset_ = nullptr;
    uint32_t *mig_ = nullptr;  // 27 per tile chunk
    uint32_t *order_ = nullptr;  // per bin-order slot: state slot in P2G level order
    uint32_t *perm_ = nullptr;   // bin-order slot -> slot of state_ (written by the binning)
    uint2 *sup_ = nullptr;  // scan totals per 256 keys, two parities
    size_t sup_cap_ = 0;
    float *state_buf0_ = nullptr;  // parity 0 = this buffer is the current state
    TileChunk *tchunks_ = nullptr;
    uint32_t *ntchunks_ = nullptr;
    uint2 *tile_chunk_ = nullptr;  // key tile -> (first tile chunk, count)

    Tuning tun_{};
    char *adj_ws_ = nullptr;  // backward() workspace, kept between calls
    size_t adj_ws_cap_ = 0;

    // True when key_/rank_/count_ hold the binning counts of state_ (left by
    // the last g2p); false after anything else changed the particles.
    bool keys_ready_ = false;

    // A captured substep is valid exactly while the kernel arguments it was
    // captured with are the current ones (they are value-initialised, so
    // padding compares equal too).
    struct GraphEntry {
        BatchArgs args;
        BinBuffers bins;
        DeviceScene stage;  // the reference-shaped path's arguments
        int path;
        bool keys_ready;
        hipGraphExec_t exec;
    };
    static constexpr int kMaxGraphs = 6;
    GraphEntry graphs_[kMaxGraphs] = {};
    int ngraphs_ = 0;
    // the stream the cached graphs were launched on: drop_graphs() waits for
    // it, so no executable graph is destroyed while a launch of it may run
    hipStream_t graph_stream_ = nullptr;
    bool graph_launched_ = false;
};

}  // namespace mpm
