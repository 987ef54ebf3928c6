// mpm_path.h -- the production substep of one scene: kernel sequence,
// auxiliary buffers and the hipGraphs that replay it.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_fused.h"
#include "mpm_internal.h"

namespace mpm {

struct Tuning {
    int path = 0;          // 0: binned LDS-window kernels, 1: reference-shaped stage kernels
    int ppt = 2;           // particles per thread in P2G/G2P blocks (1, 2, 4)
    int wmax = 2048;       // LDS window budget in cells
    int sort_period = 1;   // re-bin every k substeps
};

class FusedPath {
public:
    FusedPath() = default;
    FusedPath(const FusedPath &) = delete;
    FusedPath &operator=(const FusedPath &) = delete;
    ~FusedPath() { release(); }

    // One substep of the scene on `st`.  May swap s.state / s.id with the
    // path's spare buffers (after binning).
    hipError_t substep(DeviceScene &s, hipStream_t st);

    // Runs `substeps` substeps eagerly with HIP events around each kernel
    // group; ms[i] = summed time of slot i, calls[i] = launches of slot i.
    hipError_t time_stages(DeviceScene &s, hipStream_t st, unsigned substeps, float *ms,
                           int *calls, const char **names, int capacity, int *count);

    // Scene buffers or sizes changed: drop graphs and aux buffers.
    void invalidate();
    // Particle storage was rewritten in caller order: re-bin next substep.
    void particles_changed() { since_sort_ = 0; }
    void release();

    const Tuning &tuning() const { return tun_; }
    hipError_t set_tuning(const Tuning &t);

    // Number of substeps issued since the last binning pass.
    unsigned since_sort() const { return since_sort_; }

private:
    hipError_t ensure_buffers(const DeviceScene &s);
    hipError_t launch_substep(DeviceScene &s, hipStream_t st, bool sort, hipEvent_t *ev);
    FusedArgs args(const DeviceScene &s) const;
    void drop_graphs();

    struct GraphEntry {
        bool sort;
        float *state;
        hipGraphExec_t exec;
    };
    GraphEntry graphs_[4] = {};
    int ngraphs_ = 0;
    DeviceScene captured_{};

    Tuning tun_{};
    unsigned since_sort_ = 0;

    // aux buffers
    uint32_t cap_ = 0;       // particle capacity of the aux buffers
    uint64_t key_cap_ = 0;   // key capacity
    float *spare_state_ = nullptr;
    uint32_t *spare_id_ = nullptr;
    uint32_t *key_ = nullptr, *rank_ = nullptr, *count_ = nullptr, *offset_ = nullptr,
             *block_sums_ = nullptr;
    KeyGeom kg_{};
};

}  // namespace mpm
