// mpm_path.h -- the production substep of one scene: kernel sequence,
// auxiliary buffers and the hipGraph that replays it.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_internal.h"

namespace mpm {

class FusedPath {
public:
    FusedPath() = default;
    FusedPath(const FusedPath &) = delete;
    FusedPath &operator=(const FusedPath &) = delete;
    ~FusedPath() { release(); }

    // One substep of the scene on `st` (replayed from a hipGraph after the
    // first call for a given scene shape).
    hipError_t substep(const DeviceScene &s, hipStream_t st);

    // Runs `substeps` substeps eagerly with HIP events around each kernel.
    hipError_t time_stages(const DeviceScene &s, hipStream_t st, unsigned substeps, float *ms,
                           const char **names, int capacity, int *count);

    // Scene buffers or sizes changed: drop the captured graph.
    void invalidate();
    // Particle data was overwritten from the host.
    void particles_changed() {}
    void release();

private:
    hipError_t launch_sequence(const DeviceScene &s, hipStream_t st, hipEvent_t *ev, int *nev);

    hipGraphExec_t exec_ = nullptr;
    DeviceScene captured_{};
};

}  // namespace mpm
