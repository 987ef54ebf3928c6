// mpm_internal.h -- launch interface between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.h"

namespace mpm {

// Device buffers of one scene.
struct DeviceScene {
    float *state;        // NCOMP x stride floats (SoA particle state)
    uint32_t *id;        // storage slot -> caller particle index
    uint32_t stride;     // >= n, multiple of 64
    uint32_t n;          // particles
    int4 *grid;          // G cells: (vx, vy, vz, mass) i32
    uint32_t G;          // cells
    float4 *posvel;      // 2 x n float4: (position, 0), (v, 0)
    Uniforms *uni;       // device copy of the box uniforms
    Params prm;
    float init_box[3];
};

// Reference-shaped kernels: one launch per WGSL kernel, one thread per
// particle or cell, global i32 atomics (mls-mpm/mls-mpm.ts:248-265).
hipError_t launch_stage(int stage, const DeviceScene &s, hipStream_t st);

// The same kernels over n particles in the reference's 80-B records (read and
// written in place) with PosVel in record order: the allocation-free step.
hipError_t launch_stage_records(int stage, float *records, uint32_t n, int4 *grid, uint32_t G,
                                const Uniforms *uni, const Params &prm, float4 *posvel,
                                hipStream_t st);

// Scratch of the allocation-free step on the current device: uniforms at
// *uni, then max_cells grid cells at *grid (a static device array; its
// address is looked up once per device).
hipError_t noalloc_scratch(Uniforms **uni, int4 **grid, uint32_t *max_cells);

// Writes the uniforms with a one-thread kernel so the update is ordered on
// the stream like queue.writeBuffer (mls-mpm/mls-mpm.ts:274).
hipError_t launch_set_uniforms(Uniforms *dst, const Uniforms &src, hipStream_t st);

// SoA <-> 80-B reference layout conversion on the device; unpack resets the
// storage order to the caller's (id = identity), pack writes caller order.
hipError_t launch_unpack_particles(const float *packed, const DeviceScene &s, hipStream_t st);
hipError_t launch_pack_particles(float *packed, const DeviceScene &s, hipStream_t st);

}  // namespace mpm
