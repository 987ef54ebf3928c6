// mpm_device.h -- per-particle MLS-MPM arithmetic shared by the gfx950 kernels.
//
// Every float operation is one IEEE binary32 operation in the order the
// reference WGSL writes it; the library is compiled with -ffp-contract=off and
// the pragma below, so no mul+add is fused and results are bit-reproducible
// against a correctly rounded CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mpm {

// Particle state, structure-of-arrays: component k of particle i lives at
// state[k * stride + i].  C is column-major: C[col][row] at k = 6 + 3*col + row.
enum : int { PX = 0, PY, PZ, VX, VY, VZ, C00, NCOMP = 15 };

// Simulation "uniforms" (the reference's init_box_size / real_box_size
// uniform buffers, mls-mpm/mls-mpm.ts:131-142) plus values derived from them.
struct Uniforms {
    float real_box[3];
    float init_box[3];
    int32_t Y, Z;        // i32(init_box.y), i32(init_box.z): grid strides (p2g_1.wgsl:53-56)
    int32_t wall_hi[3];  // i32(ceil(real_box) - 3)                        (updateGrid.wgsl:41-43)
    int32_t pad;
};

struct Params {
    float stiffness, rest_density, viscosity, dt, mult;
};

// Loads through the constant address space: scalar loads, which the compiler
// may also move out of loops.  Only for data that no thread writes while the
// kernel runs (scene descriptors; the chunk list and counts of a binning,
// written by an earlier kernel).
template <typename T>
__device__ __forceinline__ T load_const(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T *)(uintptr_t)p;
#else
    return *p;  // (the host pass of a device function; never executed)
#endif
}

// i32(f32): truncate toward zero, saturating, NaN -> 0.  That is exactly
// what gfx950's V_CVT_I32_F32 does, so the conversion is that one
// instruction (C's (int)f leaves out-of-range values undefined and would
// make the compiler emit range checks).
__device__ __forceinline__ int32_t f2i_sat(float f) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// encodeFixedPoint -- mls-mpm/p2g_1.wgsl:15-17
__device__ __forceinline__ int32_t encode(float f, float mult) { return f2i_sat(f * mult); }

// decodeFixedPoint -- mls-mpm/p2g_2.wgsl:22-24 (IEEE division)
__device__ __forceinline__ float decode(int32_t i, float mult) { return (float)i / mult; }

// Quadratic B-spline stencil of one particle -- mls-mpm/p2g_1.wgsl:29-34.
struct Stencil {
    float ci[3];    // floor(position)
    float w[3][3];  // w[k][axis]
};

__device__ __forceinline__ void stencil_init(const float x[3], Stencil &s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float ci = floorf(x[a]);
        float d = x[a] - (ci + 0.5f);
        s.ci[a] = ci;
        s.w[0][a] = (0.5f * (0.5f - d)) * (0.5f - d);
        s.w[1][a] = 0.75f - d * d;
        s.w[2][a] = (0.5f * (0.5f + d)) * (0.5f + d);
    }
}

__device__ __forceinline__ float node_weight(const Stencil &s, int gx, int gy, int gz) {
    return (s.w[gx][0] * s.w[gy][1]) * s.w[gz][2];
}

// cell_x, cell_dist and the i32 cell_index of node (gx,gy,gz) -- p2g_1.wgsl:42-56.
// Index arithmetic wraps like WGSL i32 arithmetic.
__device__ __forceinline__ int32_t node_index(const Stencil &s, const float x[3], int gx,
                                              int gy, int gz, int32_t Y, int32_t Z,
                                              float dist[3]) {
    const int g[3] = {gx, gy, gz};
    int32_t ic[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float cx = (s.ci[a] + (float)g[a]) - 1.0f;
        dist[a] = (cx + 0.5f) - x[a];
        ic[a] = f2i_sat(cx);
    }
    uint32_t idx = (uint32_t)ic[0] * (uint32_t)Y * (uint32_t)Z + (uint32_t)ic[1] * (uint32_t)Z +
                   (uint32_t)ic[2];
    return (int32_t)idx;
}

__device__ __forceinline__ bool in_grid(int32_t idx, uint32_t G) {
    return idx >= 0 && (uint32_t)idx < G;
}

}  // namespace mpm
