// render.hip -- the screen-space fluid renderer of WebGPU-Ocean
// (render/fluidRender.ts, render/*.wgsl) as gfx950 compute kernels behind the
// C ABI of include/render.h.
//
// A frame (FluidRenderer.execute, fluidRender.ts:424-484):
//   k_project    per particle: its impostor's screen rectangle and the number
//                of kTile^2 screen tiles it touches
//   scan         pair offsets (rocPRIM exclusive scan); the host reads the
//                total to size the pair buffers (one wait per frame)
//   k_emit       (tile, particle) pairs in particle order
//   sort         stable radix sort of the pairs by tile (rocPRIM): every
//                tile's list stays in particle (= draw) order
//   k_ranges     each tile's run of the sorted pairs
//   k_raster     one block per 8x8-pixel tile, one lane per pixel, over the
//                tile's list: the depth test (lowest depth, ties to the
//                earlier instance) and the additive r16float thickness blend
//                in draw order (the depth map and thickness map passes draw
//                the same quads, fluidRender.ts:424-448); in sphere mode the
//                winner's shaded colour
//   k_bilateral  4 x (along x, along y) on the depth map (spatial weights
//                from a table made once by k_filter_weights)
//   k_gaussian   1 x (along x, along y) on the thickness, r16float (weights
//                from the same table)
//   k_fluid      per pixel shading with the cubemap
// Fragment rules (the rasteriser's choices, fixed here and in the oracle):
// pixel centres inside [xl, xr) x [yt, yb) of the impostor's screen
// rectangle, uv linear across it (the quad is parallel to the image plane),
// frag_depth clamped to [0, 1], impostors outside the clip volume
// (clip z not in [0, w]: nearer than the near plane or beyond the far
// plane) not drawn.
// textureLoad coordinates are clamped into the image (Tint's robustness).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "render.h"
#include "div_rn.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string t_error;

int fail(int status, const std::string &msg) {
    t_error = msg;
    return status;
}

#define R_TRY(expr)                                                                      \
    do {                                                                                 \
        const hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? RENDER_ERR_OUT_OF_MEMORY             \
                                                  : RENDER_ERR_DEVICE,                   \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

constexpr int kThreads = 256;

struct Consts {
    int width, height;
    float depth_threshold, max_filter_size, projected_particle_constant;
};

// mat4x4f * vec4f, column by column: ((c0 x + c1 y) + c2 z) + c3 w
__device__ __forceinline__ void mv4(const float *m, const float v[4], float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__device__ __forceinline__ void normalize3(const float v[3], float out[3]) {
    const float l = sqrtf(dot3(v, v));
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = v[i] / l;
}

// textureLoad(tex, vec2u(coord), 0), coordinates clamped into the image
__device__ __forceinline__ int tex_index(float fx, float fy, int w, int h) {
    int ix = fx <= 0.0f ? 0 : (fx >= (float)w ? w - 1 : (int)fx);
    int iy = fy <= 0.0f ? 0 : (fy >= (float)h ? h - 1 : (int)fy);
    return iy * w + ix;
}

// exp and pow: WGSL bounds their error (exp: 3 + 2|x| ulp; pow as
// exp2(y log2 x)) without fixing their values; here one deterministic formula
// each, the same as the oracle's (oracle_render_exp / _pow5 / _pow250).
// exp(x): x = k ln2 + r, k = rint(x log2 e), ln2 split into 0.693359375 +
// (-2.12194440e-4); exp(r), |r| <= 0.35, by its degree-7 Taylor polynomial in
// Horner form (one fused multiply-add per step); times 2^k.  0 below -86 (results
// stay normal), inf above 88, NaN through; within 0.9 ulp of exp elsewhere.
__device__ __forceinline__ float det_exp(float x) {
    if (x != x) return x;
    if (!(x > -86.0f)) return 0.0f;
    if (x > 88.0f) return __builtin_inff();
    const float k = rintf(x * 1.44269502f);
    const float r = fmaf(k, 2.12194440e-4f, x - k * 0.693359375f);
    float p = 1.98412698e-4f;  // 1/7!
    p = fmaf(p, r, 1.38888889e-3f);
    p = fmaf(p, r, 8.33333333e-3f);
    p = fmaf(p, r, 4.16666667e-2f);
    p = fmaf(p, r, 1.66666667e-1f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

// log(x) for x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
// s = (m - 1) / (m + 1), by its series to s^9; e ln2 with ln2 split as for
// exp.  -inf at 0, NaN below.
__device__ __forceinline__ float det_log(float x) {
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e;
    float m = frexpf(x, &e);  // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f), s2 = s * s;
    float p = 1.11111111e-1f;  // 1/9
    p = fmaf(p, s2, 1.42857143e-1f);
    p = fmaf(p, s2, 2.00000000e-1f);
    p = fmaf(p, s2, 3.33333333e-1f);
    p = fmaf(p, s2, 1.0f);
    const float fe = (float)e;
    return fe * 0.693359375f + fmaf(2.0f * s, p, fe * -2.12194440e-4f);
}

// pow(x, 5) and pow(x, 250) for x >= 0 (fluid.wgsl:70,80): ((x x)(x x)) x and
// exp(250 log x), within WGSL's bound for pow (inherited from exp2(y log2 x)).
__device__ __forceinline__ float det_pow5(float x) { return ((x * x) * (x * x)) * x; }

__device__ __forceinline__ float det_pow250(float x) { return det_exp(250.0f * det_log(x)); }

// f32 -> r16float -> f32 (round to nearest even)
__device__ __forceinline__ float f16_round(float f) { return (float)(_Float16)f; }

// The canvas (bgra8unorm) stores a colour clamped to [0, 1], NaN as 0: the
// output is that value before its 8-bit quantisation.  NaN arises where both
// horizontal (or vertical) neighbours of a fluid pixel are background (depth
// 0 there gives an infinite view position, fluid.wgsl:22-27).
__device__ __forceinline__ float unorm(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// ---- impostors (depthMap.wgsl:34-79, thicknessMap.wgsl:25-61) ----

struct Impostor {
    float view[3];
    float xl, xr, yt, yb;
    int px0, px1, py0, py1;
};

__device__ __forceinline__ bool impostor(const render_uniforms &u, const float p[3], int w, int h,
                                         Impostor &im) {
    const float v4[4] = {p[0], p[1], p[2], 1.0f};
    float vp[4];
    mv4(u.view, v4, vp);
#pragma unroll
    for (int i = 0; i < 3; ++i) im.view[i] = vp[i];
    float fx[2], fy[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const float cx = (c ? 0.5f : -0.5f) * u.sphere_size;
        const float cy = (c ? 0.5f : -0.5f) * u.sphere_size;
        const float q[4] = {vp[0] + cx, vp[1] + cy, vp[2] + 0.0f, 1.0f};
        float clip[4];
        mv4(u.projection, q, clip);
        // the quad is parallel to the image plane: its corners share clip z and w,
        // so the clip volume (0 <= z <= w) keeps all of it or none
        if (!(clip[3] > 0.0f) || !(clip[2] >= 0.0f) || !(clip[2] <= clip[3])) return false;
        const float nx = clip[0] / clip[3], ny = clip[1] / clip[3];
        fx[c] = ((nx + 1.0f) * 0.5f) * (float)w;
        fy[c] = ((1.0f - ny) * 0.5f) * (float)h;
    }
    im.xl = fx[0];
    im.xr = fx[1];
    im.yt = fy[1];
    im.yb = fy[0];
    if (!(im.xr > im.xl) || !(im.yb > im.yt)) return false;
    double a = ceil((double)im.xl - 0.5), b = ceil((double)im.xr - 0.5);
    double c2 = ceil((double)im.yt - 0.5), d = ceil((double)im.yb - 0.5);
    if (a < 0) a = 0;
    if (c2 < 0) c2 = 0;
    if (b > w) b = w;
    if (d > h) d = h;
    if (a >= b || c2 >= d) return false;
    im.px0 = (int)a;
    im.px1 = (int)b;
    im.py0 = (int)c2;
    im.py1 = (int)d;
    return true;
}

__device__ __forceinline__ float frag_depth(const render_uniforms &u, const Impostor &im,
                                            const float nxy[2], float r2, float rvp[3],
                                            float normal[3]) {
    const float nz = sqrtf(1.0f - r2);
    normal[0] = nxy[0];
    normal[1] = nxy[1];
    normal[2] = nz;
    const float radius = u.sphere_size / 2.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) rvp[i] = im.view[i] + normal[i] * radius;
    const float q[4] = {rvp[0], rvp[1], rvp[2], 1.0f};
    float clip[4];
    mv4(u.projection, q, clip);
    float d = clip[2] / clip[3];
    d = d > 0.0f ? d : 0.0f;
    if (d > 1.0f) d = 1.0f;
    return d;
}

__device__ __forceinline__ void value_to_color(float value, float out[3]) {
    const float col[5][3] = {{0.0f, 0.4f, 0.8f},
                             {35.0f / 256.0f, 161.0f / 256.0f, 165.0f / 256.0f},
                             {95.0f / 256.0f, 254.0f / 256.0f, 150.0f / 256.0f},
                             {243.0f / 256.0f, 250.0f / 256.0f, 49.0f / 256.0f},
                             {255.0f / 256.0f, 165.0f / 256.0f, 0.0f / 256.0f}};
    int k;
    float t;
    if (0.0f <= value && value < 0.25f) {
        k = 0;
        t = value / 0.25f;
    } else if (0.25f <= value && value < 0.50f) {
        k = 1;
        t = (value - 0.25f) / 0.25f;
    } else if (0.50f <= value && value < 0.75f) {
        k = 2;
        t = (value - 0.50f) / 0.25f;
    } else {
        k = 3;
        t = (value - 0.75f) / 0.25f;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = col[k][i] * (1.0f - t) + col[k + 1][i] * t;
}

// ---- impostor raster ----
//
// The depth map and thickness map passes draw the same quads (depthMap.wgsl,
// thicknessMap.wgsl), so one raster serves both.  Screen tiles of kTile^2
// pixels; each drawn impostor is listed in every tile its pixel rectangle
// touches, the lists in particle order (a stable sort by tile of pairs
// emitted in particle order), and one thread per pixel walks its tile's list:
// the depth test keeps the first of the lowest depths, the thickness blends
// in list order, i.e. draw order, as the raster pipeline's blend unit does.

constexpr int kTile = 8;
constexpr int kTileShift = 3;
static_assert(kTile * kTile == 64, "a raster tile is one wave's lanes");

// Lower bound of the depth of any fragment of an impostor, less a margin, for
// a projection whose clip z and w depend on view z alone (the perspective
// matrix: entries 2, 3, 6, 7 zero, w = -z): a fragment's view z is at most
// view z + radius (its normal's z is at most 1), and the depth decreases with
// view z, up to a few ulps of rounding; the margin 2^-20 covers them.  With
// another projection the bound is -1 (it never skips).
__device__ __forceinline__ float depth_floor(const render_uniforms &u, const Impostor &im,
                                             int standard_projection) {
    if (!standard_projection) return -1.0f;
    const float nxy[2] = {0.0f, 0.0f};
    float rvp[3], nrm[3];
    return frag_depth(u, im, nxy, 0.0f, rvp, nrm) - 0x1p-20f;
}

struct TileRect {
    int tx0, ty0, nx;
    uint32_t count;
};

__device__ __forceinline__ TileRect tile_rect(const Impostor &im, bool drawn) {
    TileRect r;
    r.tx0 = im.px0 >> kTileShift;
    r.ty0 = im.py0 >> kTileShift;
    r.nx = ((im.px1 - 1) >> kTileShift) - r.tx0 + 1;
    r.count = drawn ? (uint32_t)r.nx * (uint32_t)(((im.py1 - 1) >> kTileShift) - r.ty0 + 1) : 0u;
    return r;
}

// Impostor record of a particle: its screen rectangle, its view position and
// its depth floor, and the number of tiles its pixel rectangle touches
// (count[n] = 0: the scan's total lands in offset[n]).
__global__ __launch_bounds__(kThreads) void k_project(const float4 *__restrict__ posvel, uint32_t n,
                                                      render_uniforms u, int w, int h,
                                                      int standard_projection,
                                                      float4 *__restrict__ rec,
                                                      uint32_t *__restrict__ count) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i > n) return;
    Impostor im{};
    bool drawn = false;
    if (i < n) {
        const float4 pv = posvel[2 * (size_t)i];
        const float p[3] = {pv.x, pv.y, pv.z};
        drawn = impostor(u, p, w, h, im);
        if (drawn) {
            rec[2 * (size_t)i] = make_float4(im.xl, im.xr, im.yt, im.yb);
            rec[2 * (size_t)i + 1] = make_float4(im.view[0], im.view[1], im.view[2],
                                                 depth_floor(u, im, standard_projection));
        }
    }
    count[i] = tile_rect(im, drawn).count;
}

// (tile, particle) pairs of every drawn impostor at its offset, tiles in
// row-major order: the pair array is in particle order.
__global__ __launch_bounds__(kThreads) void k_emit(const float4 *__restrict__ posvel, uint32_t n,
                                                   render_uniforms u, int w, int h, int tiles_x,
                                                   const uint64_t *__restrict__ offset,
                                                   uint32_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t o = offset[i];
    if (offset[i + 1] == o) return;
    const float4 pv = posvel[2 * (size_t)i];
    const float p[3] = {pv.x, pv.y, pv.z};
    Impostor im{};
    impostor(u, p, w, h, im);
    for (int ty = im.py0 >> kTileShift; ty <= (im.py1 - 1) >> kTileShift; ++ty)
        for (int tx = im.px0 >> kTileShift; tx <= (im.px1 - 1) >> kTileShift; ++tx) {
            keys[o] = (uint32_t)(ty * tiles_x + tx);
            vals[o] = i;
            ++o;
        }
}

__global__ __launch_bounds__(kThreads) void k_clear_ranges(uint2 *__restrict__ range, uint32_t ntiles) {
    for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < ntiles; t += gridDim.x * kThreads)
        range[t] = make_uint2(0u, 0u);
}

// [begin, end) of every tile's run in the sorted pairs
__global__ __launch_bounds__(kThreads) void k_ranges(const uint32_t *__restrict__ keys,
                                                     uint32_t total, uint2 *__restrict__ range) {
    for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < total; q += gridDim.x * kThreads) {
        const uint32_t k = keys[q];
        if (q == 0 || keys[q - 1] != k) range[k].x = q;
        if (q + 1 == total || keys[q + 1] != k) range[k].y = q + 1;
    }
}

// Raster blocks: kRasterWaves waves per tile, every lane of a wave one pixel.
// A tile's list (in particle order = draw order) is walked in chunks of kChunk
// entries: each wave computes the fragments of its kChunkPer entries
// (thickness contributions into LDS, 0 where an entry has no fragment, and its
// own depth winner per pixel); then wave 0 blends the chunk's contributions in
// list order.  The waves' winners merge at the end as (depth, particle index)
// keys: the lowest depth, ties to the earlier instance, whichever wave saw it.
constexpr int kRasterWaves = 8;
constexpr int kRasterThreads = 64 * kRasterWaves;
constexpr int kChunkPer = 32;
constexpr int kChunk = kRasterWaves * kChunkPer;
constexpr int kCStride = kChunk + 1;  // odd row stride: conflict-free in both phases
static_assert(kChunk <= kRasterThreads - 64, "waves 1.. stage a whole chunk during the blend");

// The rectangle's extents and their rounded reciprocals, (xr - xl, 1 / that,
// yb - yt, 1 / that), for div_rn.
__device__ __forceinline__ float4 rect_extents(const float4 rc) {
    const float bx = rc.y - rc.x, by = rc.w - rc.z;
    return make_float4(bx, 1.0f / bx, by, 1.0f / by);
}

// Fragment of record (rect, view) at pixel centre (cx, cy): 0 without a
// fragment, else its thickness contribution; d, z, normal when the depth test
// may pass (floor not above bound), else d = 2.  uv = ((cx - xl) / (xr - xl),
// (yb - cy) / (yb - yt)), the quotients correctly rounded (div_rn from the
// extents' reciprocals); n.xy = 2 uv - 1; no fragment where |n.xy|^2 > 1.
__device__ __forceinline__ float raster_fragment(const render_uniforms &u, const float4 rc,
                                                 const float4 ext, const float4 vw, float cx,
                                                 float cy, float bound, float &d, float &z,
                                                 float nrm[3]) {
    d = 2.0f;
    if (!(cx >= rc.x && cx < rc.y && cy >= rc.z && cy < rc.w)) return 0.0f;
    const float uvx = div_rn(cx - rc.x, ext.x, ext.y);
    const float uvy = div_rn(rc.w - cy, ext.z, ext.w);
    float nxy[2];
    nxy[0] = uvx * 2.0f - 1.0f;
    nxy[1] = uvy * 2.0f - 1.0f;
    const float r2 = nxy[0] * nxy[0] + nxy[1] * nxy[1];
    if (r2 > 1.0f) return 0.0f;
    const float nz = sqrtf(1.0f - r2);
    Impostor im;
    if (!(vw.w > bound)) {
        im.view[0] = vw.x;
        im.view[1] = vw.y;
        im.view[2] = vw.z;
        float rvp[3];
        d = frag_depth(u, im, nxy, r2, rvp, nrm);
        z = rvp[2];
    }
    return 0.05f * nz;
}

// One block per tile: the depth map (the winner's view z; 0 = no fragment) or
// the sphere pass's colour, and the thickness map.
__global__ __launch_bounds__(kRasterThreads) void k_raster(
    const float4 *__restrict__ posvel, const float4 *__restrict__ rec,
    const uint32_t *__restrict__ vals, const uint2 *__restrict__ range, render_uniforms u, int w,
    int h, int tiles_x, float *__restrict__ depth, float *__restrict__ thick16,
    float4 *__restrict__ color, int sphere_mode) {
    __shared__ float s_c[64 * kCStride];
    __shared__ float4 s_rect[kChunk], s_ext[kChunk], s_view[kChunk];
    __shared__ uint32_t s_idx[kChunk];
    __shared__ unsigned long long s_key[64];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int tx = (int)(blockIdx.x % (uint32_t)tiles_x), ty = (int)(blockIdx.x / (uint32_t)tiles_x);
    const int px = tx * kTile + (lane & (kTile - 1));
    const int py = ty * kTile + (lane >> kTileShift);
    const bool on = px < w && py < h;
    const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    const uint2 rg = range[blockIdx.x];
    if (wave == 0) s_key[lane] = ~0ull;
    // best_d, best_i: this wave's winner (its entries in list order, so the
    // strict test keeps the first of equal depths); bound: the lowest depth
    // any wave has found, for the depth floor's skip only (a skipped
    // fragment is deeper than a found one: it can neither win nor tie)
    float best_d = 2.0f, bound = 2.0f, t = 0.0f;
    uint32_t best_i = 0xffffffffu;
    // Chunk records are staged by the waves that are idle while wave 0 blends:
    // the first chunk before the loop, each next one during the blend of the
    // current one (phase 2 reads only s_c).
    auto stage = [&](uint32_t base, int tid) {
        const int m = rg.y - base < (uint32_t)kChunk ? (int)(rg.y - base) : kChunk;
        if (tid >= 0 && tid < m) {
            const uint32_t i = vals[base + tid];
            const float4 rc = rec[2 * (size_t)i];
            s_idx[tid] = i;
            s_rect[tid] = rc;
            s_ext[tid] = rect_extents(rc);
            s_view[tid] = rec[2 * (size_t)i + 1];
        }
    };
    if (rg.x < rg.y) stage(rg.x, (int)threadIdx.x);
    for (uint32_t base = rg.x; base < rg.y; base += kChunk) {
        const int m = rg.y - base < (uint32_t)kChunk ? (int)(rg.y - base) : kChunk;
        __syncthreads();  // this chunk's records are staged; the last blend is done with s_c
        const int k1 = (wave + 1) * kChunkPer < m ? (wave + 1) * kChunkPer : m;
        for (int k = wave * kChunkPer; k < k1; ++k) {
            float d, z, nrm[3];
            const float c = raster_fragment(u, s_rect[k], s_ext[k], s_view[k], cx, cy,
                                            fminf(best_d, bound), d, z, nrm);
            if (d < best_d) {
                best_d = d;
                best_i = s_idx[k];
            }
            s_c[lane * kCStride + k] = c;
        }
        if (best_i != 0xffffffffu)
            atomicMin(&s_key[lane], ((unsigned long long)__float_as_uint(best_d) << 32) | best_i);
        __syncthreads();  // s_c complete; the records are free for the next chunk
        {
            const unsigned long long kb = s_key[lane];
            if (kb != ~0ull) bound = fminf(bound, __uint_as_float((uint32_t)(kb >> 32)));
        }
        if (wave == 0) {
            if (!sphere_mode) {
                // the thickness target's blends, one fragment at a time in list
                // order (adding 0 leaves an f16 value unchanged)
                const float *row = s_c + lane * kCStride;
                for (int k = 0; k < m; ++k) t = f16_round(t + row[k]);
            }
        } else if (base + kChunk < rg.y) {
            stage(base + kChunk, (int)threadIdx.x - 64);
        }
    }
    if (best_i != 0xffffffffu)
        atomicMin(&s_key[lane], ((unsigned long long)__float_as_uint(best_d) << 32) | best_i);
    __syncthreads();
    if (wave != 0 || !on) return;
    const size_t pi = (size_t)py * w + px;
    const unsigned long long key = s_key[lane];
    float z = 0.0f, nrm[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t wi = (uint32_t)(key & 0xffffffffu);
    if (key != ~0ull) {
        // the winner's fragment again, for its view z and normal
        const float4 rc = rec[2 * (size_t)wi];
        float d;
        raster_fragment(u, rc, rect_extents(rc), rec[2 * (size_t)wi + 1], cx, cy, 2.0f, d, z,
                        nrm);
    }
    depth[pi] = z;
    thick16[pi] = sphere_mode ? 0.0f : t;
    if (sphere_mode) {
        float4 c = make_float4(0.8f, 0.8f, 0.8f, 1.0f);
        if (key != ~0ull) {
            // sphere.wgsl: colour by speed, Lambert term towards normalize(1, 1, 1)
            const float4 vv = posvel[2 * (size_t)wi + 1];
            const float v[3] = {vv.x, vv.y, vv.z};
            const float speed = sqrtf(dot3(v, v));
            const float one[3] = {1.0f, 1.0f, 1.0f};
            float ln[3], col[3];
            normalize3(one, ln);
            const float dd = dot3(nrm, ln);
            const float diffuse = dd > 0.0f ? dd : 0.0f;
            value_to_color(speed / 1.5f, col);
            c = make_float4(unorm(col[0] * diffuse), unorm(col[1] * diffuse),
                            unorm(col[2] * diffuse), 1.0f);
        }
        color[pi] = c;
    }
}

// Per-pixel kernels: 2D blocks of kPixW x kPixH pixels (x, y straight from the
// block and thread indices).
constexpr int kPixW = 64, kPixH = 4;
static_assert(kPixW * kPixH == kThreads, "per-pixel blocks");

// The filters' data-independent weights, evaluated once with the same
// expressions (and so the same bits) as inline: the bilateral's spatial weight
// exp(-rr / two_sigma) per filter size (1..100) and tap distance, the
// gaussian's 61 weights (sigma 10) and their sum in tap order.
constexpr int kMaxFilter = 100;
constexpr int kBilateralTable = (kMaxFilter + 1) * (kMaxFilter + 1);
constexpr int kGaussHalf = 30;
constexpr int kGaussTable = 2 * kGaussHalf + 2;

__global__ __launch_bounds__(kThreads) void k_filter_weights(float *__restrict__ bw,
                                                             float *__restrict__ gw) {
    for (int q = (int)threadIdx.x; q < kBilateralTable; q += kThreads) {
        const int fs = q / (kMaxFilter + 1), k = q % (kMaxFilter + 1);
        const float sigma = (float)fs / 3.0f;
        const float two_sigma = 2.0f * sigma * sigma;
        const float co = (float)k;
        const float rr = co * co + co * co;
        bw[q] = fs ? det_exp(-rr / two_sigma) : 0.0f;
    }
    const float sigma = (float)kGaussHalf / 3.0f;
    const float two_sigma = 2.0f * sigma * sigma;
    for (int k = (int)threadIdx.x; k <= 2 * kGaussHalf; k += kThreads) {
        const float co = (float)(k - kGaussHalf);
        gw[k] = det_exp(-co * co / two_sigma);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float wsum = 0.0f;
        for (int k = 0; k <= 2 * kGaussHalf; ++k) wsum += gw[k];
        gw[2 * kGaussHalf + 1] = wsum;
    }
}

// Largest filter size whose taps are read from LDS: a block stages its rows
// (along x) or columns (along y) with an apron of its own largest filter
// size; blocks with a larger one gather from global memory.
constexpr int kApron = 40;

// One pixel of bilateral.wgsl:20-58: taps k = -fs..fs, sample(k) the
// neighbour's |depth| (clamped into the image).
template <typename Sample>
__device__ __forceinline__ float bilateral_pixel(float depth, int fs, const float *wrow,
                                                 float two_sigma_depth,
                                                 float inv_two_sigma_depth, Sample sample) {
    float sum = 0.0f, wsum = 0.0f;
#pragma unroll 4
    for (int k = -fs; k <= fs; ++k) {
        const float sampled = sample(k);
        const float wt = wrow[k < 0 ? -k : k];
        const float r_depth = sampled - depth;
        const float wd =
            det_exp(div_rn(-r_depth * r_depth, two_sigma_depth, inv_two_sigma_depth));
        sum += sampled * wt * wd;
        wsum += wt * wd;
    }
    return sum / wsum;
}

// bilateral.wgsl:20-58 along x (kAlongX) or y.  Tap k of pixel (x, y) reads
// |src| at column clamp(x + k) (along x) or row clamp(y + k) (along y), which
// is textureLoad's clamped coordinate of (x + 0.5 + k, y + 0.5).  A block
// is one 64-pixel row segment per wave, 4 rows; every wave's lanes run the
// loop of the lane with the largest filter, so the taps' latency is what
// bounds the pass: staged through LDS, the block's loads are issued at once.
template <bool kAlongX>
__global__ __launch_bounds__(kThreads) void k_bilateral(const float *__restrict__ src,
                                                        float *__restrict__ dst, Consts c,
                                                        const float *__restrict__ bw) {
    constexpr int kRowS = kPixW + 2 * kApron;  // along x: LDS row stride
    constexpr int kS = kAlongX ? kPixH * kRowS : (kPixH + 2 * kApron) * kPixW;
    __shared__ float s[kS];
    __shared__ int s_wmax[kPixH];
    const int w = c.width, h = c.height;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int x0 = (int)blockIdx.x * kPixW, y0 = (int)blockIdx.y * kPixH;
    const int x = x0 + tx, y = y0 + ty;
    const bool inb = x < w && y < h;
    const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
    const float depth = inb ? fabsf(src[tex_index(iu, iv, w, h)]) : 0.0f;
    const bool active = inb && !(depth >= 1e4f || depth <= 0.0f);
    int fs = 0;
    if (active) {
        const float q = ceilf(c.projected_particle_constant / depth);
        const int fsz = q >= 2147483648.0f ? 2147483647 : (int)q;
        const int mfs = (int)c.max_filter_size;
        fs = mfs < fsz ? mfs : fsz;
    }
    // the block's largest filter size + 1 (0: no pixel to filter); wave ty
    // is row ty of the block
    int m = active ? fs + 1 : 0;
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if (tx == 0) s_wmax[ty] = m;
    __syncthreads();
    int g = s_wmax[0];
    for (int i = 1; i < kPixH; ++i) g = s_wmax[i] > g ? s_wmax[i] : g;
    if (g == 0) {  // (block-uniform) background only
        if (inb) dst[(size_t)y * w + x] = depth;
        return;
    }
    const int F = g - 1;
    const float sigma_depth = c.depth_threshold / 3.0f;
    const float two_sigma_depth = 2.0f * sigma_depth * sigma_depth;
    // every tap divides by two_sigma_depth: the correctly rounded quotient
    // from its rounded reciprocal (div_rn; the numerators are 0 or squares
    // of depth differences, >= ~1e-14)
    const float inv_two_sigma_depth = 1.0f / two_sigma_depth;
    const float *wrow = bw + fs * (kMaxFilter + 1);
    if (F > kApron) {  // (block-uniform) gather from global memory
        if (!active) {
            if (inb) dst[(size_t)y * w + x] = depth;
            return;
        }
        dst[(size_t)y * w + x] = bilateral_pixel(
            depth, fs, wrow, two_sigma_depth, inv_two_sigma_depth, [&](int k) {
                const float co = (float)k;
                return kAlongX ? fabsf(src[tex_index(iu + co, iv, w, h)])
                               : fabsf(src[tex_index(iu, iv + co, w, h)]);
            });
        return;
    }
    if (kAlongX) {  // row y0 + ty (clamped), columns x0 - F .. x0 + 63 + F (clamped)
        const float *srow = src + (size_t)(y < h ? y : h - 1) * w;
        for (int j = tx; j < kPixW + 2 * F; j += kPixW) {
            const int cx = x0 - F + j;
            s[ty * kRowS + j] = fabsf(srow[cx < 0 ? 0 : (cx >= w ? w - 1 : cx)]);
        }
    } else {  // rows y0 - F .. y0 + 3 + F (clamped), column x (clamped)
        const int cx = x < w ? x : w - 1;
        for (int j = ty; j < kPixH + 2 * F; j += kPixH) {
            const int cy = y0 - F + j;
            s[j * kPixW + tx] = fabsf(src[(size_t)(cy < 0 ? 0 : (cy >= h ? h - 1 : cy)) * w + cx]);
        }
    }
    __syncthreads();
    if (!active) {
        if (inb) dst[(size_t)y * w + x] = depth;
        return;
    }
    const float *centre = kAlongX ? s + ty * kRowS + tx + F : s + (ty + F) * kPixW + tx;
    constexpr int kStep = kAlongX ? 1 : kPixW;
    dst[(size_t)y * w + x] =
        bilateral_pixel(depth, fs, wrow, two_sigma_depth, inv_two_sigma_depth,
                        [&](int k) { return centre[k * kStep]; });
}

// gaussian.wgsl:15-42 along x (kAlongX) or y (r16float target).  Tap k of
// pixel (x, y) reads src at column clamp(x + k) or row clamp(y + k)
// (textureLoad's clamped coordinate of (x + 0.5 + k, y + 0.5)); a block with
// any nonzero pixel stages its rows or columns with the 30-pixel apron in
// LDS, read by the fixed 61-tap loop.
template <bool kAlongX>
__global__ __launch_bounds__(kThreads) void k_gaussian(const float *__restrict__ src,
                                                       float *__restrict__ dst, int w, int h,
                                                       const float *__restrict__ gw) {
    constexpr int kRowS = kPixW + 2 * kGaussHalf;  // along x: LDS row stride
    constexpr int kS = kAlongX ? kPixH * kRowS : (kPixH + 2 * kGaussHalf) * kPixW;
    __shared__ float s[kS];
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int x0 = (int)blockIdx.x * kPixW, y0 = (int)blockIdx.y * kPixH;
    const int x = x0 + tx, y = y0 + ty;
    const bool inb = x < w && y < h;
    const float t = inb ? src[tex_index((float)x + 0.5f, (float)y + 0.5f, w, h)] : 0.0f;
    const bool active = inb && t != 0.0f;
    if (!__syncthreads_or(active)) {  // (block-uniform) nothing to blur
        if (inb) dst[(size_t)y * w + x] = 0.0f;
        return;
    }
    if (kAlongX) {  // row y (clamped), columns x0 - 30 .. x0 + 93 (clamped)
        const float *srow = src + (size_t)(y < h ? y : h - 1) * w;
        for (int j = tx; j < kRowS; j += kPixW) {
            const int cx = x0 - kGaussHalf + j;
            s[ty * kRowS + j] = srow[cx < 0 ? 0 : (cx >= w ? w - 1 : cx)];
        }
    } else {  // rows y0 - 30 .. y0 + 33 (clamped), column x (clamped)
        const int cx = x < w ? x : w - 1;
        for (int j = ty; j < kPixH + 2 * kGaussHalf; j += kPixH) {
            const int cy = y0 - kGaussHalf + j;
            s[j * kPixW + tx] = src[(size_t)(cy < 0 ? 0 : (cy >= h ? h - 1 : cy)) * w + cx];
        }
    }
    __syncthreads();
    if (!active) {
        if (inb) dst[(size_t)y * w + x] = 0.0f;
        return;
    }
    const float *centre =
        kAlongX ? s + ty * kRowS + tx + kGaussHalf : s + (ty + kGaussHalf) * kPixW + tx;
    constexpr int kStep = kAlongX ? 1 : kPixW;
    const float wsum = gw[2 * kGaussHalf + 1];
    float sum = 0.0f;
#pragma unroll
    for (int k = -kGaussHalf; k <= kGaussHalf; ++k) sum += centre[k * kStep] * gw[k + kGaussHalf];
    dst[(size_t)y * w + x] = f16_round(sum / wsum);
}

// textureSampleLevel(envmap, linear sampler, dir, 0): the major-axis face,
// bilinear, clamp to edge
__device__ __forceinline__ void cube_sample(const float4 *__restrict__ cube, int size,
                                            const float d[3], float rgb[3]) {
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        face = d[0] >= 0.0f ? 0 : 1;
        ma = ax;
        sc = d[0] >= 0.0f ? -d[2] : d[2];
        tc = -d[1];
    } else if (ay >= az) {
        face = d[1] >= 0.0f ? 2 : 3;
        ma = ay;
        sc = d[0];
        tc = d[1] >= 0.0f ? d[2] : -d[2];
    } else {
        face = d[2] >= 0.0f ? 4 : 5;
        ma = az;
        sc = d[2] >= 0.0f ? d[0] : -d[0];
        tc = -d[1];
    }
    const float s = (sc / ma + 1.0f) * 0.5f, t = (tc / ma + 1.0f) * 0.5f;
    const float u = s * (float)size - 0.5f, v = t * (float)size - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    int i0 = (int)fu, j0 = (int)fv, i1 = i0 + 1, j1 = j0 + 1;
    i0 = min(max(i0, 0), size - 1);
    j0 = min(max(j0, 0), size - 1);
    i1 = min(max(i1, 0), size - 1);
    j1 = min(max(j1, 0), size - 1);
    const float4 *fb = cube + (size_t)face * size * size;
    const float4 t00 = fb[(size_t)j0 * size + i0], t10 = fb[(size_t)j0 * size + i1];
    const float4 t01 = fb[(size_t)j1 * size + i0], t11 = fb[(size_t)j1 * size + i1];
    const float c00[3] = {t00.x, t00.y, t00.z}, c10[3] = {t10.x, t10.y, t10.z};
    const float c01[3] = {t01.x, t01.y, t01.z}, c11[3] = {t11.x, t11.y, t11.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float top = c00[k] * (1.0f - a) + c10[k] * a, bot = c01[k] * (1.0f - a) + c11[k] * a;
        rgb[k] = top * (1.0f - b) + bot * b;
    }
}

__device__ __forceinline__ void view_pos_from_uv_depth(const render_uniforms &u, float tx,
                                                       float ty, float depth, float out[3]) {
    float ndc[4] = {tx * 2.0f - 1.0f, 1.0f - 2.0f * ty, 0.0f, 1.0f};
    ndc[2] = -u.projection[2 * 4 + 2] + u.projection[3 * 4 + 2] / depth;
    ndc[3] = 1.0f;
    float eye[4];
    mv4(u.inv_projection, ndc, eye);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = eye[i] / eye[3];
}

// fluid.wgsl:38-84
__global__ __launch_bounds__(kThreads) void k_fluid(const float *__restrict__ depth_map,
                                                    const float *__restrict__ thick,
                                                    const float4 *__restrict__ cube, int cube_size,
                                                    render_uniforms u, int w, int h,
                                                    float4 *__restrict__ color) {
    const int x = (int)(blockIdx.x * kPixW + threadIdx.x);
    const int y = (int)(blockIdx.y * kPixH + threadIdx.y);
    if (x >= w || y >= h) return;
    const size_t pi = (size_t)y * w + x;
    const float bg[3] = {0.8f, 0.8f, 0.8f};
    const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
    const float uvx = iu / (float)w, uvy = iv / (float)h;
    const float depth = fabsf(depth_map[tex_index(iu, iv, w, h)]);
    if (depth >= 1e4f || depth <= 0.0f) {
        color[pi] = make_float4(bg[0], bg[1], bg[2], 1.0f);
        return;
    }
    float vp[3];
    view_pos_from_uv_depth(u, uvx, uvy, depth, vp);
    const float tsx = u.texel_size[0], tsy = u.texel_size[1];
    float p[3], ddx[3], ddy[3], ddx2[3], ddy2[3];
    view_pos_from_uv_depth(u, uvx + tsx, uvy + 0.0f,
                           fabsf(depth_map[tex_index(iu + 1.0f, iv + 0.0f, w, h)]), p);
#pragma unroll
    for (int i = 0; i < 3; ++i) ddx[i] = p[i] - vp[i];
    view_pos_from_uv_depth(u, uvx + 0.0f, uvy + tsy,
                           fabsf(depth_map[tex_index(iu + 0.0f, iv + 1.0f, w, h)]), p);
#pragma unroll
    for (int i = 0; i < 3; ++i) ddy[i] = p[i] - vp[i];
    view_pos_from_uv_depth(u, uvx + -tsx, uvy + 0.0f,
                           fabsf(depth_map[tex_index(iu + -1.0f, iv + 0.0f, w, h)]), p);
#pragma unroll
    for (int i = 0; i < 3; ++i) ddx2[i] = vp[i] - p[i];
    view_pos_from_uv_depth(u, uvx + 0.0f, uvy + -tsy,
                           fabsf(depth_map[tex_index(iu + 0.0f, iv + -1.0f, w, h)]), p);
#pragma unroll
    for (int i = 0; i < 3; ++i) ddy2[i] = vp[i] - p[i];
    if (fabsf(ddx[2]) > fabsf(ddx2[2]))
        for (int i = 0; i < 3; ++i) ddx[i] = ddx2[i];
    if (fabsf(ddy[2]) > fabsf(ddy2[2]))
        for (int i = 0; i < 3; ++i) ddy[i] = ddy2[i];
    const float cr[3] = {ddx[1] * ddy[2] - ddx[2] * ddy[1], ddx[2] * ddy[0] - ddx[0] * ddy[2],
                         ddx[0] * ddy[1] - ddx[1] * ddy[0]};
    float normal[3];
    normalize3(cr, normal);
#pragma unroll
    for (int i = 0; i < 3; ++i) normal[i] = -normal[i];
    float ray[3];
    normalize3(vp, ray);
    const float ld4[4] = {0.0f, 0.0f, -1.0f, 0.0f};
    float l4[4], light[3];
    mv4(u.view, ld4, l4);
    normalize3(l4, light);
    const float hv[3] = {light[0] - ray[0], light[1] - ray[1], light[2] - ray[2]};
    float H[3];
    normalize3(hv, H);
    const float hn = dot3(H, normal);
    const float specular = det_pow250(hn > 0.0f ? hn : 0.0f);
    const float density = 1.5f;
    const float thickness = thick[tex_index(iu, iv, w, h)];
    const float dc[3] = {0.085f, 0.6375f, 0.9f};
    float refr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) refr[i] = bg[i] * det_exp(-density * thickness * (1.0f - dc[i]));
    const float F0 = 0.02f;
    const float nr[3] = {-ray[0], -ray[1], -ray[2]};
    float fr = F0 + (1.0f - F0) * det_pow5(1.0f - dot3(normal, nr));
    fr = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
    const float dn = dot3(normal, ray);
    const float refl[4] = {ray[0] - 2.0f * dn * normal[0], ray[1] - 2.0f * dn * normal[1],
                           ray[2] - 2.0f * dn * normal[2], 0.0f};
    float rw[4];
    mv4(u.inv_view, refl, rw);
    float rc[3];
    cube_sample(cube, cube_size, rw, rc);
    color[pi] = make_float4(unorm(1.0f * specular + (refr[0] * (1.0f - fr) + rc[0] * fr)),
                            unorm(1.0f * specular + (refr[1] * (1.0f - fr) + rc[1] * fr)),
                            unorm(1.0f * specular + (refr[2] * (1.0f - fr) + rc[2] * fr)),
                            1.0f);
}

unsigned blocks_for(size_t count) {
    const size_t b = (count + kThreads - 1) / kThreads;
    return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

struct render_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Consts c{};
    int cube_size = 0;
    int tiles_x = 0, tiles_y = 0;
    int key_bits = 1;
    float4 *cube = nullptr;
    float *depth = nullptr, *fdepth = nullptr, *tmp = nullptr;
    float *thick16 = nullptr, *fthick = nullptr;
    float4 *color = nullptr;
    float *bilateral_w = nullptr;  // spatial weights per filter size and tap
    float *gauss_w = nullptr;      // the gaussian's 61 weights and their sum
    uint2 *range = nullptr;        // per tile [begin, end) in the sorted pairs
    uint64_t *total = nullptr;     // pinned host word: the frame's pair count
    // per particle (grown on demand)
    float4 *rec = nullptr;         // 2 per particle
    uint32_t *count = nullptr;     // n + 1
    uint64_t *offset = nullptr;    // n + 1
    uint32_t part_cap = 0;
    // (tile, particle) pairs (grown on demand)
    uint32_t *keys = nullptr, *vals = nullptr, *keys_sorted = nullptr, *vals_sorted = nullptr;
    size_t pair_cap = 0;
    void *scratch = nullptr;       // rocPRIM temporary storage
    size_t scratch_bytes = 0;
    float4 *staging = nullptr;     // host PosVel records (render_execute_host)
    uint32_t staging_cap = 0;
};

namespace {

// Upper bound of a frame's (tile, particle) pairs: 2^31 (16 bytes each).
constexpr uint64_t kMaxPairs = 1ull << 31;

// The radix sort always takes rocPRIM's onesweep path (no merge sort below a
// size limit): two 8-bit passes for up to 65536 tiles.
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

void release(render_ctx *r) {
    void *ptrs[] = {r->cube, r->depth, r->fdepth, r->tmp, r->thick16, r->fthick, r->color,
                    r->bilateral_w, r->gauss_w, r->range, r->rec, r->count, r->offset, r->keys,
                    r->vals, r->keys_sorted, r->vals_sorted, r->scratch, r->staging};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (r->total) (void)hipHostFree(r->total);
    if (r->stream) (void)hipStreamDestroy(r->stream);
}

// Grows a device buffer to `bytes` (contents not kept).
int grow(void **p, size_t bytes) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    R_TRY(hipMalloc(p, bytes));
    return RENDER_OK;
}

#define R_CHECK(expr)                  \
    do {                               \
        const int rc_ = (expr);        \
        if (rc_ != RENDER_OK) return rc_; \
    } while (0)

int ensure_scratch(render_ctx *r, size_t bytes) {
    if (bytes <= r->scratch_bytes) return RENDER_OK;
    R_TRY(hipStreamSynchronize(r->stream));  // the previous frame may still use it
    r->scratch_bytes = 0;
    R_CHECK(grow(&r->scratch, bytes));
    r->scratch_bytes = bytes;
    return RENDER_OK;
}

int frame(render_ctx *r, const float4 *pv, uint32_t n, const render_uniforms &u, int sphere) {
    const int w = r->c.width, h = r->c.height;
    const size_t np = (size_t)w * h;
    const dim3 pgrid((unsigned)((w + kPixW - 1) / kPixW), (unsigned)((h + kPixH - 1) / kPixH));
    const dim3 pblock(kPixW, kPixH);
    const uint32_t ntiles = (uint32_t)(r->tiles_x * r->tiles_y);
    hipStream_t st = r->stream;
    if (n + 1 > r->part_cap) {
        R_TRY(hipStreamSynchronize(st));
        r->part_cap = 0;
        R_CHECK(grow((void **)&r->rec, sizeof(float4) * 2 * ((size_t)n + 1)));
        R_CHECK(grow((void **)&r->count, sizeof(uint32_t) * ((size_t)n + 1)));
        R_CHECK(grow((void **)&r->offset, sizeof(uint64_t) * ((size_t)n + 1)));
        r->part_cap = n + 1;
    }
    // The perspective form for which the depth floor holds (column-major):
    // clip z and w do not depend on view x, y; w = -z exactly; depth
    // -P10 - P14 / z decreases with z (P14 < 0); |P10| <= 2 bounds the rounding
    // of each depth evaluation by 2.4e-7, within the floor's margin 2^-20.
    const float *P = u.projection;
    const int standard = P[2] == 0.0f && P[6] == 0.0f && P[3] == 0.0f && P[7] == 0.0f &&
                         P[11] == -1.0f && P[15] == 0.0f && P[14] < 0.0f &&
                         fabsf(P[10]) <= 2.0f;
    // impostors, their tile counts and the pairs' offsets
    k_project<<<n / kThreads + 1, kThreads, 0, st>>>(pv, n, u, w, h, standard, r->rec, r->count);
    size_t bytes = 0;
    R_TRY(rocprim::exclusive_scan(nullptr, bytes, r->count, r->offset, (uint64_t)0, (size_t)n + 1,
                                  rocprim::plus<uint64_t>(), st));
    R_CHECK(ensure_scratch(r, bytes));
    R_TRY(rocprim::exclusive_scan(r->scratch, bytes, r->count, r->offset, (uint64_t)0,
                                  (size_t)n + 1, rocprim::plus<uint64_t>(), st));
    R_TRY(hipMemcpyAsync(r->total, r->offset + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    R_TRY(hipStreamSynchronize(st));
    const uint64_t total = *r->total;
    if (total > kMaxPairs)
        return fail(RENDER_ERR_OUT_OF_MEMORY,
                    "the impostors cover more than 2^31 tile slots (camera too close to the particles)");
    if (total > r->pair_cap) {
        const size_t cap = (size_t)total + total / 4 + 4096;
        r->pair_cap = 0;
        R_CHECK(grow((void **)&r->keys, 4 * cap));
        R_CHECK(grow((void **)&r->vals, 4 * cap));
        R_CHECK(grow((void **)&r->keys_sorted, 4 * cap));
        R_CHECK(grow((void **)&r->vals_sorted, 4 * cap));
        r->pair_cap = cap;
    }
    k_clear_ranges<<<blocks_for(ntiles), kThreads, 0, st>>>(r->range, ntiles);
    if (total) {
        k_emit<<<(n + kThreads - 1) / kThreads, kThreads, 0, st>>>(pv, n, u, w, h, r->tiles_x,
                                                                   r->offset, r->keys, r->vals);
        bytes = 0;
        R_TRY(rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, r->keys, r->keys_sorted,
                                                    r->vals, r->vals_sorted, (size_t)total, 0,
                                                    r->key_bits, st));
        R_CHECK(ensure_scratch(r, bytes));
        R_TRY(rocprim::radix_sort_pairs<SortConfig>(r->scratch, bytes, r->keys, r->keys_sorted,
                                                    r->vals, r->vals_sorted, (size_t)total, 0,
                                                    r->key_bits, st));
        k_ranges<<<blocks_for(total), kThreads, 0, st>>>(r->keys_sorted, (uint32_t)total, r->range);
    }
    k_raster<<<ntiles, kRasterThreads, 0, st>>>(pv, r->rec, r->vals_sorted, r->range, u, w, h,
                                               r->tiles_x, r->depth, r->thick16, r->color, sphere);
    if (!sphere) {
        // 4 x (bilateral x, bilateral y) (fluidRender.ts:432-444)
        const float *src = r->depth;
        for (int it = 0; it < 4; ++it) {
            k_bilateral<true><<<pgrid, pblock, 0, st>>>(src, r->tmp, r->c, r->bilateral_w);
            k_bilateral<false><<<pgrid, pblock, 0, st>>>(r->tmp, r->fdepth, r->c, r->bilateral_w);
            src = r->fdepth;
        }
        // 1 x (gaussian x, gaussian y) (fluidRender.ts:450-462)
        k_gaussian<true><<<pgrid, pblock, 0, st>>>(r->thick16, r->tmp, w, h, r->gauss_w);
        k_gaussian<false><<<pgrid, pblock, 0, st>>>(r->tmp, r->fthick, w, h, r->gauss_w);
        k_fluid<<<pgrid, pblock, 0, st>>>(r->fdepth, r->fthick, r->cube, r->cube_size, u, w, h,
                                         r->color);
    } else {
        R_TRY(hipMemcpyAsync(r->fdepth, r->depth, np * sizeof(float), hipMemcpyDeviceToDevice, st));
        R_TRY(hipMemsetAsync(r->fthick, 0, np * sizeof(float), st));
    }
    R_TRY(hipGetLastError());
    return RENDER_OK;
}

}  // namespace

extern "C" {

const char *render_last_error(void) { return t_error.c_str(); }

int render_create(const render_params *params, const float *cubemap, uint32_t cube_size,
                  render_ctx **out) {
    if (!params || !out || (!cubemap && cube_size))
        return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (params->width == 0 || params->height == 0 || params->width > 16384 ||
        params->height > 16384)
        return fail(RENDER_ERR_INVALID_ARGUMENT, "width and height must be in [1, 16384]");
    if (!(params->radius > 0.0) || !(params->fov > 0.0 && params->fov < M_PI))
        return fail(RENDER_ERR_INVALID_ARGUMENT, "radius must be > 0 and fov in (0, pi)");
    if (cube_size == 0 || cube_size > 8192)
        return fail(RENDER_ERR_INVALID_ARGUMENT, "cube_size must be in [1, 8192]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(RENDER_ERR_NO_DEVICE, "no HIP device");
    render_ctx *r = new render_ctx();
    const double radius = params->radius, diameter = 2.0 * radius;
    r->c.width = (int)params->width;
    r->c.height = (int)params->height;
    // fluidRender.ts:38-52 (JavaScript numbers, rounded to f32 as pipeline constants)
    r->c.depth_threshold = (float)(radius * 10.0);
    r->c.max_filter_size = 100.0f;
    r->c.projected_particle_constant =
        (float)((12.0 * diameter * 0.05 * ((double)params->height / 2.0)) / std::tan(params->fov / 2.0));
    r->cube_size = (int)cube_size;
    const size_t np = (size_t)params->width * params->height;
    const size_t cube_bytes = sizeof(float4) * 6 * (size_t)cube_size * cube_size;
    hipError_t e = hipGetDevice(&r->device);
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    r->tiles_x = (int)((params->width + kTile - 1) / kTile);
    r->tiles_y = (int)((params->height + kTile - 1) / kTile);
    const uint32_t ntiles = (uint32_t)(r->tiles_x * r->tiles_y);
    alloc((void **)&r->cube, cube_bytes);
    while (r->key_bits < 32 && (ntiles - 1) >> r->key_bits) ++r->key_bits;
    alloc((void **)&r->range, ntiles * sizeof(uint2));
    alloc((void **)&r->bilateral_w, sizeof(float) * kBilateralTable);
    alloc((void **)&r->gauss_w, sizeof(float) * kGaussTable);
    alloc((void **)&r->depth, np * sizeof(float));
    alloc((void **)&r->fdepth, np * sizeof(float));
    alloc((void **)&r->tmp, np * sizeof(float));
    alloc((void **)&r->thick16, np * sizeof(float));
    alloc((void **)&r->fthick, np * sizeof(float));
    alloc((void **)&r->color, np * sizeof(float4));
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&r->total, sizeof(uint64_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemcpy(r->cube, cubemap, cube_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_filter_weights<<<1, kThreads, 0, r->stream>>>(r->bilateral_w, r->gauss_w);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) {
        release(r);
        delete r;
        return fail(e == hipErrorOutOfMemory ? RENDER_ERR_OUT_OF_MEMORY : RENDER_ERR_DEVICE,
                    std::string("render_create: ") + hipGetErrorString(e));
    }
    *out = r;
    return RENDER_OK;
}

int render_destroy(render_ctx *r) {
    if (!r) return RENDER_OK;
    (void)hipSetDevice(r->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    release(r);
    delete r;
    return RENDER_OK;
}

int render_execute(render_ctx *r, const void *posvel_device, uint32_t n,
                   const render_uniforms *uniforms, int sphere_mode) {
    if (!r || !uniforms || (!posvel_device && n)) return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((uintptr_t)posvel_device % 16)
        return fail(RENDER_ERR_INVALID_ARGUMENT, "posvel_device must be 16-byte aligned");
    R_TRY(hipSetDevice(r->device));
    return frame(r, static_cast<const float4 *>(posvel_device), n, *uniforms, sphere_mode);
}

int render_execute_host(render_ctx *r, const void *posvel, uint32_t n,
                        const render_uniforms *uniforms, int sphere_mode) {
    if (!r || !uniforms || (!posvel && n)) return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    R_TRY(hipSetDevice(r->device));
    if (n > r->staging_cap) {
        R_TRY(hipStreamSynchronize(r->stream));
        if (r->staging) (void)hipFree(r->staging);
        r->staging = nullptr;
        r->staging_cap = 0;
        R_TRY(hipMalloc(&r->staging, 32 * (size_t)n));
        r->staging_cap = n;
    }
    if (n) {
        // ordered after the previous frame's reads of the staging buffer; the
        // caller's records are consumed before returning
        R_TRY(hipMemcpyAsync(r->staging, posvel, 32 * (size_t)n, hipMemcpyHostToDevice, r->stream));
        R_TRY(hipStreamSynchronize(r->stream));
    }
    return frame(r, r->staging, n, *uniforms, sphere_mode);
}

int render_read_image(render_ctx *r, int image, float *dst) {
    if (!r || !dst) return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t np = (size_t)r->c.width * r->c.height;
    const void *src = nullptr;
    size_t bytes = np * sizeof(float);
    switch (image) {
    case RENDER_IMAGE_DEPTH: src = r->depth; break;
    case RENDER_IMAGE_FILTERED_DEPTH: src = r->fdepth; break;
    case RENDER_IMAGE_THICKNESS: src = r->thick16; break;
    case RENDER_IMAGE_FILTERED_THICKNESS: src = r->fthick; break;
    case RENDER_IMAGE_COLOR:
        src = r->color;
        bytes = np * sizeof(float4);
        break;
    default: return fail(RENDER_ERR_INVALID_ARGUMENT, "unknown image");
    }
    R_TRY(hipSetDevice(r->device));
    R_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, r->stream));
    R_TRY(hipStreamSynchronize(r->stream));
    return RENDER_OK;
}

int render_color_device_ptr(render_ctx *r, void **out) {
    if (!r || !out) return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = r->color;
    return RENDER_OK;
}

int render_time_frames(render_ctx *r, const void *posvel_device, uint32_t n,
                       const render_uniforms *uniforms, int sphere_mode, uint32_t frames,
                       float *ms) {
    if (!r || !uniforms || !ms || (!posvel_device && n))
        return fail(RENDER_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((uintptr_t)posvel_device % 16)
        return fail(RENDER_ERR_INVALID_ARGUMENT, "posvel_device must be 16-byte aligned");
    R_TRY(hipSetDevice(r->device));
    hipEvent_t e0, e1;
    R_TRY(hipEventCreate(&e0));
    R_TRY(hipEventCreate(&e1));
    int rc = RENDER_OK;
    hipError_t e = hipEventRecord(e0, r->stream);
    for (uint32_t f = 0; f < frames && e == hipSuccess && rc == RENDER_OK; ++f)
        rc = frame(r, static_cast<const float4 *>(posvel_device), n, *uniforms, sphere_mode);
    if (e == hipSuccess) e = hipEventRecord(e1, r->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != RENDER_OK) return rc;
    if (e != hipSuccess) return fail(RENDER_ERR_DEVICE, std::string("timing: ") + hipGetErrorString(e));
    return RENDER_OK;
}

}  // extern "C"
