// render.hip -- the screen-space fluid renderer of WebGPU-Ocean
// (render/fluidRender.ts, render/*.wgsl) as gfx950 compute kernels behind the
// C ABI of include/render.h.
//
// A frame (FluidRenderer.execute, fluidRender.ts:424-484):
//   k_clear      depth keys, thickness accumulators
//   k_splat      one thread per particle walks the pixels of its impostor
//                rectangle: every fragment adds its thickness (the additive
//                thickness pass: no depth test, fluidRender.ts:130-149) with a
//                float atomic, and enters the depth test as a 64-bit key
//                (depth bits << 32 | particle index) through atomicMin -- the
//                lowest depth wins, ties to the earlier instance, as the
//                raster pipeline's less-than test in draw order
//   k_resolve    per pixel: the winner's view-space z (the depth map, r32float)
//                or, in sphere mode, its shaded colour; the thickness rounded
//                to f16 (the r16float target)
//   k_bilateral  4 x (along x, along y) on the depth map
//   k_gaussian   1 x (along x, along y) on the thickness, r16float
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

#include "render.h"

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
constexpr uint64_t kNoFragment = ~0ull;

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

__device__ __forceinline__ bool fragment(const Impostor &im, int px, int py, float nxy[2],
                                         float &r2) {
    const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    const float uvx = (cx - im.xl) / (im.xr - im.xl);
    const float uvy = (im.yb - cy) / (im.yb - im.yt);
    nxy[0] = uvx * 2.0f - 1.0f;
    nxy[1] = uvy * 2.0f - 1.0f;
    r2 = nxy[0] * nxy[0] + nxy[1] * nxy[1];
    return !(r2 > 1.0f);
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

// ---- kernels ----

__global__ __launch_bounds__(kThreads) void k_clear(uint64_t *key, float *thick, size_t np) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < np;
         i += (size_t)gridDim.x * kThreads) {
        key[i] = kNoFragment;
        thick[i] = 0.0f;
    }
}

// Both impostor passes over the same fragments: the thickness adds (no depth
// test in the reference's thickness pipeline) and the depth test.
__global__ __launch_bounds__(kThreads) void k_splat(const float4 *__restrict__ posvel, uint32_t n,
                                                    render_uniforms u, int w, int h,
                                                    uint64_t *key, float *thick) {
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const float4 pv = posvel[2 * (size_t)i];
        const float p[3] = {pv.x, pv.y, pv.z};
        Impostor im;
        if (!impostor(u, p, w, h, im)) continue;
        for (int py = im.py0; py < im.py1; ++py)
            for (int px = im.px0; px < im.px1; ++px) {
                float nxy[2], r2;
                if (!fragment(im, px, py, nxy, r2)) continue;
                const size_t pi = (size_t)py * w + px;
                float rvp[3], nrm[3];
                const float d = frag_depth(u, im, nxy, r2, rvp, nrm);
                const uint64_t k = ((uint64_t)__float_as_uint(d) << 32) | i;
                if (k < __hip_atomic_load(&key[pi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMin((unsigned long long *)&key[pi], (unsigned long long)k);
                unsafeAtomicAdd(&thick[pi], 0.05f * sqrtf(1.0f - r2));
            }
    }
}

// Per pixel: the depth map (the winner's view z; 0 = no fragment) or the
// sphere pass's colour, and the thickness map as r16float.
__global__ __launch_bounds__(kThreads) void k_resolve(const float4 *__restrict__ posvel,
                                                      render_uniforms u, int w, int h,
                                                      const uint64_t *key, const float *thick,
                                                      float *depth, float *thick16, float4 *color,
                                                      int sphere_mode) {
    const size_t np = (size_t)w * h;
    for (size_t pi = (size_t)blockIdx.x * kThreads + threadIdx.x; pi < np;
         pi += (size_t)gridDim.x * kThreads) {
        const uint64_t k = key[pi];
        float z = 0.0f;
        float4 c = make_float4(0.8f, 0.8f, 0.8f, 1.0f);
        if (k != kNoFragment) {
            const uint32_t i = (uint32_t)(k & 0xffffffffu);
            const float4 pv = posvel[2 * (size_t)i];
            const float p[3] = {pv.x, pv.y, pv.z};
            Impostor im;
            impostor(u, p, w, h, im);
            float nxy[2], r2, rvp[3], nrm[3];
            fragment(im, (int)(pi % w), (int)(pi / w), nxy, r2);
            frag_depth(u, im, nxy, r2, rvp, nrm);
            z = rvp[2];
            if (sphere_mode) {
                const float4 vv = posvel[2 * (size_t)i + 1];
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
        }
        depth[pi] = z;
        thick16[pi] = sphere_mode ? 0.0f : f16_round(thick[pi]);
        if (sphere_mode) color[pi] = c;
    }
}

// bilateral.wgsl:20-58 along `dir`
__global__ __launch_bounds__(kThreads) void k_bilateral(const float *__restrict__ src,
                                                        float *__restrict__ dst, Consts c,
                                                        float dirx, float diry) {
    const int w = c.width, h = c.height;
    const size_t np = (size_t)w * h;
    for (size_t pi = (size_t)blockIdx.x * kThreads + threadIdx.x; pi < np;
         pi += (size_t)gridDim.x * kThreads) {
        const int x = (int)(pi % w), y = (int)(pi / w);
        const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
        const float depth = fabsf(src[tex_index(iu, iv, w, h)]);
        if (depth >= 1e4f || depth <= 0.0f) {
            dst[pi] = depth;
            continue;
        }
        const float q = ceilf(c.projected_particle_constant / depth);
        const int fsz = q >= 2147483648.0f ? 2147483647 : (int)q;
        const int mfs = (int)c.max_filter_size;
        const int filter_size = mfs < fsz ? mfs : fsz;
        const float sigma = (float)filter_size / 3.0f;
        const float two_sigma = 2.0f * sigma * sigma;
        const float sigma_depth = c.depth_threshold / 3.0f;
        const float two_sigma_depth = 2.0f * sigma_depth * sigma_depth;
        float sum = 0.0f, wsum = 0.0f;
        for (int k = -filter_size; k <= filter_size; ++k) {
            const float co = (float)k;
            const float sampled = fabsf(src[tex_index(iu + co * dirx, iv + co * diry, w, h)]);
            const float rr = co * co + co * co;
            const float wt = expf(-rr / two_sigma);
            const float r_depth = sampled - depth;
            const float wd = expf(-r_depth * r_depth / two_sigma_depth);
            sum += sampled * wt * wd;
            wsum += wt * wd;
        }
        dst[pi] = sum / wsum;
    }
}

// gaussian.wgsl:15-42 along `dir` (r16float target)
__global__ __launch_bounds__(kThreads) void k_gaussian(const float *__restrict__ src,
                                                       float *__restrict__ dst, int w, int h,
                                                       float dirx, float diry) {
    const size_t np = (size_t)w * h;
    for (size_t pi = (size_t)blockIdx.x * kThreads + threadIdx.x; pi < np;
         pi += (size_t)gridDim.x * kThreads) {
        const int x = (int)(pi % w), y = (int)(pi / w);
        const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
        const float t = src[tex_index(iu, iv, w, h)];
        if (t == 0.0f) {
            dst[pi] = 0.0f;
            continue;
        }
        const int filter_size = 30;
        const float sigma = (float)filter_size / 3.0f;
        const float two_sigma = 2.0f * sigma * sigma;
        float sum = 0.0f, wsum = 0.0f;
        for (int k = -filter_size; k <= filter_size; ++k) {
            const float co = (float)k;
            const float sampled = src[tex_index(iu + dirx * co, iv + diry * co, w, h)];
            const float wt = expf(-co * co / two_sigma);
            sum += sampled * wt;
            wsum += wt;
        }
        dst[pi] = f16_round(sum / wsum);
    }
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
    const size_t np = (size_t)w * h;
    for (size_t pi = (size_t)blockIdx.x * kThreads + threadIdx.x; pi < np;
         pi += (size_t)gridDim.x * kThreads) {
        const int x = (int)(pi % w), y = (int)(pi / w);
        const float bg[3] = {0.8f, 0.8f, 0.8f};
        const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
        const float uvx = iu / (float)w, uvy = iv / (float)h;
        const float depth = fabsf(depth_map[tex_index(iu, iv, w, h)]);
        if (depth >= 1e4f || depth <= 0.0f) {
            color[pi] = make_float4(bg[0], bg[1], bg[2], 1.0f);
            continue;
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
        const float specular = powf(hn > 0.0f ? hn : 0.0f, 250.0f);
        const float density = 1.5f;
        const float thickness = thick[tex_index(iu, iv, w, h)];
        const float dc[3] = {0.085f, 0.6375f, 0.9f};
        float refr[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) refr[i] = bg[i] * expf(-density * thickness * (1.0f - dc[i]));
        const float F0 = 0.02f;
        const float nr[3] = {-ray[0], -ray[1], -ray[2]};
        float fr = F0 + (1.0f - F0) * powf(1.0f - dot3(normal, nr), 5.0f);
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
    float4 *cube = nullptr;
    uint64_t *key = nullptr;
    float *thick = nullptr, *depth = nullptr, *fdepth = nullptr, *tmp = nullptr;
    float *thick16 = nullptr, *fthick = nullptr;
    float4 *color = nullptr;
    float4 *staging = nullptr;  // host PosVel records (render_execute_host)
    uint32_t staging_cap = 0;
};

namespace {

void release(render_ctx *r) {
    void *ptrs[] = {r->cube, r->key, r->thick, r->depth, r->fdepth, r->tmp, r->thick16,
                    r->fthick, r->color, r->staging};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (r->stream) (void)hipStreamDestroy(r->stream);
}

int frame(render_ctx *r, const float4 *pv, uint32_t n, const render_uniforms &u, int sphere) {
    const int w = r->c.width, h = r->c.height;
    const size_t np = (size_t)w * h;
    const unsigned pb = blocks_for(np);
    hipStream_t st = r->stream;
    k_clear<<<pb, kThreads, 0, st>>>(r->key, r->thick, np);
    if (n) k_splat<<<blocks_for(n), kThreads, 0, st>>>(pv, n, u, w, h, r->key, r->thick);
    k_resolve<<<pb, kThreads, 0, st>>>(pv, u, w, h, r->key, r->thick, r->depth, r->thick16,
                                       r->color, sphere);
    if (!sphere) {
        // 4 x (bilateral x, bilateral y) (fluidRender.ts:432-444)
        const float *src = r->depth;
        for (int it = 0; it < 4; ++it) {
            k_bilateral<<<pb, kThreads, 0, st>>>(src, r->tmp, r->c, 1.0f, 0.0f);
            k_bilateral<<<pb, kThreads, 0, st>>>(r->tmp, r->fdepth, r->c, 0.0f, 1.0f);
            src = r->fdepth;
        }
        // 1 x (gaussian x, gaussian y) (fluidRender.ts:450-462)
        k_gaussian<<<pb, kThreads, 0, st>>>(r->thick16, r->tmp, w, h, 1.0f, 0.0f);
        k_gaussian<<<pb, kThreads, 0, st>>>(r->tmp, r->fthick, w, h, 0.0f, 1.0f);
        k_fluid<<<pb, kThreads, 0, st>>>(r->fdepth, r->fthick, r->cube, r->cube_size, u, w, h,
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
    alloc((void **)&r->cube, cube_bytes);
    alloc((void **)&r->key, np * sizeof(uint64_t));
    alloc((void **)&r->thick, np * sizeof(float));
    alloc((void **)&r->depth, np * sizeof(float));
    alloc((void **)&r->fdepth, np * sizeof(float));
    alloc((void **)&r->tmp, np * sizeof(float));
    alloc((void **)&r->thick16, np * sizeof(float));
    alloc((void **)&r->fthick, np * sizeof(float));
    alloc((void **)&r->color, np * sizeof(float4));
    if (e == hipSuccess) e = hipMemcpy(r->cube, cubemap, cube_bytes, hipMemcpyHostToDevice);
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
