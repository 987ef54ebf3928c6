/* render_oracle.c -- CPU restatement of WebGPU-Ocean's screen-space fluid
 * renderer: render/depthMap.wgsl, bilateral.wgsl, thicknessMap.wgsl,
 * gaussian.wgsl, fluid.wgsl, sphere.wgsl and the pass sequence of
 * render/fluidRender.ts:332-484.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP renderer (compiled with
 * -ffp-contract=off: every float operation rounds on its own).
 *
 * The reference runs these as raster pipelines; this restatement fixes what
 * a rasteriser leaves to the implementation:
 *   - a fragment is a pixel whose centre lies in the impostor's screen
 *     rectangle [xl, xr) x [yt, yb) (left/top edges in, the top-left rule of
 *     the quad's two triangles), its uv interpolated linearly (the quad is
 *     parallel to the image plane, so perspective-correct = linear);
 *   - the depth test keeps the lowest depth, ties to the lower instance (the
 *     earlier draw); frag_depth is clamped to [0, 1]; impostors outside
 *     the clip volume (clip z not in [0, w]: nearer than the near plane or
 *     beyond the far plane) are not drawn;
 *   - the thickness map (r16float target, additive blending) blends each
 *     fragment in draw order: the f32 sum of the stored value and the
 *     fragment's output, rounded to f16 (round to nearest even) and stored;
 *   - textureLoad coordinates are clamped into the image (Tint's
 *     robustness), u32 conversions of negative coordinates give 0;
 *   - the cubemap sample is the major-axis face, bilinear, clamp to edge;
 *   - exp and pow (WGSL bounds their error, exp to 3 + 2|x| ulp, pow as
 *     exp2(y log2 x), without fixing their values) are one deterministic
 *     formula each, the same in the HIP kernels: oracle_render_exp (range
 *     reduction by ln 2 in two parts, degree-7 Taylor polynomial, 0 below
 *     x = -86), pow(x, 5) = ((x x)(x x)) x, pow(x, 250) =
 *     exp(250 log x) with oracle_render_log (atanh series).
 * mat4 x vec4 products are evaluated column by column,
 * ((c0 x + c1 y) + c2 z) + c3 w.
 */
#include "render_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void mv4(const float *m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* exp(x): x = k ln2 + r with k = rint(x log2 e) and r from ln2 split into
 * 0.693359375 + (-2.12194440e-4) (k ln2_hi exact); exp(r), |r| <= 0.35, by
 * its Taylor polynomial to degree 7 (truncation < 6e-9) in Horner form, one
 * fused multiply-add (one rounding) per step; times 2^k.  0 below -86 (results stay normal),
 * inf above 88, NaN through. */
float oracle_render_exp(float x) {
    if (x != x) return x;
    if (!(x > -86.0f)) return 0.0f;
    if (x > 88.0f) return INFINITY;
    const float k = rintf(x * 1.44269502f);
    const float r = fmaf(k, 2.12194440e-4f, x - k * 0.693359375f);
    float p = 1.98412698e-4f;           /* 1/7! */
    p = fmaf(p, r, 1.38888889e-3f);         /* 1/6! */
    p = fmaf(p, r, 8.33333333e-3f);         /* 1/5! */
    p = fmaf(p, r, 4.16666667e-2f);         /* 1/4! */
    p = fmaf(p, r, 1.66666667e-1f);         /* 1/3! */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

/* log(x) for x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m =
 * 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.172, by its series to s^9
 * (truncation < 2e-9); e ln2 with ln2 split as for exp.  -inf at 0, NaN
 * below. */
float oracle_render_log(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = frexpf(x, &e);          /* x = m 2^e, m in [0.5, 1) */
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f), s2 = s * s;
    float p = 1.11111111e-1f;          /* 1/9 */
    p = fmaf(p, s2, 1.42857143e-1f);       /* 1/7 */
    p = fmaf(p, s2, 2.00000000e-1f);       /* 1/5 */
    p = fmaf(p, s2, 3.33333333e-1f);       /* 1/3 */
    p = fmaf(p, s2, 1.0f);
    const float fe = (float)e;
    return fe * 0.693359375f + fmaf(2.0f * s, p, fe * -2.12194440e-4f);
}

/* pow(x, 5) and pow(x, 250) for x >= 0 (fluid.wgsl:70,80): ((x x)(x x)) x,
 * and exp(250 log x), whose error (about |250 log x| / 2 ulp) stays within
 * WGSL's bound for pow, 3 + 2 |250 log2 x| ulp. */
float oracle_render_pow5(float x) { return ((x * x) * (x * x)) * x; }

float oracle_render_pow250(float x) { return oracle_render_exp(250.0f * oracle_render_log(x)); }

static void normalize3(const float v[3], float out[3]) {
    const float l = sqrtf(dot3(v, v));
    for (int i = 0; i < 3; ++i) out[i] = v[i] / l;
}

/* f32 -> f16 (round to nearest even) -> f32, as an r16float store. */
float oracle_f16_round(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = x & 0x80000000u;
    uint32_t a = x & 0x7fffffffu;
    uint32_t h;
    if (a >= 0x7f800000u) {                  /* inf / nan */
        h = a > 0x7f800000u ? 0x7e00u : 0x7c00u;
    } else if (a >= 0x477ff000u) {           /* rounds to >= 65520: inf */
        h = 0x7c00u;
    } else if (a < 0x38800000u) {            /* below the smallest normal f16: subnormal */
        const uint32_t shift = 126u - (a >> 23);  /* 14 .. */
        if (shift > 24) {
            h = 0;
        } else {
            const uint32_t m = (a & 0x7fffffu) | 0x800000u;
            const uint32_t s = shift;  /* right shift of the 24-bit significand to units of 2^-24 */
            const uint32_t q = m >> s;
            const uint32_t rem = m & ((1u << s) - 1u), half = 1u << (s - 1u);
            h = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
        }
    } else {
        const uint32_t m = a & 0x7fffffu;
        const uint32_t e = (a >> 23) - 112u;  /* rebias 127 -> 15 */
        uint32_t q = (e << 10) | (m >> 13);
        const uint32_t rem = m & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
        h = q;
    }
    /* back to f32 */
    const uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
    uint32_t y;
    if (he == 0x1fu) {
        y = 0x7f800000u | (hm << 13);
    } else if (he == 0) {
        if (hm == 0) {
            y = 0;
        } else {
            float v = (float)hm * (1.0f / 16777216.0f);  /* hm * 2^-24 */
            memcpy(&y, &v, 4);
        }
    } else {
        y = ((he + 112u) << 23) | (hm << 13);
    }
    y |= sign;
    float r;
    memcpy(&r, &y, 4);
    return r;
}

/* textureLoad(tex, vec2u(coord), 0) with coordinates clamped into the image */
static int tex_index(float fx, float fy, int w, int h) {
    long ix = fx <= 0.0f ? 0 : (long)fx, iy = fy <= 0.0f ? 0 : (long)fy;  /* u32(): trunc, >= 0 */
    if (ix > w - 1) ix = w - 1;
    if (iy > h - 1) iy = h - 1;
    return (int)(iy * w + ix);
}

/* ---- impostor fragments (depthMap.wgsl:34-79, thicknessMap.wgsl:25-61,
 * sphere.wgsl:33-109) ---- */

typedef struct {
    float view[3];            /* view_position */
    float xl, xr, yt, yb;     /* screen rectangle (framebuffer pixels, y down) */
    int px0, px1, py0, py1;   /* covered pixel ranges [px0, px1) x [py0, py1) */
} Impostor;

static int impostor(const oracle_render_uniforms *u, const float p[3], int w, int h,
                    Impostor *im) {
    const float v4[4] = {p[0], p[1], p[2], 1.0f};
    float vp[4];
    mv4(u->view, v4, vp);
    for (int i = 0; i < 3; ++i) im->view[i] = vp[i];
    float fx[2], fy[2];
    for (int c = 0; c < 2; ++c) {
        const float cx = (c ? 0.5f : -0.5f) * u->sphere_size;  /* corner x of the quad */
        const float cy = (c ? 0.5f : -0.5f) * u->sphere_size;
        const float q[4] = {vp[0] + cx, vp[1] + cy, vp[2] + 0.0f, 1.0f};
        float clip[4];
        mv4(u->projection, q, clip);
        /* the quad is parallel to the image plane: its corners share clip z and w,
         * so the clip volume (0 <= z <= w) keeps all of it or none */
        if (!(clip[3] > 0.0f) || !(clip[2] >= 0.0f) || !(clip[2] <= clip[3])) return 0;
        const float nx = clip[0] / clip[3], ny = clip[1] / clip[3];
        fx[c] = ((nx + 1.0f) * 0.5f) * (float)w;
        fy[c] = ((1.0f - ny) * 0.5f) * (float)h;
    }
    im->xl = fx[0];
    im->xr = fx[1];
    im->yt = fy[1];  /* corner y +0.5 is the upper edge */
    im->yb = fy[0];
    if (!(im->xr > im->xl) || !(im->yb > im->yt)) return 0;
    /* pixel centres px + 0.5 in [xl, xr) */
    double a = ceil((double)im->xl - 0.5), b = ceil((double)im->xr - 0.5);
    double c2 = ceil((double)im->yt - 0.5), d = ceil((double)im->yb - 0.5);
    if (a < 0) a = 0;
    if (c2 < 0) c2 = 0;
    if (b > w) b = w;
    if (d > h) d = h;
    if (a >= b || c2 >= d) return 0;
    im->px0 = (int)a;
    im->px1 = (int)b;
    im->py0 = (int)c2;
    im->py1 = (int)d;
    return 1;
}

/* The fragment of impostor im at pixel (px, py): normalxy and r2 (returns 0
 * when discarded). */
static int fragment(const Impostor *im, int px, int py, float nxy[2], float *r2) {
    const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    const float uvx = (cx - im->xl) / (im->xr - im->xl);
    const float uvy = (im->yb - cy) / (im->yb - im->yt);
    nxy[0] = uvx * 2.0f - 1.0f;
    nxy[1] = uvy * 2.0f - 1.0f;
    *r2 = nxy[0] * nxy[0] + nxy[1] * nxy[1];
    return !(*r2 > 1.0f);
}

/* depth and real_view_pos of a fragment (depthMap.wgsl:68-75) */
static float frag_depth(const oracle_render_uniforms *u, const Impostor *im, const float nxy[2],
                        float r2, float rvp[3], float normal[3]) {
    const float nz = sqrtf(1.0f - r2);
    normal[0] = nxy[0];
    normal[1] = nxy[1];
    normal[2] = nz;
    const float radius = u->sphere_size / 2.0f;
    for (int i = 0; i < 3; ++i) rvp[i] = im->view[i] + normal[i] * radius;
    const float q[4] = {rvp[0], rvp[1], rvp[2], 1.0f};
    float clip[4];
    mv4(u->projection, q, clip);
    float d = clip[2] / clip[3];
    d = d > 0.0f ? d : 0.0f;  /* clamped to [0, 1]; -0 and NaN to +0 */
    if (d > 1.0f) d = 1.0f;
    return d;
}

/* sphere.wgsl:62-86 */
static void value_to_color(float value, float out[3]) {
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
    for (int i = 0; i < 3; ++i) out[i] = col[k][i] * (1.0f - t) + col[k + 1][i] * t;
}

/* ---- filters (bilateral.wgsl:20-58, gaussian.wgsl:15-42) ---- */

void oracle_bilateral_pass(const float *src, float *dst, int w, int h, const float dir[2],
                           const oracle_render_consts *c, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
            const float depth = fabsf(src[tex_index(iu, iv, w, h)]);
            if (depth >= 1e4f || depth <= 0.0f) {
                dst[y * w + x] = depth;
                continue;
            }
            const float q = ceilf(c->projected_particle_constant / depth);
            const int fsz = q >= 2147483648.0f ? 2147483647 : (int)q;
            const int mfs = (int)c->max_filter_size;
            const int filter_size = mfs < fsz ? mfs : fsz;
            const float sigma = (float)filter_size / 3.0f;
            const float two_sigma = 2.0f * sigma * sigma;
            const float sigma_depth = c->depth_threshold / 3.0f;
            const float two_sigma_depth = 2.0f * sigma_depth * sigma_depth;
            float sum = 0.0f, wsum = 0.0f;
            for (int k = -filter_size; k <= filter_size; ++k) {
                const float co = (float)k;  /* coords = vec2f(f32(x)): both components */
                const float sx = iu + co * dir[0], sy = iv + co * dir[1];
                const float sampled = fabsf(src[tex_index(sx, sy, w, h)]);
                const float rr = co * co + co * co;
                const float wt = oracle_render_exp(-rr / two_sigma);
                const float r_depth = sampled - depth;
                const float wd = oracle_render_exp(-r_depth * r_depth / two_sigma_depth);
                sum += sampled * wt * wd;
                wsum += wt * wd;
            }
            dst[y * w + x] = sum / wsum;
        }
}

void oracle_gaussian_pass(const float *src, float *dst, int w, int h, const float dir[2],
                          int threads) {
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
            const float t = src[tex_index(iu, iv, w, h)];
            if (t == 0.0f) {
                dst[y * w + x] = 0.0f;
                continue;
            }
            const int filter_size = 30;
            const float sigma = (float)filter_size / 3.0f;
            const float two_sigma = 2.0f * sigma * sigma;
            float sum = 0.0f, wsum = 0.0f;
            for (int k = -filter_size; k <= filter_size; ++k) {
                const float co = (float)k;
                const float sampled = src[tex_index(iu + dir[0] * co, iv + dir[1] * co, w, h)];
                const float wt = oracle_render_exp(-co * co / two_sigma);
                sum += sampled * wt;
                wsum += wt;
            }
            dst[y * w + x] = oracle_f16_round(sum / wsum);  /* r16float target */
        }
}

/* ---- cubemap (fluid.wgsl:80, textureSampleLevel with a linear sampler) ---- */

void oracle_cube_sample(const float *cubemap, int size, const float d[3], float rgb[3]) {
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
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (i1 > size - 1) i1 = size - 1;
    if (j1 > size - 1) j1 = size - 1;
    if (i0 > size - 1) i0 = size - 1;
    if (j0 > size - 1) j0 = size - 1;
    if (i1 < 0) i1 = 0;
    if (j1 < 0) j1 = 0;
    const float *fb = cubemap + (size_t)face * size * size * 4;
    for (int k = 0; k < 3; ++k) {
        const float t00 = fb[((size_t)j0 * size + i0) * 4 + k], t10 = fb[((size_t)j0 * size + i1) * 4 + k];
        const float t01 = fb[((size_t)j1 * size + i0) * 4 + k], t11 = fb[((size_t)j1 * size + i1) * 4 + k];
        const float top = t00 * (1.0f - a) + t10 * a, bot = t01 * (1.0f - a) + t11 * a;
        rgb[k] = top * (1.0f - b) + bot * b;
    }
}

/* ---- fluid shading (fluid.wgsl:20-84) ---- */

static void view_pos_from_uv_depth(const oracle_render_uniforms *u, float tx, float ty,
                                   float depth, float out[3]) {
    float ndc[4] = {tx * 2.0f - 1.0f, 1.0f - 2.0f * ty, 0.0f, 1.0f};
    ndc[2] = -u->projection[2 * 4 + 2] + u->projection[3 * 4 + 2] / depth;
    ndc[3] = 1.0f;
    float eye[4];
    mv4(u->inv_projection, ndc, eye);
    for (int i = 0; i < 3; ++i) out[i] = eye[i] / eye[3];
}

/* The canvas (bgra8unorm) stores a colour clamped to [0, 1], NaN as 0: the
 * output is that value before its 8-bit quantisation.  NaN arises where both
 * horizontal (or vertical) neighbours of a fluid pixel are background (depth
 * 0 there gives an infinite view position, fluid.wgsl:22-27). */
static float unorm(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

static void fluid_pixel(const oracle_render_uniforms *u, const float *depth_map,
                        const float *thick, const float *cubemap, int cube_size, int w, int h,
                        int x, int y, float out[4]) {
    const float bg[3] = {0.8f, 0.8f, 0.8f};
    const float iu = (float)x + 0.5f, iv = (float)y + 0.5f;
    const float uvx = iu / (float)w, uvy = iv / (float)h;
    const float depth = fabsf(depth_map[tex_index(iu, iv, w, h)]);
    out[3] = 1.0f;
    if (depth >= 1e4f || depth <= 0.0f) {
        for (int i = 0; i < 3; ++i) out[i] = bg[i];
        return;
    }
    float vp[3];
    view_pos_from_uv_depth(u, uvx, uvy, depth, vp);
    const float tsx = u->texel_size[0], tsy = u->texel_size[1];
    float p[3], ddx[3], ddy[3], ddx2[3], ddy2[3];
    view_pos_from_uv_depth(u, uvx + tsx, uvy + 0.0f,
                           fabsf(depth_map[tex_index(iu + 1.0f, iv + 0.0f, w, h)]), p);
    for (int i = 0; i < 3; ++i) ddx[i] = p[i] - vp[i];
    view_pos_from_uv_depth(u, uvx + 0.0f, uvy + tsy,
                           fabsf(depth_map[tex_index(iu + 0.0f, iv + 1.0f, w, h)]), p);
    for (int i = 0; i < 3; ++i) ddy[i] = p[i] - vp[i];
    view_pos_from_uv_depth(u, uvx + -tsx, uvy + 0.0f,
                           fabsf(depth_map[tex_index(iu + -1.0f, iv + 0.0f, w, h)]), p);
    for (int i = 0; i < 3; ++i) ddx2[i] = vp[i] - p[i];
    view_pos_from_uv_depth(u, uvx + 0.0f, uvy + -tsy,
                           fabsf(depth_map[tex_index(iu + 0.0f, iv + -1.0f, w, h)]), p);
    for (int i = 0; i < 3; ++i) ddy2[i] = vp[i] - p[i];
    if (fabsf(ddx[2]) > fabsf(ddx2[2])) memcpy(ddx, ddx2, sizeof ddx);
    if (fabsf(ddy[2]) > fabsf(ddy2[2])) memcpy(ddy, ddy2, sizeof ddy);
    const float cr[3] = {ddx[1] * ddy[2] - ddx[2] * ddy[1], ddx[2] * ddy[0] - ddx[0] * ddy[2],
                         ddx[0] * ddy[1] - ddx[1] * ddy[0]};
    float normal[3];
    normalize3(cr, normal);
    for (int i = 0; i < 3; ++i) normal[i] = -normal[i];
    float ray[3];
    normalize3(vp, ray);
    const float ld4[4] = {0.0f, 0.0f, -1.0f, 0.0f};
    float l4[4], light[3];
    mv4(u->view, ld4, l4);
    normalize3(l4, light);
    const float hv[3] = {light[0] - ray[0], light[1] - ray[1], light[2] - ray[2]};
    float H[3];
    normalize3(hv, H);
    const float hn = dot3(H, normal);
    const float specular = oracle_render_pow250(hn > 0.0f ? hn : 0.0f);
    const float density = 1.5f;
    const float thickness = thick[tex_index(iu, iv, w, h)];
    const float dc[3] = {0.085f, 0.6375f, 0.9f};
    float refr[3];
    for (int i = 0; i < 3; ++i) refr[i] = bg[i] * oracle_render_exp(-density * thickness * (1.0f - dc[i]));
    const float F0 = 0.02f;
    const float nr[3] = {-ray[0], -ray[1], -ray[2]};
    float fr = F0 + (1.0f - F0) * oracle_render_pow5(1.0f - dot3(normal, nr));
    fr = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
    const float dn = dot3(normal, ray);
    const float refl[4] = {ray[0] - 2.0f * dn * normal[0], ray[1] - 2.0f * dn * normal[1],
                           ray[2] - 2.0f * dn * normal[2], 0.0f};
    float rw[4];
    mv4(u->inv_view, refl, rw);
    float rc[3];
    oracle_cube_sample(cubemap, cube_size, rw, rc);
    for (int i = 0; i < 3; ++i) out[i] = unorm(1.0f * specular + (refr[i] * (1.0f - fr) + rc[i] * fr));
}

/* ---- the frame (fluidRender.ts:424-484) ---- */

void oracle_fluid_pass(const oracle_render_uniforms *u, const float *depth, const float *thickness,
                       const float *cubemap, int cube_size, int w, int h, float *color,
                       int threads) {
    (void)threads;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            fluid_pixel(u, depth, thickness, cubemap, cube_size, w, h, x, y,
                        color + 4 * ((size_t)y * w + x));
}

void oracle_render_frame(const float *posvel, uint32_t n, const oracle_render_uniforms *u,
                         const oracle_render_consts *c, const float *cubemap, int cube_size,
                         int sphere_mode, oracle_render_images *out, int threads) {
    const int w = c->width, h = c->height;
    const size_t np = (size_t)w * h;
    float *zbuf = (float *)malloc(np * sizeof(float));
    uint32_t *who = (uint32_t *)malloc(np * sizeof(uint32_t));
    float *thick = (float *)calloc(np, sizeof(float));
    for (size_t i = 0; i < np; ++i) {
        zbuf[i] = 1.0f;  /* depthClearValue */
        who[i] = 0xffffffffu;
        out->depth[i] = 0.0f;  /* color clearValue */
    }
    /* depth map pass and thickness map pass: the same fragments */
    for (uint32_t i = 0; i < n; ++i) {
        Impostor im;
        if (!impostor(u, posvel + 8 * (size_t)i, w, h, &im)) continue;
        for (int py = im.py0; py < im.py1; ++py)
            for (int px = im.px0; px < im.px1; ++px) {
                float nxy[2], r2;
                if (!fragment(&im, px, py, nxy, &r2)) continue;
                const size_t pi = (size_t)py * w + px;
                float rvp[3], nrm[3];
                const float d = frag_depth(u, &im, nxy, r2, rvp, nrm);
                if (d < zbuf[pi]) {
                    zbuf[pi] = d;
                    who[pi] = i;
                    out->depth[pi] = rvp[2];
                }
                thick[pi] = oracle_f16_round(thick[pi] + 0.05f * sqrtf(1.0f - r2));
            }
    }
    if (sphere_mode) {
        /* sphere.wgsl: the winning fragment's colour, background 0.8 */
        for (size_t pi = 0; pi < np; ++pi) {
            float *o = out->color + 4 * pi;
            o[0] = o[1] = o[2] = 0.8f;
            o[3] = 1.0f;
            if (who[pi] == 0xffffffffu) continue;
            const float *pv = posvel + 8 * (size_t)who[pi];
            Impostor im;
            impostor(u, pv, w, h, &im);
            float nxy[2], r2, rvp[3], nrm[3];
            fragment(&im, (int)(pi % w), (int)(pi / w), nxy, &r2);
            frag_depth(u, &im, nxy, r2, rvp, nrm);
            const float speed = sqrtf(dot3(pv + 4, pv + 4));
            const float one[3] = {1.0f, 1.0f, 1.0f};
            float ln[3], col[3];
            normalize3(one, ln);
            const float dd = dot3(nrm, ln);
            const float diffuse = dd > 0.0f ? dd : 0.0f;
            value_to_color(speed / 1.5f, col);
            for (int k = 0; k < 3; ++k) o[k] = unorm(col[k] * diffuse);
        }
        memcpy(out->filtered_depth, out->depth, np * sizeof(float));
        memset(out->thickness, 0, np * sizeof(float));
        memset(out->filtered_thickness, 0, np * sizeof(float));
        free(zbuf);
        free(who);
        free(thick);
        return;
    }
    /* 4 x (bilateral x, bilateral y) on the depth map */
    float *tmp = (float *)malloc(np * sizeof(float));
    memcpy(out->filtered_depth, out->depth, np * sizeof(float));
    const float dx[2] = {1.0f, 0.0f}, dy[2] = {0.0f, 1.0f};
    for (int it = 0; it < 4; ++it) {
        oracle_bilateral_pass(out->filtered_depth, tmp, w, h, dx, c, threads);
        oracle_bilateral_pass(tmp, out->filtered_depth, w, h, dy, c, threads);
    }
    /* thickness (r16float) and 1 x (gaussian x, gaussian y) */
    memcpy(out->thickness, thick, np * sizeof(float));
    oracle_gaussian_pass(out->thickness, tmp, w, h, dx, threads);
    oracle_gaussian_pass(tmp, out->filtered_thickness, w, h, dy, threads);
    oracle_fluid_pass(u, out->filtered_depth, out->filtered_thickness, cubemap, cube_size, w, h,
                      out->color, threads);
    free(tmp);
    free(zbuf);
    free(who);
    free(thick);
}
