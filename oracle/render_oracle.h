/* render_oracle.h -- CPU restatement of WebGPU-Ocean's screen-space fluid
 * renderer (the render/ WGSL shaders and render/fluidRender.ts).  TEST INFRASTRUCTURE ONLY:
 * the checker of the HIP renderer, never part of the product path. */
#ifndef RENDER_ORACLE_H
#define RENDER_ORACLE_H

#include <stdint.h>

typedef struct {
    float texel_size[2];
    float sphere_size;
    float pad;
    float inv_projection[16];
    float projection[16];
    float view[16];
    float inv_view[16];
} oracle_render_uniforms;

typedef struct {
    int width, height;
    float depth_threshold;             /* radius * 10           (fluidRender.ts:42,49) */
    float max_filter_size;             /* 100                   (fluidRender.ts:40,50) */
    float projected_particle_constant; /* (fluidRender.ts:51)                          */
} oracle_render_consts;

/* Images, width*height floats each (color: 4 per pixel).  threads: OpenMP
 * threads of the per-pixel passes (the splat is serial, in instance order). */
typedef struct {
    float *depth, *filtered_depth, *thickness, *filtered_thickness, *color;
} oracle_render_images;

float oracle_f16_round(float f);

/* One frame (FluidRenderer.execute, fluidRender.ts:332-484). cubemap: 6 faces
 * of size*size RGBA f32 (+X, -X, +Y, -Y, +Z, -Z). posvel: n * 8 floats. */
void oracle_render_frame(const float *posvel, uint32_t n, const oracle_render_uniforms *u,
                         const oracle_render_consts *c, const float *cubemap, int cube_size,
                         int sphere_mode, oracle_render_images *out, int threads);

/* Single pieces, for closed-form tests. */
float oracle_render_exp(float x);
float oracle_render_log(float x);
float oracle_render_pow5(float x);
float oracle_render_pow250(float x);
void oracle_cube_sample(const float *cubemap, int size, const float dir[3], float rgb[3]);
void oracle_bilateral_pass(const float *src, float *dst, int w, int h, const float dir[2],
                           const oracle_render_consts *c, int threads);
/* The fluid pass (fluid.wgsl) over a filtered depth map and a filtered
 * thickness map: color, 4 floats per pixel. */
void oracle_fluid_pass(const oracle_render_uniforms *u, const float *depth, const float *thickness,
                       const float *cubemap, int cube_size, int w, int h, float *color,
                       int threads);
void oracle_gaussian_pass(const float *src, float *dst, int w, int h, const float dir[2],
                          int threads);

#endif
