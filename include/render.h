/*
 * render.h -- C ABI of the screen-space fluid renderer in libmpm355: an
 * MI355X (gfx950) implementation of WebGPU-Ocean's FluidRenderer
 * (render/fluidRender.ts:10-488) as compute kernels.
 *
 * What a Node N-API addon (or any FFI) binds to replace the reference's
 * FluidRenderer class.  Plain pointers and sizes only; every function returns
 * a render_status (0 = OK) and never throws or aborts across the boundary;
 * render_last_error() returns a message for the calling thread.
 *
 * Threading: one handle per host thread; all device work of a handle is
 * ordered on that handle's HIP stream.
 *
 * Input layouts (identical to the reference's GPU buffers):
 *   posvel    32 B  position f32[3] @0, v f32[3] @16 (render/depthMap.wgsl:26-29):
 *                   what mpm_posvel_device_ptr / sph_posvel_device_ptr return
 *   uniforms 272 B  render_uniforms below (common.ts:1-9)
 *   cubemap         6 faces +X, -X, +Y, -Y, +Z, -Z, each size x size texels of
 *                   RGBA f32, rows top to bottom (the reference's cubemap
 *                   texture, main.ts:82-103, as linear floats)
 * Output: width x height RGBA f32, rows top to bottom: the canvas's colour
 * before its 8-bit quantisation (clamped to [0, 1], NaN as 0, as a unorm
 * target stores it).
 *
 * The passes are the reference's (fluidRender.ts:424-484): the sphere
 * impostors' depth map with a less-than depth test, 4 x (bilateral filter
 * along x, then y), the additive thickness map (r16float), 1 x (gaussian filter
 * along x, then y), and the fluid shading pass; or, in sphere mode, the
 * sphere pass.  The impostors are rasterised in screen tiles, each tile
 * walking its impostors in draw order: a fragment is a pixel whose centre
 * lies in the impostor's screen rectangle, left and top edges included (the
 * rectangle's top-left rule), with its uv interpolated linearly (the quad is
 * parallel to the image plane, so perspective-correct interpolation is
 * linear); ties of the depth test go to the lower particle index (the
 * earlier-drawn instance); the thickness target blends one fragment at a
 * time in draw order, rounding to f16 after each, as a raster pipeline's
 * blend unit does.  Frames are deterministic.
 *
 * render_execute waits once per frame for the device (the number of
 * (tile, impostor) pairs sizes the tile lists); the rest of the frame runs
 * asynchronously.  A frame whose impostors touch more than 2^31 tile slots
 * (a camera right at the particles) fails with RENDER_ERR_OUT_OF_MEMORY.
 */
#ifndef RENDER355_H
#define RENDER355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RENDER_OK = 0,
    RENDER_ERR_INVALID_ARGUMENT = 1,
    RENDER_ERR_OUT_OF_MEMORY = 3,
    RENDER_ERR_DEVICE = 4, /* a HIP runtime call failed */
    RENDER_ERR_NO_DEVICE = 5,
} render_status;

/* The reference's render uniform buffer (common.ts:1-9): column-major 4x4
 * matrices, as WGSL's mat4x4f. */
typedef struct {
    float texel_size[2];          /* 1 / width, 1 / height (main.ts:106) */
    float sphere_size;            /* particle diameter (mls-mpm/mls-mpm.ts:226) */
    float pad;
    float inv_projection[16];
    float projection[16];
    float view[16];
    float inv_view[16];
} render_uniforms;

/* Constructor arguments of FluidRenderer (fluidRender.ts:35-53). */
typedef struct {
    uint32_t width, height; /* canvas.width, canvas.height */
    double radius;          /* particle radius (main.ts:138, 0.6 for MLS-MPM) */
    double fov;             /* vertical field of view, radians (main.ts:137) */
} render_params;

/* Intermediate images a caller (or a test) can read back after a frame. */
typedef enum {
    RENDER_IMAGE_DEPTH = 0,          /* the depth map pass's r32float: view-space z, 0 = none */
    RENDER_IMAGE_FILTERED_DEPTH = 1, /* after the 4 bilateral iterations */
    RENDER_IMAGE_THICKNESS = 2,      /* the thickness map (r16float values, as f32) */
    RENDER_IMAGE_FILTERED_THICKNESS = 3,
    RENDER_IMAGE_COLOR = 4,          /* the output: RGBA f32 (4 floats per pixel) */
} render_image;

typedef struct render_ctx render_ctx;

const char *render_last_error(void);

/* new FluidRenderer(device, canvas, format, radius, fov, posvelBuffer,
 * renderUniformBuffer, cubemapTextureView) (fluidRender.ts:35).  The cubemap
 * is copied to the device: 6 * cube_size^2 * 4 floats. */
int render_create(const render_params *params, const float *cubemap, uint32_t cube_size,
                  render_ctx **out);

int render_destroy(render_ctx *r);

/* FluidRenderer.execute(context, encoder, numParticles, sphereRenderFl)
 * (fluidRender.ts:332): renders the first n PosVel records of the device
 * buffer `posvel_device` (n * 32 bytes, 16-byte aligned, e.g.
 * mpm_posvel_device_ptr) with the given uniforms.  Asynchronous on the
 * handle's stream after one wait for the device (see above); the buffer must
 * be complete when called (synchronize its producer first, e.g.
 * mpm_synchronize). */
int render_execute(render_ctx *r, const void *posvel_device, uint32_t n,
                   const render_uniforms *uniforms, int sphere_mode);

/* The same from host PosVel records (copied to the device first; the
 * records may be freed when the call returns). */
int render_execute_host(render_ctx *r, const void *posvel, uint32_t n,
                        const render_uniforms *uniforms, int sphere_mode);

/* Copies an image of the last frame to host memory (synchronizes):
 * width * height floats, or 4 floats per pixel for RENDER_IMAGE_COLOR. */
int render_read_image(render_ctx *r, int image, float *dst);

/* Device address of the output image (RGBA f32), valid until render_destroy. */
int render_color_device_ptr(render_ctx *r, void **out);

/* Milliseconds of `frames` render_execute calls (device events), for benchmarks. */
int render_time_frames(render_ctx *r, const void *posvel_device, uint32_t n,
                       const render_uniforms *uniforms, int sphere_mode, uint32_t frames,
                       float *ms);

#ifdef __cplusplus
}
#endif

#endif /* RENDER355_H */
