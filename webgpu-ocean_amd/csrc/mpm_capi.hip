// mpm_capi.hip -- the C ABI (include/mpm.h): handle lifetime, validation,
// stream ordering and error reporting.  mpm_sim mirrors MLSMPMSimulator
// (mls-mpm/mls-mpm.ts:12-276) as a batch of one scene; mpm_batch holds many
// independent scenes advanced by the same kernels.  Nothing throws across
// the boundary.
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mpm.h"
#include "mpm_batch.h"
#include "mpm_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? MPM_ERR_OUT_OF_MEMORY : MPM_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(call)                                       \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

// A device context shared by both handle kinds.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    mpm::Batch batch;
};

int ctx_open(Context &c) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(MPM_ERR_NO_DEVICE, "no HIP device is visible");
    HIP_TRY(hipGetDevice(&c.device));
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    return MPM_OK;
}

void ctx_close(Context &c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    c.batch.release();
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
}

int select_device(const Context &c) {
    HIP_TRY(hipSetDevice(c.device));
    return MPM_OK;
}

// Every call that touches the scene goes through this: after a failed reset
// the handle holds no scene and no buffers, and must not launch anything.
int ready(const Context &c) {
    if (!c.batch.ready())
        return fail(MPM_ERR_NOT_READY,
                    "the handle holds no scene (its last reset failed); call mpm_reset again");
    return select_device(c);
}

int check_params(const mpm_params &p) {
    const float vals[5] = {p.stiffness, p.rest_density, p.dynamic_viscosity, p.dt,
                           p.fixed_point_multiplier};
    for (float f : vals)
        if (!std::isfinite(f)) return fail(MPM_ERR_INVALID_ARGUMENT, "params must be finite");
    if (!(p.fixed_point_multiplier > 0.0f) || !(p.rest_density != 0.0f))
        return fail(MPM_ERR_INVALID_ARGUMENT,
                    "fixed_point_multiplier must be > 0 and rest_density non-zero");
    return MPM_OK;
}

int check_box(const float b[3], const char *what) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b[a]))
            return fail(MPM_ERR_INVALID_ARGUMENT, std::string(what) + " must be finite");
    return MPM_OK;
}

// Validates scenes and the packed totals (32-bit slots, cells and keys).
int check_scenes(const mpm::SceneInit *sc, uint32_t count) {
    uint64_t tn = 0, tg = 0, tk = 0;
    for (uint32_t s = 0; s < count; ++s) {
        int rc = check_box(sc[s].init_box, "init_box");
        if (rc) return rc;
        for (int a = 0; a < 3; ++a)
            if (sc[s].init_box[a] < 1.0f || sc[s].init_box[a] > 1048576.0f)
                return fail(MPM_ERR_INVALID_ARGUMENT, "init_box components must lie in [1, 2^20]");
        if (sc[s].n && !sc[s].particles)
            return fail(MPM_ERR_INVALID_ARGUMENT, "particles is NULL");
        // gridCount = ceil(bx)*ceil(by)*ceil(bz) (mls-mpm/mls-mpm.ts:229); i32 cell indices.
        uint64_t G = (uint64_t)std::ceil(sc[s].init_box[0]) *
                     (uint64_t)std::ceil(sc[s].init_box[1]) *
                     (uint64_t)std::ceil(sc[s].init_box[2]);
        if (G > (uint64_t)INT32_MAX)
            return fail(MPM_ERR_GRID_TOO_LARGE,
                        "gridCount " + std::to_string(G) +
                            " exceeds the 2^31-1 cells addressable by i32 cell indices");
        uint64_t nk = 1;
        for (int a = 0; a < 3; ++a) nk *= ((uint64_t)std::ceil(sc[s].init_box[a]) + 3) / 4;
        tn += sc[s].n;
        tg += G;
        tk += nk * 64;
    }
    if (tn > 0xfffff000ull || tg > 0xffffffffull || tk > 0xffffffffull)
        return fail(MPM_ERR_GRID_TOO_LARGE, "batch exceeds 2^32 particles, cells or bin keys");
    return MPM_OK;
}

mpm::Params to_params(const mpm_params &p) {
    return mpm::Params{p.stiffness, p.rest_density, p.dynamic_viscosity, p.dt,
                       p.fixed_point_multiplier};
}

void to_c_tuning(const mpm::Tuning &t, mpm_tuning *out) {
    out->path = t.path;
    out->particles_per_thread = t.ppt;
    out->grid_blocks = t.grid_blocks;
    out->xcd_remap = t.xcd_remap;
    out->tile_edge = 1 << t.tile_bits;
}

int from_c_tuning(const mpm_tuning *in, mpm::Batch &b, hipStream_t st) {
    mpm::Tuning t;
    t.path = in->path;
    t.ppt = in->particles_per_thread;
    t.grid_blocks = in->grid_blocks;
    t.xcd_remap = in->xcd_remap;
    t.tile_bits = in->tile_edge == 8 ? 3 : in->tile_edge == 4 ? 2 : -1;
    const hipError_t e = t.tile_bits < 0 ? hipErrorInvalidValue : b.set_tuning(t, st);
    if (e == hipErrorInvalidValue)
        return fail(MPM_ERR_INVALID_ARGUMENT,
                    "tuning out of range: path 0|1 (1 only for one scene), "
                    "particles_per_thread 1|2|4, grid_blocks 0..2^20, xcd_remap 0|1, "
                    "tile_edge 4|8");
    if (e != hipSuccess) return hip_fail(e, "set_tuning");
    return MPM_OK;
}

int time_stages_impl(Context &c, uint32_t substeps, float *stage_ms, int *calls,
                     const char **names, int capacity, int *count) {
    if (!stage_ms || !count) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c.stream));
    hipError_t e =
        c.batch.time_stages(c.stream, substeps, stage_ms, calls, names, capacity, count);
    if (e != hipSuccess) return hip_fail(e, "time_stages");
    return MPM_OK;
}

int time_substeps(Context &c, uint32_t substeps, float *ms) {
    int rc = ready(c);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(hipEventRecord(e0, c.stream));
    hipError_t e = hipSuccess;
    for (uint32_t k = 0; k < substeps && e == hipSuccess; ++k) e = c.batch.substep(c.stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c.stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "substep");
    return MPM_OK;
}

}  // namespace

struct mpm_sim {
    Context ctx;
    mpm_params params{};
};

struct mpm_batch {
    Context ctx;
};

extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

const char *mpm_last_error(void) { return g_last_error.c_str(); }

int mpm_default_params(mpm_params *out) {
    if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "out is NULL");
    out->stiffness = 3.0f;
    out->rest_density = 4.0f;
    out->dynamic_viscosity = 0.1f;
    out->dt = 0.2f;
    out->fixed_point_multiplier = 1e7f;
    return MPM_OK;
}

// ---- single simulator -------------------------------------------------------

int mpm_create(const mpm_params *params, const void *particles, uint32_t n,
               const float init_box[3], mpm_sim **out) {
    if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    mpm_params p;
    if (params)
        p = *params;
    else
        mpm_default_params(&p);
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_box(init_box, "init_box");
    if (rc) return rc;
    mpm::SceneInit si{particles, n, {init_box[0], init_box[1], init_box[2]}, to_params(p)};
    rc = check_scenes(&si, 1);
    if (rc) return rc;
    mpm_sim *sim = new (std::nothrow) mpm_sim();
    if (!sim) return fail(MPM_ERR_OUT_OF_MEMORY, "host allocation failed");
    sim->params = p;
    rc = ctx_open(sim->ctx);
    if (!rc) {
        hipError_t e = sim->ctx.batch.init(&si, 1, sim->ctx.stream);
        if (e != hipSuccess) rc = hip_fail(e, "create");
    }
    if (rc) {
        mpm_destroy(sim);
        return rc;
    }
    *out = sim;
    return MPM_OK;
}

int mpm_destroy(mpm_sim *sim) {
    if (!sim) return MPM_OK;
    ctx_close(sim->ctx);
    delete sim;
    return MPM_OK;
}

int mpm_reset(mpm_sim *sim, const void *particles, uint32_t n, const float init_box[3]) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = check_box(init_box, "init_box");
    if (rc) return rc;
    mpm::SceneInit si{particles, n, {init_box[0], init_box[1], init_box[2]},
                      to_params(sim->params)};
    rc = check_scenes(&si, 1);
    if (rc) return rc;
    rc = select_device(sim->ctx);  // (a reset is what makes a failed handle usable again)
    if (rc) return rc;
    hipError_t e = sim->ctx.batch.init(&si, 1, sim->ctx.stream);
    if (e != hipSuccess) return hip_fail(e, "reset");
    return MPM_OK;
}

int mpm_set_box(mpm_sim *sim, const float real_box[3]) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = check_box(real_box, "real_box");
    if (rc) return rc;
    rc = ready(sim->ctx);
    if (rc) return rc;
    HIP_TRY(sim->ctx.batch.set_box(0, real_box, sim->ctx.stream));
    return MPM_OK;
}

int mpm_step(mpm_sim *sim, uint32_t substeps) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    for (uint32_t k = 0; k < substeps; ++k) HIP_TRY(sim->ctx.batch.substep(sim->ctx.stream));
    return MPM_OK;
}

int mpm_run_stage(mpm_sim *sim, int stage) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    if (stage < 0 || stage >= MPM_STAGE_COUNT)
        return fail(MPM_ERR_INVALID_ARGUMENT, "stage out of range");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    HIP_TRY(mpm::launch_stage(stage, sim->ctx.batch.stage_view(), sim->ctx.stream));
    HIP_TRY(sim->ctx.batch.external_change(sim->ctx.stream));
    return MPM_OK;
}

int mpm_synchronize(mpm_sim *sim) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = select_device(sim->ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->ctx.stream));
    return MPM_OK;
}

int mpm_read_posvel(mpm_sim *sim, void *dst, uint32_t n) {
    if (!sim || (!dst && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (n > sim->ctx.batch.scene(0).n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds numParticles");
    HIP_TRY(sim->ctx.batch.read_posvel(0, dst, n, sim->ctx.stream));
    return MPM_OK;
}

int mpm_posvel_device_ptr(mpm_sim *sim, void **out) {
    if (!sim || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    int rc = ready(sim->ctx);
    if (rc) return rc;
    *out = sim->ctx.batch.posvel_ptr(0);
    return MPM_OK;
}

int mpm_read_particles(mpm_sim *sim, void *dst, uint32_t n) {
    if (!sim || (!dst && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (n > sim->ctx.batch.scene(0).n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds numParticles");
    HIP_TRY(sim->ctx.batch.read_particles(0, dst, n, sim->ctx.stream));
    return MPM_OK;
}

int mpm_write_particles(mpm_sim *sim, const void *src, uint32_t n) {
    if (!sim || (!src && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (n != sim->ctx.batch.scene(0).n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n must equal numParticles (use mpm_reset)");
    HIP_TRY(sim->ctx.batch.write_particles(0, src, n, sim->ctx.stream));
    return MPM_OK;
}

int mpm_read_grid(mpm_sim *sim, int32_t *dst, uint64_t ncells) {
    if (!sim || (!dst && ncells)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (ncells > sim->ctx.batch.scene(0).G)
        return fail(MPM_ERR_INVALID_ARGUMENT, "ncells exceeds gridCount");
    HIP_TRY(sim->ctx.batch.read_grid(0, dst, ncells, sim->ctx.stream));
    return MPM_OK;
}

int mpm_read_grid_velocity(mpm_sim *sim, float *dst, uint64_t ncells) {
    if (!sim || (!dst && ncells)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (ncells > sim->ctx.batch.scene(0).G)
        return fail(MPM_ERR_INVALID_ARGUMENT, "ncells exceeds gridCount");
    HIP_TRY(sim->ctx.batch.read_vel(0, dst, ncells, sim->ctx.stream));
    return MPM_OK;
}

int mpm_write_grid(mpm_sim *sim, const int32_t *src, uint64_t ncells) {
    if (!sim || (!src && ncells)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    if (ncells > sim->ctx.batch.scene(0).G)
        return fail(MPM_ERR_INVALID_ARGUMENT, "ncells exceeds gridCount");
    HIP_TRY(sim->ctx.batch.write_grid(0, src, ncells, sim->ctx.stream));
    return MPM_OK;
}

int mpm_get_counts(mpm_sim *sim, uint32_t *num_particles, uint64_t *grid_count) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    if (!sim->ctx.batch.ready()) {
        if (num_particles) *num_particles = 0;
        if (grid_count) *grid_count = 0;
        return fail(MPM_ERR_NOT_READY, "the handle holds no scene (its last reset failed)");
    }
    if (num_particles) *num_particles = sim->ctx.batch.scene(0).n;
    if (grid_count) *grid_count = sim->ctx.batch.scene(0).G;
    return MPM_OK;
}

int mpm_time_steps(mpm_sim *sim, uint32_t frames, uint32_t substeps_per_frame, float *ms) {
    if (!sim || !ms) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    return time_substeps(sim->ctx, frames * substeps_per_frame, ms);
}

int mpm_time_stages(mpm_sim *sim, uint32_t substeps, float *stage_ms, int *calls,
                    const char **names, int capacity, int *count) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    return time_stages_impl(sim->ctx, substeps, stage_ms, calls, names, capacity, count);
}

int mpm_get_tuning(mpm_sim *sim, mpm_tuning *out) {
    if (!sim || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    to_c_tuning(sim->ctx.batch.tuning(), out);
    return MPM_OK;
}

int mpm_set_tuning(mpm_sim *sim, const mpm_tuning *in) {
    if (!sim || !in) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->ctx.stream));
    return from_c_tuning(in, sim->ctx.batch, sim->ctx.stream);
}

int mpm_autotune(mpm_sim *sim, uint32_t substeps_per_trial, mpm_tuning *chosen) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    mpm::Tuning t;
    HIP_TRY(sim->ctx.batch.autotune(sim->ctx.stream, substeps_per_trial ? substeps_per_trial : 4,
                                    &t));
    if (chosen) to_c_tuning(t, chosen);
    return MPM_OK;
}

int mpm_backward(mpm_sim *sim, const mpm_adjoint_config *cfg, float *grad_v0, float *grad_x0,
                 double *loss, float *grad_stiffness, float *grad_viscosity) {
    if (!sim || !cfg) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (cfg->substeps == 0 || cfg->checkpoint_every == 0)
        return fail(MPM_ERR_INVALID_ARGUMENT, "substeps and checkpoint_every must be >= 1");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(cfg->target[a]))
            return fail(MPM_ERR_INVALID_ARGUMENT, "target must be finite");
    if (!std::isfinite(cfg->ke_weight)) return fail(MPM_ERR_INVALID_ARGUMENT, "ke_weight must be finite");
    int rc = ready(sim->ctx);
    if (rc) return rc;
    mpm::AdjointConfig ac{cfg->substeps, cfg->checkpoint_every,
                          {cfg->target[0], cfg->target[1], cfg->target[2]}, cfg->ke_weight};
    mpm::AdjointResult res;
    res.grad_v0 = grad_v0;
    res.grad_x0 = grad_x0;
    HIP_TRY(sim->ctx.batch.backward(ac, sim->ctx.stream, res));
    if (loss) *loss = res.loss;
    if (grad_stiffness) *grad_stiffness = res.grad_stiffness;
    if (grad_viscosity) *grad_viscosity = res.grad_viscosity;
    return MPM_OK;
}

// ---- allocation-free step ------------------------------------------------------

int mpm_step_noalloc(const mpm_params *params, void *particles, void *posvel, uint32_t n,
                     const float init_box[3], const float real_box[3], uint32_t substeps,
                     void *stream) {
    mpm_params p;
    if (params)
        p = *params;
    else
        mpm_default_params(&p);
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_box(init_box, "init_box");
    if (rc) return rc;
    rc = check_box(real_box, "real_box");
    if (rc) return rc;
    if (n && (!particles || !posvel))
        return fail(MPM_ERR_INVALID_ARGUMENT, "particles and posvel must be device buffers");
    for (int a = 0; a < 3; ++a)
        if (init_box[a] < 1.0f || init_box[a] > 1048576.0f)
            return fail(MPM_ERR_INVALID_ARGUMENT, "init_box components must lie in [1, 2^20]");
    const uint64_t G = (uint64_t)std::ceil(init_box[0]) * (uint64_t)std::ceil(init_box[1]) *
                       (uint64_t)std::ceil(init_box[2]);
    mpm::Uniforms *uni = nullptr;
    int4 *grid = nullptr;
    uint32_t max_cells = 0;
    HIP_TRY(mpm::noalloc_scratch(&uni, &grid, &max_cells));
    if (G > max_cells)
        return fail(MPM_ERR_GRID_TOO_LARGE,
                    "gridCount " + std::to_string(G) + " exceeds the " +
                        std::to_string(max_cells) +
                        " cells of the allocation-free scratch grid (64^3, the reference's cap)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const mpm::Params prm = to_params(p);
    HIP_TRY(mpm::launch_set_uniforms(uni, mpm::derive_uniforms(init_box, real_box), st));
    float *rec = static_cast<float *>(particles);
    float4 *pv = static_cast<float4 *>(posvel);
    for (uint32_t k = 0; k < substeps; ++k)
        for (int stage = 0; stage < MPM_STAGE_COUNT; ++stage)
            HIP_TRY(mpm::launch_stage_records(stage, rec, n, grid, (uint32_t)G, uni, prm, pv, st));
    return MPM_OK;
}

// ---- batch of independent scenes ---------------------------------------------

int mpm_batch_create(const mpm_scene *scenes, uint32_t count, mpm_batch **out) {
    if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!scenes || count == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "no scenes");
    std::vector<mpm::SceneInit> si(count);
    for (uint32_t s = 0; s < count; ++s) {
        int rc = check_params(scenes[s].params);
        if (rc) return rc;
        rc = check_box(scenes[s].init_box, "init_box");
        if (rc) return rc;
        si[s] = mpm::SceneInit{scenes[s].particles, scenes[s].n,
                               {scenes[s].init_box[0], scenes[s].init_box[1],
                                scenes[s].init_box[2]},
                               to_params(scenes[s].params)};
    }
    int rc = check_scenes(si.data(), count);
    if (rc) return rc;
    mpm_batch *b = new (std::nothrow) mpm_batch();
    if (!b) return fail(MPM_ERR_OUT_OF_MEMORY, "host allocation failed");
    rc = ctx_open(b->ctx);
    if (!rc) {
        hipError_t e = b->ctx.batch.init(si.data(), count, b->ctx.stream);
        if (e != hipSuccess) rc = hip_fail(e, "batch create");
    }
    if (rc) {
        mpm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MPM_OK;
}

int mpm_batch_destroy(mpm_batch *b) {
    if (!b) return MPM_OK;
    ctx_close(b->ctx);
    delete b;
    return MPM_OK;
}

static int check_scene_index(mpm_batch *b, uint32_t scene) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (scene >= b->ctx.batch.scenes()) return fail(MPM_ERR_INVALID_ARGUMENT, "scene out of range");
    return MPM_OK;
}

int mpm_batch_step(mpm_batch *b, uint32_t substeps) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, "batch is NULL");
    int rc = ready(b->ctx);
    if (rc) return rc;
    for (uint32_t k = 0; k < substeps; ++k) HIP_TRY(b->ctx.batch.substep(b->ctx.stream));
    return MPM_OK;
}

int mpm_batch_set_box(mpm_batch *b, uint32_t scene, const float real_box[3]) {
    int rc = check_scene_index(b, scene);
    if (rc) return rc;
    rc = check_box(real_box, "real_box");
    if (rc) return rc;
    rc = ready(b->ctx);
    if (rc) return rc;
    HIP_TRY(b->ctx.batch.set_box(scene, real_box, b->ctx.stream));
    return MPM_OK;
}

int mpm_batch_synchronize(mpm_batch *b) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, "batch is NULL");
    int rc = ready(b->ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->ctx.stream));
    return MPM_OK;
}

int mpm_batch_get_counts(mpm_batch *b, uint32_t scene, uint32_t *num_particles,
                         uint64_t *grid_count) {
    int rc = check_scene_index(b, scene);
    if (rc) return rc;
    if (num_particles) *num_particles = b->ctx.batch.scene(scene).n;
    if (grid_count) *grid_count = b->ctx.batch.scene(scene).G;
    return MPM_OK;
}

int mpm_batch_read_posvel(mpm_batch *b, uint32_t scene, void *dst, uint32_t n) {
    int rc = check_scene_index(b, scene);
    if (rc) return rc;
    if (!dst && n) return fail(MPM_ERR_INVALID_ARGUMENT, "dst is NULL");
    if (n > b->ctx.batch.scene(scene).n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds the scene's particles");
    rc = ready(b->ctx);
    if (rc) return rc;
    HIP_TRY(b->ctx.batch.read_posvel(scene, dst, n, b->ctx.stream));
    return MPM_OK;
}

int mpm_batch_read_particles(mpm_batch *b, uint32_t scene, void *dst, uint32_t n) {
    int rc = check_scene_index(b, scene);
    if (rc) return rc;
    if (!dst && n) return fail(MPM_ERR_INVALID_ARGUMENT, "dst is NULL");
    if (n > b->ctx.batch.scene(scene).n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds the scene's particles");
    rc = ready(b->ctx);
    if (rc) return rc;
    HIP_TRY(b->ctx.batch.read_particles(scene, dst, n, b->ctx.stream));
    return MPM_OK;
}

int mpm_batch_get_tuning(mpm_batch *b, mpm_tuning *out) {
    if (!b || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    to_c_tuning(b->ctx.batch.tuning(), out);
    return MPM_OK;
}

int mpm_batch_set_tuning(mpm_batch *b, const mpm_tuning *in) {
    if (!b || !in) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ready(b->ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->ctx.stream));
    return from_c_tuning(in, b->ctx.batch, b->ctx.stream);
}

int mpm_batch_autotune(mpm_batch *b, uint32_t substeps_per_trial, mpm_tuning *chosen) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, "batch is NULL");
    int rc = ready(b->ctx);
    if (rc) return rc;
    mpm::Tuning t;
    HIP_TRY(b->ctx.batch.autotune(b->ctx.stream, substeps_per_trial ? substeps_per_trial : 4, &t));
    if (chosen) to_c_tuning(t, chosen);
    return MPM_OK;
}

int mpm_batch_time_steps(mpm_batch *b, uint32_t substeps, float *ms) {
    if (!b || !ms) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    return time_substeps(b->ctx, substeps, ms);
}

int mpm_batch_time_stages(mpm_batch *b, uint32_t substeps, float *stage_ms, int *calls,
                          const char **names, int capacity, int *count) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, "batch is NULL");
    return time_stages_impl(b->ctx, substeps, stage_ms, calls, names, capacity, count);
}

}  // extern "C"
