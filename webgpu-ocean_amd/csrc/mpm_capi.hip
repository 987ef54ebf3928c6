// mpm_capi.hip -- the C ABI (include/mpm.h): handle lifetime, buffers,
// uniforms and the per-frame step.  Mirrors MLSMPMSimulator
// (mls-mpm/mls-mpm.ts:12-276); nothing here throws across the boundary.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "mpm.h"
#include "mpm_internal.h"
#include "mpm_path.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? MPM_ERR_OUT_OF_MEMORY : MPM_ERR_DEVICE, \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

int32_t f2i_sat_host(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

}  // namespace

struct mpm_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    mpm::Params prm{};
    mpm::Uniforms uni{};
    mpm::DeviceScene s{};
    uint32_t cap = 0;          // particle capacity (state stride)
    uint64_t grid_cap = 0;     // allocated cells
    float *staging = nullptr;  // 80-B records, device
    mpm::FusedPath path;       // production substep
};

namespace {

int select_device(mpm_sim *sim) {
    HIP_TRY(hipSetDevice(sim->device));
    return MPM_OK;
}

int check_box(const float b[3], const char *what) {
    if (!b) return fail(MPM_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b[a]))
            return fail(MPM_ERR_INVALID_ARGUMENT, std::string(what) + " must be finite");
    return MPM_OK;
}

// Uniforms derived from the two boxes (mls-mpm/p2g_1.wgsl:53-56,
// updateGrid.wgsl:37-43).
void derive_uniforms(mpm::Uniforms &u, const float init_box[3], const float real_box[3]) {
    for (int a = 0; a < 3; ++a) {
        u.init_box[a] = init_box[a];
        u.real_box[a] = real_box[a];
        u.wall_hi[a] = f2i_sat_host(std::ceil(real_box[a]) - 3.0f);
    }
    u.Y = f2i_sat_host(init_box[1]);
    u.Z = f2i_sat_host(init_box[2]);
    u.pad = 0;
}

void free_buffers(mpm_sim *sim) {
    if (sim->s.state) (void)hipFree(sim->s.state);
    if (sim->s.id) (void)hipFree(sim->s.id);
    if (sim->s.grid) (void)hipFree(sim->s.grid);
    if (sim->s.posvel) (void)hipFree(sim->s.posvel);
    if (sim->staging) (void)hipFree(sim->staging);
    sim->s.state = nullptr;
    sim->s.id = nullptr;
    sim->s.grid = nullptr;
    sim->s.posvel = nullptr;
    sim->staging = nullptr;
    sim->cap = 0;
    sim->grid_cap = 0;
}

// (Re)initialises a scene: reference reset(), mls-mpm/mls-mpm.ts:225-243.
int reset_scene(mpm_sim *sim, const void *particles, uint32_t n, const float init_box[3]) {
    int rc = check_box(init_box, "init_box");
    if (rc) return rc;
    if (n && !particles) return fail(MPM_ERR_INVALID_ARGUMENT, "particles is NULL");
    for (int a = 0; a < 3; ++a)
        if (init_box[a] < 1.0f || init_box[a] > 1048576.0f)
            return fail(MPM_ERR_INVALID_ARGUMENT, "init_box components must lie in [1, 2^20]");
    // gridCount = ceil(bx)*ceil(by)*ceil(bz) (:229); cell indices are i32.
    uint64_t G = (uint64_t)std::ceil(init_box[0]) * (uint64_t)std::ceil(init_box[1]) *
                 (uint64_t)std::ceil(init_box[2]);
    if (G > (uint64_t)INT32_MAX)
        return fail(MPM_ERR_GRID_TOO_LARGE, "gridCount " + std::to_string(G) +
                                                " exceeds the 2^31-1 cells addressable by i32 "
                                                "cell indices");
    rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->stream));
    sim->path.invalidate();

    uint32_t cap = (n + 63u) & ~63u;
    if (cap == 0) cap = 64;
    if (cap > sim->cap || G > sim->grid_cap) {
        free_buffers(sim);
        HIP_TRY(hipMalloc(&sim->s.state, sizeof(float) * mpm::NCOMP * (size_t)cap));
        HIP_TRY(hipMalloc(&sim->s.id, sizeof(uint32_t) * (size_t)cap));
        HIP_TRY(hipMalloc(&sim->s.posvel, 2 * sizeof(float4) * (size_t)cap));
        HIP_TRY(hipMalloc(&sim->staging, 80 * (size_t)cap));
        HIP_TRY(hipMalloc(&sim->s.grid, sizeof(int4) * (size_t)(G ? G : 1)));
        sim->cap = cap;
        sim->grid_cap = G ? G : 1;
    }
    sim->s.stride = sim->cap;
    sim->s.n = n;
    sim->s.G = (uint32_t)G;
    sim->s.prm = sim->prm;
    for (int a = 0; a < 3; ++a) sim->s.init_box[a] = init_box[a];
    derive_uniforms(sim->uni, init_box, init_box);
    HIP_TRY(mpm::launch_set_uniforms(sim->s.uni, sim->uni, sim->stream));
    HIP_TRY(hipMemsetAsync(sim->s.grid, 0, sizeof(int4) * (size_t)sim->s.G, sim->stream));
    HIP_TRY(hipMemsetAsync(sim->s.posvel, 0, 2 * sizeof(float4) * (size_t)sim->cap, sim->stream));
    if (n) {
        HIP_TRY(hipMemcpyAsync(sim->staging, particles, 80 * (size_t)n, hipMemcpyHostToDevice,
                               sim->stream));
        HIP_TRY(mpm::launch_unpack_particles(sim->staging, sim->s, sim->stream));
    }
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

}  // namespace

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

int mpm_create(const mpm_params *params, const void *particles, uint32_t n,
               const float init_box[3], mpm_sim **out) {
    if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    mpm_params p;
    if (params) {
        p = *params;
    } else {
        mpm_default_params(&p);
    }
    const float vals[5] = {p.stiffness, p.rest_density, p.dynamic_viscosity, p.dt,
                           p.fixed_point_multiplier};
    for (float f : vals)
        if (!std::isfinite(f)) return fail(MPM_ERR_INVALID_ARGUMENT, "params must be finite");
    if (!(p.fixed_point_multiplier > 0.0f) || !(p.rest_density != 0.0f))
        return fail(MPM_ERR_INVALID_ARGUMENT,
                    "fixed_point_multiplier must be > 0 and rest_density non-zero");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(MPM_ERR_NO_DEVICE, "no HIP device is visible");
    mpm_sim *sim = new (std::nothrow) mpm_sim();
    if (!sim) return fail(MPM_ERR_OUT_OF_MEMORY, "host allocation failed");
    sim->prm = mpm::Params{p.stiffness, p.rest_density, p.dynamic_viscosity, p.dt,
                           p.fixed_point_multiplier};
    int rc = MPM_OK;
    hipError_t e = hipGetDevice(&sim->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&sim->s.uni, sizeof(mpm::Uniforms));
    if (e != hipSuccess) {
        rc = fail(e == hipErrorOutOfMemory ? MPM_ERR_OUT_OF_MEMORY : MPM_ERR_DEVICE,
                  std::string("device setup: ") + hipGetErrorString(e));
    } else {
        rc = reset_scene(sim, particles, n, init_box);
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
    (void)hipSetDevice(sim->device);
    if (sim->stream) (void)hipStreamSynchronize(sim->stream);
    sim->path.release();
    free_buffers(sim);
    if (sim->s.uni) (void)hipFree(sim->s.uni);
    if (sim->stream) (void)hipStreamDestroy(sim->stream);
    delete sim;
    return MPM_OK;
}

int mpm_reset(mpm_sim *sim, const void *particles, uint32_t n, const float init_box[3]) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    return reset_scene(sim, particles, n, init_box);
}

int mpm_set_box(mpm_sim *sim, const float real_box[3]) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = check_box(real_box, "real_box");
    if (rc) return rc;
    rc = select_device(sim);
    if (rc) return rc;
    derive_uniforms(sim->uni, sim->uni.init_box, real_box);
    HIP_TRY(mpm::launch_set_uniforms(sim->s.uni, sim->uni, sim->stream));
    return MPM_OK;
}

int mpm_step(mpm_sim *sim, uint32_t substeps) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = select_device(sim);
    if (rc) return rc;
    for (uint32_t k = 0; k < substeps; ++k) {
        hipError_t e = sim->path.substep(sim->s, sim->stream);
        if (e != hipSuccess)
            return fail(MPM_ERR_DEVICE, std::string("substep: ") + hipGetErrorString(e));
    }
    return MPM_OK;
}

int mpm_run_stage(mpm_sim *sim, int stage) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    if (stage < 0 || stage >= MPM_STAGE_COUNT)
        return fail(MPM_ERR_INVALID_ARGUMENT, "stage out of range");
    int rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(mpm::launch_stage(stage, sim->s, sim->stream));
    return MPM_OK;
}

int mpm_synchronize(mpm_sim *sim) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_read_posvel(mpm_sim *sim, void *dst, uint32_t n) {
    if (!sim || (!dst && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > sim->s.n) return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds numParticles");
    int rc = select_device(sim);
    if (rc) return rc;
    if (n)
        HIP_TRY(hipMemcpyAsync(dst, sim->s.posvel, 32 * (size_t)n, hipMemcpyDeviceToHost,
                               sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_posvel_device_ptr(mpm_sim *sim, void **out) {
    if (!sim || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = sim->s.posvel;
    return MPM_OK;
}

int mpm_read_particles(mpm_sim *sim, void *dst, uint32_t n) {
    if (!sim || (!dst && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > sim->s.n) return fail(MPM_ERR_INVALID_ARGUMENT, "n exceeds numParticles");
    int rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(mpm::launch_pack_particles(sim->staging, sim->s, sim->stream));
    if (n)
        HIP_TRY(hipMemcpyAsync(dst, sim->staging, 80 * (size_t)n, hipMemcpyDeviceToHost,
                               sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_write_particles(mpm_sim *sim, const void *src, uint32_t n) {
    if (!sim || (!src && n)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n != sim->s.n)
        return fail(MPM_ERR_INVALID_ARGUMENT, "n must equal numParticles (use mpm_reset)");
    int rc = select_device(sim);
    if (rc) return rc;
    if (n) {
        HIP_TRY(hipMemcpyAsync(sim->staging, src, 80 * (size_t)n, hipMemcpyHostToDevice,
                               sim->stream));
        HIP_TRY(mpm::launch_unpack_particles(sim->staging, sim->s, sim->stream));
    }
    sim->path.particles_changed();
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_read_grid(mpm_sim *sim, int32_t *dst, uint64_t ncells) {
    if (!sim || (!dst && ncells)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (ncells > sim->s.G) return fail(MPM_ERR_INVALID_ARGUMENT, "ncells exceeds gridCount");
    int rc = select_device(sim);
    if (rc) return rc;
    if (ncells)
        HIP_TRY(hipMemcpyAsync(dst, sim->s.grid, 16 * (size_t)ncells, hipMemcpyDeviceToHost,
                               sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_write_grid(mpm_sim *sim, const int32_t *src, uint64_t ncells) {
    if (!sim || (!src && ncells)) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (ncells > sim->s.G) return fail(MPM_ERR_INVALID_ARGUMENT, "ncells exceeds gridCount");
    int rc = select_device(sim);
    if (rc) return rc;
    if (ncells)
        HIP_TRY(hipMemcpyAsync(sim->s.grid, src, 16 * (size_t)ncells, hipMemcpyHostToDevice,
                               sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return MPM_OK;
}

int mpm_get_counts(mpm_sim *sim, uint32_t *num_particles, uint64_t *grid_count) {
    if (!sim) return fail(MPM_ERR_INVALID_ARGUMENT, "sim is NULL");
    if (num_particles) *num_particles = sim->s.n;
    if (grid_count) *grid_count = sim->s.G;
    return MPM_OK;
}

int mpm_time_steps(mpm_sim *sim, uint32_t frames, uint32_t substeps_per_frame, float *ms) {
    if (!sim || !ms) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = select_device(sim);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    HIP_TRY(hipEventRecord(e0, sim->stream));
    for (uint32_t f = 0; f < frames; ++f) {
        rc = mpm_step(sim, substeps_per_frame);
        if (rc) break;
    }
    HIP_TRY(hipEventRecord(e1, sim->stream));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int mpm_time_stages(mpm_sim *sim, uint32_t substeps, float *stage_ms, int *calls,
                    const char **names, int capacity, int *count) {
    if (!sim || !stage_ms || !count) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->stream));
    hipError_t e = sim->path.time_stages(sim->s, sim->stream, substeps, stage_ms, calls, names,
                                         capacity, count);
    if (e != hipSuccess)
        return fail(MPM_ERR_DEVICE, std::string("time_stages: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_get_tuning(mpm_sim *sim, mpm_tuning *out) {
    if (!sim || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    const mpm::Tuning &t = sim->path.tuning();
    out->path = t.path;
    out->particles_per_thread = t.ppt;
    out->window_cells = t.wmax;
    out->sort_period = t.sort_period;
    return MPM_OK;
}

int mpm_set_tuning(mpm_sim *sim, const mpm_tuning *in) {
    if (!sim || !in) return fail(MPM_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = select_device(sim);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sim->stream));
    mpm::Tuning t;
    t.path = in->path;
    t.ppt = in->particles_per_thread;
    t.wmax = in->window_cells;
    t.sort_period = in->sort_period;
    if (sim->path.set_tuning(t) != hipSuccess)
        return fail(MPM_ERR_INVALID_ARGUMENT,
                    "tuning out of range: path 0|1, particles_per_thread 1|2|4, "
                    "window_cells 27..8192, sort_period >= 1");
    return MPM_OK;
}

}  // extern "C"
