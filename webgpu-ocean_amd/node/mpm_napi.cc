// mpm_napi.cc -- Node N-API addon over libmpm355's C ABIs (include/mpm.h,
// include/sph.h, include/render.h).
//
// Binds exactly the entry points the JS MLSMPMSimulator, SPHSimulator and
// FluidRenderer need; a failing call throws a JS Error carrying
// mpm_last_error() / sph_last_error() / render_last_error().  Built with g++ against
// the system Node headers (N-API >= 6 for BigInt device pointers).
#define NAPI_VERSION 6
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "mpm.h"
#include "render.h"
#include "sph.h"

namespace {

#define NAPI_CALL(env, call)                                                 \
    do {                                                                     \
        if ((call) != napi_ok) {                                             \
            napi_throw_error((env), nullptr, "N-API call failed: " #call);   \
            return nullptr;                                                  \
        }                                                                    \
    } while (0)

napi_value throw_status(napi_env env, int rc) {
    std::string msg = std::string("mpm error ") + std::to_string(rc) + ": " + mpm_last_error();
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

#define MPM_CALL(env, call)                    \
    do {                                       \
        int rc_ = (call);                      \
        if (rc_ != MPM_OK) return throw_status((env), rc_); \
    } while (0)

template <int N>
bool get_args(napi_env env, napi_callback_info info, napi_value (&argv)[N], size_t *argc) {
    *argc = N;
    return napi_get_cb_info(env, info, argc, argv, nullptr, nullptr) == napi_ok;
}

// The JS handle is an external owning a Holder, so destroy() can release the
// simulator early and the GC finalizer then finds nothing left to free.
struct Holder {
    mpm_sim *sim;
};

bool get_holder(napi_env env, napi_value v, Holder **out) {
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "expected a simulator handle");
        return false;
    }
    *out = static_cast<Holder *>(p);
    return true;
}

bool get_handle(napi_env env, napi_value v, mpm_sim **out) {
    Holder *h;
    if (!get_holder(env, v, &h)) return false;
    if (!h->sim) {
        napi_throw_error(env, nullptr, "simulator was destroyed");
        return false;
    }
    *out = h->sim;
    return true;
}

// Raw bytes of an ArrayBuffer or TypedArray (null/undefined -> empty).
bool get_bytes(napi_env env, napi_value v, void **data, size_t *len) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_null || t == napi_undefined) {
        *data = nullptr;
        *len = 0;
        return true;
    }
    bool is_ab = false, is_ta = false;
    napi_is_arraybuffer(env, v, &is_ab);
    if (is_ab) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
    napi_is_typedarray(env, v, &is_ta);
    if (is_ta) {
        napi_typedarray_type tt;
        size_t n = 0, off = 0;
        napi_value ab;
        void *base = nullptr;
        if (napi_get_typedarray_info(env, v, &tt, &n, &base, &ab, &off) != napi_ok) return false;
        size_t elem = 1;
        switch (tt) {
        case napi_int16_array: case napi_uint16_array: elem = 2; break;
        case napi_int32_array: case napi_uint32_array: case napi_float32_array: elem = 4; break;
        case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: elem = 8; break;
        default: elem = 1;
        }
        *data = base;
        *len = n * elem;
        return true;
    }
    napi_throw_type_error(env, nullptr, "expected an ArrayBuffer or TypedArray");
    return false;
}

bool get_f3(napi_env env, napi_value v, float out[3]) {
    bool is_arr = false;
    napi_is_array(env, v, &is_arr);
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    if (!is_arr && !is_ta) {
        napi_throw_type_error(env, nullptr, "expected a 3-element array");
        return false;
    }
    for (uint32_t i = 0; i < 3; ++i) {
        napi_value e;
        double d = 0;
        if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) {
            napi_throw_type_error(env, nullptr, "expected 3 numbers");
            return false;
        }
        out[i] = (float)d;
    }
    return true;
}

bool get_u32(napi_env env, napi_value v, uint32_t *out) {
    if (napi_get_value_uint32(env, v, out) != napi_ok) {
        napi_throw_type_error(env, nullptr, "expected an unsigned integer");
        return false;
    }
    return true;
}

bool get_params(napi_env env, napi_value v, mpm_params *p, bool *given) {
    mpm_default_params(p);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    *given = false;
    if (t != napi_object) return true;
    *given = true;
    const char *names[5] = {"stiffness", "restDensity", "dynamicViscosity", "dt",
                            "fixedPointMultiplier"};
    float *fields[5] = {&p->stiffness, &p->rest_density, &p->dynamic_viscosity, &p->dt,
                        &p->fixed_point_multiplier};
    for (int i = 0; i < 5; ++i) {
        bool has = false;
        napi_has_named_property(env, v, names[i], &has);
        if (!has) continue;
        napi_value e;
        double d;
        napi_get_named_property(env, v, names[i], &e);
        if (napi_get_value_double(env, e, &d) != napi_ok) {
            napi_throw_type_error(env, nullptr, "params fields must be numbers");
            return false;
        }
        *fields[i] = (float)d;
    }
    return true;
}

void finalize_sim(napi_env, void *data, void *) {
    Holder *h = static_cast<Holder *>(data);
    mpm_destroy(h->sim);
    delete h;
}

// create(params|null, particles ArrayBuffer|null, n, initBox[3]) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    size_t argc;
    if (!get_args(env, info, argv, &argc) || argc < 4) {
        napi_throw_type_error(env, nullptr, "create(params, particles, n, initBox)");
        return nullptr;
    }
    mpm_params p;
    bool given;
    void *data;
    size_t len;
    uint32_t n;
    float box[3];
    if (!get_params(env, argv[0], &p, &given) || !get_bytes(env, argv[1], &data, &len) ||
        !get_u32(env, argv[2], &n) || !get_f3(env, argv[3], box))
        return nullptr;
    if ((size_t)n * MPM_PARTICLE_STRUCT_SIZE > len) {
        napi_throw_range_error(env, nullptr, "particle buffer shorter than n x 80 bytes");
        return nullptr;
    }
    mpm_sim *sim = nullptr;
    MPM_CALL(env, mpm_create(&p, data, n, box, &sim));
    Holder *holder = new Holder{sim};
    napi_value h;
    if (napi_create_external(env, holder, finalize_sim, nullptr, &h) != napi_ok) {
        mpm_destroy(sim);
        delete holder;
        napi_throw_error(env, nullptr, "napi_create_external failed");
        return nullptr;
    }
    return h;
}

napi_value Undefined(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

// reset(h, particles, n, initBox)
napi_value Reset(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    size_t argc;
    mpm_sim *sim;
    void *data;
    size_t len;
    uint32_t n;
    float box[3];
    if (!get_args(env, info, argv, &argc) || argc < 4 || !get_handle(env, argv[0], &sim) ||
        !get_bytes(env, argv[1], &data, &len) || !get_u32(env, argv[2], &n) ||
        !get_f3(env, argv[3], box))
        return nullptr;
    if ((size_t)n * MPM_PARTICLE_STRUCT_SIZE > len) {
        napi_throw_range_error(env, nullptr, "particle buffer shorter than n x 80 bytes");
        return nullptr;
    }
    MPM_CALL(env, mpm_reset(sim, data, n, box));
    return Undefined(env);
}

// setBox(h, realBox)
napi_value SetBox(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    size_t argc;
    mpm_sim *sim;
    float box[3];
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_handle(env, argv[0], &sim) ||
        !get_f3(env, argv[1], box))
        return nullptr;
    MPM_CALL(env, mpm_set_box(sim, box));
    return Undefined(env);
}

// step(h, substeps)
napi_value Step(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    size_t argc;
    mpm_sim *sim;
    uint32_t k;
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_handle(env, argv[0], &sim) ||
        !get_u32(env, argv[1], &k))
        return nullptr;
    MPM_CALL(env, mpm_step(sim, k));
    return Undefined(env);
}

// runStage(h, stage)
napi_value RunStage(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    size_t argc;
    mpm_sim *sim;
    uint32_t k;
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_handle(env, argv[0], &sim) ||
        !get_u32(env, argv[1], &k))
        return nullptr;
    MPM_CALL(env, mpm_run_stage(sim, (int)k));
    return Undefined(env);
}

napi_value Synchronize(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    mpm_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_handle(env, argv[0], &sim))
        return nullptr;
    MPM_CALL(env, mpm_synchronize(sim));
    return Undefined(env);
}

// readPosvel(h, dst) / readParticles(h, dst) / writeParticles(h, src): n is
// taken from the simulator; the buffer must be large enough.
napi_value Transfer(napi_env env, napi_callback_info info, int which) {
    napi_value argv[2];
    size_t argc;
    mpm_sim *sim;
    void *data;
    size_t len;
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_handle(env, argv[0], &sim) ||
        !get_bytes(env, argv[1], &data, &len))
        return nullptr;
    uint32_t n = 0;
    MPM_CALL(env, mpm_get_counts(sim, &n, nullptr));
    const size_t rec = which == 0 ? MPM_POSVEL_STRUCT_SIZE : MPM_PARTICLE_STRUCT_SIZE;
    if ((size_t)n * rec > len) {
        napi_throw_range_error(env, nullptr, "buffer shorter than numParticles records");
        return nullptr;
    }
    if (which == 0) MPM_CALL(env, mpm_read_posvel(sim, data, n));
    if (which == 1) MPM_CALL(env, mpm_read_particles(sim, data, n));
    if (which == 2) MPM_CALL(env, mpm_write_particles(sim, data, n));
    return Undefined(env);
}
napi_value ReadPosvel(napi_env env, napi_callback_info info) { return Transfer(env, info, 0); }
napi_value ReadParticles(napi_env env, napi_callback_info info) { return Transfer(env, info, 1); }
napi_value WriteParticles(napi_env env, napi_callback_info info) { return Transfer(env, info, 2); }

// counts(h) -> { numParticles, gridCount }
napi_value Counts(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    mpm_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_handle(env, argv[0], &sim))
        return nullptr;
    uint32_t n = 0;
    uint64_t g = 0;
    MPM_CALL(env, mpm_get_counts(sim, &n, &g));
    napi_value o, vn, vg;
    NAPI_CALL(env, napi_create_object(env, &o));
    NAPI_CALL(env, napi_create_uint32(env, n, &vn));
    NAPI_CALL(env, napi_create_double(env, (double)g, &vg));
    NAPI_CALL(env, napi_set_named_property(env, o, "numParticles", vn));
    NAPI_CALL(env, napi_set_named_property(env, o, "gridCount", vg));
    return o;
}

// posvelDevicePtr(h) -> BigInt device address of the PosVel buffer
napi_value PosvelDevicePtr(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    mpm_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_handle(env, argv[0], &sim))
        return nullptr;
    void *p = nullptr;
    MPM_CALL(env, mpm_posvel_device_ptr(sim, &p));
    napi_value b;
    NAPI_CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &b));
    return b;
}

// timeSteps(h, frames, substepsPerFrame) -> ms (GPU events)
napi_value TimeSteps(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    size_t argc;
    mpm_sim *sim;
    uint32_t f, k;
    if (!get_args(env, info, argv, &argc) || argc < 3 || !get_handle(env, argv[0], &sim) ||
        !get_u32(env, argv[1], &f) || !get_u32(env, argv[2], &k))
        return nullptr;
    float ms = 0;
    MPM_CALL(env, mpm_time_steps(sim, f, k, &ms));
    napi_value r;
    NAPI_CALL(env, napi_create_double(env, ms, &r));
    return r;
}

// destroy(h): releases the simulator now; later calls on h throw.
napi_value Destroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    Holder *h;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_holder(env, argv[0], &h))
        return nullptr;
    mpm_sim *sim = h->sim;
    h->sim = nullptr;
    MPM_CALL(env, mpm_destroy(sim));
    return Undefined(env);
}

napi_value AbiVersion(napi_env env, napi_callback_info) {
    napi_value r;
    NAPI_CALL(env, napi_create_int32(env, mpm_abi_version(), &r));
    return r;
}

// ---- SPH (include/sph.h) ----------------------------------------------------

napi_value throw_sph(napi_env env, int rc) {
    std::string msg = std::string("sph error ") + std::to_string(rc) + ": " + sph_last_error();
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

#define SPH_CALL(env, call)                                   \
    do {                                                      \
        int rc_ = (call);                                     \
        if (rc_ != SPH_OK) return throw_sph((env), rc_);      \
    } while (0)

struct SphHolder {
    sph_sim *sim;
};

bool get_sph(napi_env env, napi_value v, SphHolder **holder, sph_sim **out) {
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "expected an SPH simulator handle");
        return false;
    }
    *holder = static_cast<SphHolder *>(p);
    if (out) {
        if (!(*holder)->sim) {
            napi_throw_error(env, nullptr, "simulator was destroyed");
            return false;
        }
        *out = (*holder)->sim;
    }
    return true;
}

// SPH parameters: JS numbers (doubles) as the reference's constants are.
bool get_sph_params(napi_env env, napi_value v, sph_params *p) {
    sph_default_params(p);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t != napi_object) return true;
    const char *names[7] = {"kernelRadius", "mass", "dt", "stiffness", "nearStiffness",
                            "restDensity", "viscosity"};
    double *fields[7] = {&p->kernel_radius, &p->mass, &p->dt, &p->stiffness,
                         &p->near_stiffness, &p->rest_density, &p->viscosity};
    for (int i = 0; i < 7; ++i) {
        bool has = false;
        napi_has_named_property(env, v, names[i], &has);
        if (!has) continue;
        napi_value e;
        napi_get_named_property(env, v, names[i], &e);
        if (napi_get_value_double(env, e, fields[i]) != napi_ok) {
            napi_throw_type_error(env, nullptr, "params fields must be numbers");
            return false;
        }
    }
    return true;
}

void finalize_sph(napi_env, void *data, void *) {
    SphHolder *h = static_cast<SphHolder *>(data);
    sph_destroy(h->sim);
    delete h;
}

// sphCreate(params|null, particles|null, n, initHalfBox[3]) -> handle
napi_value SphCreate(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    size_t argc;
    if (!get_args(env, info, argv, &argc) || argc < 4) {
        napi_throw_type_error(env, nullptr, "sphCreate(params, particles, n, initHalfBox)");
        return nullptr;
    }
    sph_params p;
    void *data;
    size_t len;
    uint32_t n;
    float box[3];
    if (!get_sph_params(env, argv[0], &p) || !get_bytes(env, argv[1], &data, &len) ||
        !get_u32(env, argv[2], &n) || !get_f3(env, argv[3], box))
        return nullptr;
    if ((size_t)n * SPH_PARTICLE_STRUCT_SIZE > len) {
        napi_throw_range_error(env, nullptr, "particle buffer shorter than n x 64 bytes");
        return nullptr;
    }
    sph_sim *sim = nullptr;
    SPH_CALL(env, sph_create(&p, data, n, box, &sim));
    SphHolder *holder = new SphHolder{sim};
    napi_value h;
    if (napi_create_external(env, holder, finalize_sph, nullptr, &h) != napi_ok) {
        sph_destroy(sim);
        delete holder;
        napi_throw_error(env, nullptr, "napi_create_external failed");
        return nullptr;
    }
    return h;
}

// sphReset(h, particles, n, initHalfBox)
napi_value SphReset(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    void *data;
    size_t len;
    uint32_t n;
    float box[3];
    if (!get_args(env, info, argv, &argc) || argc < 4 || !get_sph(env, argv[0], &hd, &sim) ||
        !get_bytes(env, argv[1], &data, &len) || !get_u32(env, argv[2], &n) ||
        !get_f3(env, argv[3], box))
        return nullptr;
    if ((size_t)n * SPH_PARTICLE_STRUCT_SIZE > len) {
        napi_throw_range_error(env, nullptr, "particle buffer shorter than n x 64 bytes");
        return nullptr;
    }
    SPH_CALL(env, sph_reset(sim, data, n, box));
    return Undefined(env);
}

// sphSetBox(h, realHalfBox) / sphStep(h, substeps) / sphSynchronize(h)
napi_value SphSetBox(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    float box[3];
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_sph(env, argv[0], &hd, &sim) ||
        !get_f3(env, argv[1], box))
        return nullptr;
    SPH_CALL(env, sph_set_box(sim, box));
    return Undefined(env);
}

napi_value SphStep(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    uint32_t k;
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_sph(env, argv[0], &hd, &sim) ||
        !get_u32(env, argv[1], &k))
        return nullptr;
    SPH_CALL(env, sph_step(sim, k));
    return Undefined(env);
}

napi_value SphSynchronize(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_sph(env, argv[0], &hd, &sim))
        return nullptr;
    SPH_CALL(env, sph_synchronize(sim));
    return Undefined(env);
}

// sphReadPosvel(h, dst) / sphReadParticles(h, dst) / sphWriteParticles(h, src)
napi_value SphTransfer(napi_env env, napi_callback_info info, int which) {
    napi_value argv[2];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    void *data;
    size_t len;
    if (!get_args(env, info, argv, &argc) || argc < 2 || !get_sph(env, argv[0], &hd, &sim) ||
        !get_bytes(env, argv[1], &data, &len))
        return nullptr;
    uint32_t n = 0;
    SPH_CALL(env, sph_get_counts(sim, &n, nullptr));
    const size_t rec = which == 0 ? SPH_POSVEL_STRUCT_SIZE : SPH_PARTICLE_STRUCT_SIZE;
    if ((size_t)n * rec > len) {
        napi_throw_range_error(env, nullptr, "buffer shorter than numParticles records");
        return nullptr;
    }
    if (which == 0) SPH_CALL(env, sph_read_posvel(sim, data, n));
    if (which == 1) SPH_CALL(env, sph_read_particles(sim, data, n));
    if (which == 2) SPH_CALL(env, sph_write_particles(sim, data, n));
    return Undefined(env);
}
napi_value SphReadPosvel(napi_env env, napi_callback_info info) { return SphTransfer(env, info, 0); }
napi_value SphReadParticles(napi_env env, napi_callback_info info) { return SphTransfer(env, info, 1); }
napi_value SphWriteParticles(napi_env env, napi_callback_info info) { return SphTransfer(env, info, 2); }

// sphCounts(h) -> { numParticles, gridCount }
napi_value SphCounts(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_sph(env, argv[0], &hd, &sim))
        return nullptr;
    uint32_t n = 0;
    uint64_t g = 0;
    SPH_CALL(env, sph_get_counts(sim, &n, &g));
    napi_value o, vn, vg;
    NAPI_CALL(env, napi_create_object(env, &o));
    NAPI_CALL(env, napi_create_uint32(env, n, &vn));
    NAPI_CALL(env, napi_create_double(env, (double)g, &vg));
    NAPI_CALL(env, napi_set_named_property(env, o, "numParticles", vn));
    NAPI_CALL(env, napi_set_named_property(env, o, "gridCount", vg));
    return o;
}

napi_value SphPosvelDevicePtr(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    SphHolder *hd;
    sph_sim *sim;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_sph(env, argv[0], &hd, &sim))
        return nullptr;
    void *p = nullptr;
    SPH_CALL(env, sph_posvel_device_ptr(sim, &p));
    napi_value b;
    NAPI_CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &b));
    return b;
}

// sphDestroy(h): releases the simulator now; later calls on h throw.
napi_value SphDestroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    SphHolder *hd;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_sph(env, argv[0], &hd, nullptr))
        return nullptr;
    sph_sim *sim = hd->sim;
    hd->sim = nullptr;
    SPH_CALL(env, sph_destroy(sim));
    return Undefined(env);
}

// ---- renderer (include/render.h) ----

#define RENDER_CALL(env, call)                                                          \
    do {                                                                                \
        int rc_ = (call);                                                               \
        if (rc_ != RENDER_OK) {                                                         \
            std::string msg = std::string("render error ") + std::to_string(rc_) + ": " + \
                              render_last_error();                                      \
            napi_throw_error((env), nullptr, msg.c_str());                              \
            return nullptr;                                                             \
        }                                                                               \
    } while (0)

bool is_exception_pending(napi_env env) {
    bool pending = false;
    napi_is_exception_pending(env, &pending);
    return pending;
}

struct RenderHolder {
    render_ctx *r;
    uint32_t width, height;
};

bool get_render(napi_env env, napi_value v, RenderHolder **out) {
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "expected a renderer handle");
        return false;
    }
    *out = static_cast<RenderHolder *>(p);
    if (!(*out)->r) {
        napi_throw_error(env, nullptr, "renderer was destroyed");
        return false;
    }
    return true;
}

void finalize_render(napi_env, void *data, void *) {
    RenderHolder *h = static_cast<RenderHolder *>(data);
    render_destroy(h->r);
    delete h;
}

// The 68 floats of the render uniform buffer (common.ts:1-9).
bool get_uniforms(napi_env env, napi_value v, render_uniforms *u) {
    void *data;
    size_t len;
    if (!get_bytes(env, v, &data, &len)) return false;
    if (len < sizeof(render_uniforms)) {
        napi_throw_range_error(env, nullptr, "render uniforms must hold 68 floats");
        return false;
    }
    std::memcpy(u, data, sizeof(render_uniforms));
    return true;
}

// renderCreate(width, height, radius, fov, cubemap Float32Array, cubeSize) -> handle
napi_value RenderCreate(napi_env env, napi_callback_info info) {
    napi_value argv[6];
    size_t argc;
    if (!get_args(env, info, argv, &argc) || argc < 6) {
        napi_throw_type_error(env, nullptr, "renderCreate(width, height, radius, fov, cubemap, size)");
        return nullptr;
    }
    render_params p;
    uint32_t size;
    void *cube;
    size_t len;
    if (!get_u32(env, argv[0], &p.width) || !get_u32(env, argv[1], &p.height) ||
        napi_get_value_double(env, argv[2], &p.radius) != napi_ok ||
        napi_get_value_double(env, argv[3], &p.fov) != napi_ok ||
        !get_bytes(env, argv[4], &cube, &len) || !get_u32(env, argv[5], &size)) {
        if (!is_exception_pending(env)) napi_throw_type_error(env, nullptr, "bad renderCreate arguments");
        return nullptr;
    }
    if ((size_t)size * size * 6 * 4 * sizeof(float) > len) {
        napi_throw_range_error(env, nullptr, "cubemap shorter than 6 x size^2 RGBA floats");
        return nullptr;
    }
    render_ctx *r = nullptr;
    RENDER_CALL(env, render_create(&p, static_cast<const float *>(cube), size, &r));
    RenderHolder *holder = new RenderHolder{r, p.width, p.height};
    napi_value h;
    if (napi_create_external(env, holder, finalize_render, nullptr, &h) != napi_ok) {
        render_destroy(r);
        delete holder;
        napi_throw_error(env, nullptr, "napi_create_external failed");
        return nullptr;
    }
    return h;
}

// renderExecute(h, posvel (BigInt device address | ArrayBuffer of host
// records), n, uniforms Float32Array(68), sphere)
napi_value RenderExecute(napi_env env, napi_callback_info info) {
    napi_value argv[5];
    size_t argc;
    RenderHolder *hd;
    uint32_t n;
    render_uniforms u;
    bool sphere = false;
    if (!get_args(env, info, argv, &argc) || argc < 5 || !get_render(env, argv[0], &hd) ||
        !get_u32(env, argv[2], &n) || !get_uniforms(env, argv[3], &u))
        return nullptr;
    napi_get_value_bool(env, argv[4], &sphere);
    napi_valuetype t;
    napi_typeof(env, argv[1], &t);
    if (t == napi_bigint) {
        uint64_t addr = 0;
        bool lossless = false;
        NAPI_CALL(env, napi_get_value_bigint_uint64(env, argv[1], &addr, &lossless));
        RENDER_CALL(env, render_execute(hd->r, (const void *)(uintptr_t)addr, n, &u, sphere));
    } else {
        void *data;
        size_t len;
        if (!get_bytes(env, argv[1], &data, &len)) return nullptr;
        if ((size_t)n * 32 > len) {
            napi_throw_range_error(env, nullptr, "PosVel buffer shorter than n x 32 bytes");
            return nullptr;
        }
        RENDER_CALL(env, render_execute_host(hd->r, data, n, &u, sphere));
    }
    return Undefined(env);
}

// renderReadImage(h, image, dst ArrayBuffer|Float32Array)
napi_value RenderReadImage(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    size_t argc;
    RenderHolder *hd;
    uint32_t image;
    void *data;
    size_t len;
    if (!get_args(env, info, argv, &argc) || argc < 3 || !get_render(env, argv[0], &hd) ||
        !get_u32(env, argv[1], &image) || !get_bytes(env, argv[2], &data, &len))
        return nullptr;
    const size_t need = (size_t)hd->width * hd->height * sizeof(float) *
                        (image == RENDER_IMAGE_COLOR ? 4 : 1);
    if (len < need) {
        napi_throw_range_error(env, nullptr, "destination shorter than the image");
        return nullptr;
    }
    RENDER_CALL(env, render_read_image(hd->r, (int)image, static_cast<float *>(data)));
    return Undefined(env);
}

// renderColorDevicePtr(h) -> BigInt
napi_value RenderColorDevicePtr(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    RenderHolder *hd;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_render(env, argv[0], &hd))
        return nullptr;
    void *p = nullptr;
    RENDER_CALL(env, render_color_device_ptr(hd->r, &p));
    napi_value b;
    NAPI_CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &b));
    return b;
}

// renderTimeFrames(h, posvel BigInt, n, uniforms, sphere, frames) -> ms
napi_value RenderTimeFrames(napi_env env, napi_callback_info info) {
    napi_value argv[6];
    size_t argc;
    RenderHolder *hd;
    uint32_t n, frames;
    render_uniforms u;
    bool sphere = false, lossless = false;
    uint64_t addr = 0;
    if (!get_args(env, info, argv, &argc) || argc < 6 || !get_render(env, argv[0], &hd) ||
        !get_u32(env, argv[2], &n) || !get_uniforms(env, argv[3], &u) ||
        !get_u32(env, argv[5], &frames))
        return nullptr;
    NAPI_CALL(env, napi_get_value_bigint_uint64(env, argv[1], &addr, &lossless));
    napi_get_value_bool(env, argv[4], &sphere);
    float ms = 0.0f;
    RENDER_CALL(env, render_time_frames(hd->r, (const void *)(uintptr_t)addr, n, &u, sphere,
                                        frames, &ms));
    napi_value out;
    NAPI_CALL(env, napi_create_double(env, ms, &out));
    return out;
}

// renderDestroy(h): releases the renderer now; later calls on h throw.
napi_value RenderDestroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc;
    RenderHolder *hd;
    if (!get_args(env, info, argv, &argc) || argc < 1 || !get_render(env, argv[0], &hd))
        return nullptr;
    render_ctx *r = hd->r;
    hd->r = nullptr;
    RENDER_CALL(env, render_destroy(r));
    return Undefined(env);
}

napi_value Init(napi_env env, napi_value exports) {
    struct {
        const char *name;
        napi_callback cb;
    } fns[] = {
        {"create", Create},           {"reset", Reset},
        {"setBox", SetBox},           {"step", Step},
        {"runStage", RunStage},       {"synchronize", Synchronize},
        {"readPosvel", ReadPosvel},   {"readParticles", ReadParticles},
        {"writeParticles", WriteParticles}, {"counts", Counts},
        {"posvelDevicePtr", PosvelDevicePtr}, {"timeSteps", TimeSteps},
        {"destroy", Destroy},         {"abiVersion", AbiVersion},
        {"sphCreate", SphCreate},     {"sphReset", SphReset},
        {"sphSetBox", SphSetBox},     {"sphStep", SphStep},
        {"sphSynchronize", SphSynchronize}, {"sphReadPosvel", SphReadPosvel},
        {"sphReadParticles", SphReadParticles}, {"sphWriteParticles", SphWriteParticles},
        {"sphCounts", SphCounts},     {"sphPosvelDevicePtr", SphPosvelDevicePtr},
        {"sphDestroy", SphDestroy},
        {"renderCreate", RenderCreate}, {"renderExecute", RenderExecute},
        {"renderReadImage", RenderReadImage}, {"renderColorDevicePtr", RenderColorDevicePtr},
        {"renderTimeFrames", RenderTimeFrames}, {"renderDestroy", RenderDestroy},
    };
    for (auto &f : fns) {
        napi_value fn;
        NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
        NAPI_CALL(env, napi_set_named_property(env, exports, f.name, fn));
    }
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
