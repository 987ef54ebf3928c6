// mpm_path.hip -- production substep sequence and its hipGraph replay.
#include <cstring>

#include "mpm_path.h"

namespace mpm {
namespace {

const char *const kStageNames[6] = {"clear_grid", "p2g_1", "p2g_2",
                                    "update_grid", "g2p", "copy_position"};

bool same_scene(const DeviceScene &a, const DeviceScene &b) {
    return std::memcmp(&a, &b, sizeof(DeviceScene)) == 0;
}

}  // namespace

hipError_t FusedPath::launch_sequence(const DeviceScene &s, hipStream_t st, hipEvent_t *ev,
                                      int *nev) {
    int k = 0;
    for (int stage = 0; stage < 6; ++stage) {
        if (ev) {
            hipError_t e = hipEventRecord(ev[k++], st);
            if (e != hipSuccess) return e;
        }
        hipError_t e = launch_stage(stage, s, st);
        if (e != hipSuccess) return e;
    }
    if (ev) {
        hipError_t e = hipEventRecord(ev[k++], st);
        if (e != hipSuccess) return e;
    }
    if (nev) *nev = k;
    return hipSuccess;
}

hipError_t FusedPath::substep(const DeviceScene &s, hipStream_t st) {
    if (exec_ && !same_scene(s, captured_)) invalidate();
    if (!exec_) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t le = launch_sequence(s, st, nullptr, nullptr);
        e = hipStreamEndCapture(st, &graph);
        if (le != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return le;
        }
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&exec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            exec_ = nullptr;
            return e;
        }
        captured_ = s;
    }
    return hipGraphLaunch(exec_, st);
}

hipError_t FusedPath::time_stages(const DeviceScene &s, hipStream_t st, unsigned substeps,
                                  float *ms, const char **names, int capacity, int *count) {
    const int nslots = 6;
    if (capacity < nslots) return hipErrorInvalidValue;
    hipEvent_t ev[nslots + 1];
    for (int i = 0; i <= nslots; ++i) {
        hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) return e;
    }
    for (int i = 0; i < nslots; ++i) {
        ms[i] = 0.0f;
        if (names) names[i] = kStageNames[i];
    }
    hipError_t e = hipSuccess;
    for (unsigned k = 0; k < substeps && e == hipSuccess; ++k) {
        int nev = 0;
        e = launch_sequence(s, st, ev, &nev);
        if (e == hipSuccess) e = hipEventSynchronize(ev[nslots]);
        for (int i = 0; i < nslots && e == hipSuccess; ++i) {
            float t = 0.0f;
            e = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            ms[i] += t;
        }
    }
    for (int i = 0; i <= nslots; ++i) (void)hipEventDestroy(ev[i]);
    *count = nslots;
    return e;
}

void FusedPath::invalidate() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    std::memset(&captured_, 0, sizeof(captured_));
}

void FusedPath::release() { invalidate(); }

}  // namespace mpm
