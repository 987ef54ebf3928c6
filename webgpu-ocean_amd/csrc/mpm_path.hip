// mpm_path.hip -- production substep sequence and its hipGraph replay.
#include <cstring>
#include <utility>

#include "mpm_path.h"

namespace mpm {
namespace {

const char *const kFusedNames[5] = {"bin", "clear_grid", "p2g_1", "p2g_2", "g2p_fused"};
const char *const kStageNames[6] = {"clear_grid", "p2g_1", "p2g_2",
                                    "update_grid", "g2p", "copy_position"};

// Fields a captured graph depends on, apart from the state/id pair.
bool same_shape(const DeviceScene &a, const DeviceScene &b) {
    return a.stride == b.stride && a.n == b.n && a.grid == b.grid && a.G == b.G &&
           a.posvel == b.posvel && a.uni == b.uni &&
           std::memcmp(&a.prm, &b.prm, sizeof(Params)) == 0 &&
           std::memcmp(a.init_box, b.init_box, sizeof(a.init_box)) == 0;
}

}  // namespace

FusedArgs FusedPath::args(const DeviceScene &s) const {
    FusedArgs a{};
    a.state = s.state;
    a.id = s.id;
    a.stride = s.stride;
    a.n = s.n;
    a.grid = s.grid;
    a.G = s.G;
    a.posvel = s.posvel;
    a.uni = s.uni;
    a.prm = s.prm;
    a.ppt = tun_.ppt;
    a.wmax = tun_.wmax;
    return a;
}

hipError_t FusedPath::ensure_buffers(const DeviceScene &s) {
    KeyGeom kg = make_key_geom(s.init_box);
    if (s.stride > cap_ || kg.nkeys > key_cap_) {
        drop_graphs();
        if (spare_state_) (void)hipFree(spare_state_);
        if (spare_id_) (void)hipFree(spare_id_);
        if (key_) (void)hipFree(key_);
        if (rank_) (void)hipFree(rank_);
        if (count_) (void)hipFree(count_);
        if (offset_) (void)hipFree(offset_);
        if (block_sums_) (void)hipFree(block_sums_);
        spare_state_ = nullptr;
        spare_id_ = key_ = rank_ = count_ = offset_ = block_sums_ = nullptr;
        cap_ = 0;
        key_cap_ = 0;
        const size_t c = s.stride, k = kg.nkeys;
        hipError_t e;
        if ((e = hipMalloc(&spare_state_, sizeof(float) * NCOMP * c)) != hipSuccess) return e;
        if ((e = hipMalloc(&spare_id_, sizeof(uint32_t) * c)) != hipSuccess) return e;
        if ((e = hipMalloc(&key_, sizeof(uint32_t) * c)) != hipSuccess) return e;
        if ((e = hipMalloc(&rank_, sizeof(uint32_t) * c)) != hipSuccess) return e;
        if ((e = hipMalloc(&count_, sizeof(uint32_t) * k)) != hipSuccess) return e;
        if ((e = hipMalloc(&offset_, sizeof(uint32_t) * k)) != hipSuccess) return e;
        if ((e = hipMalloc(&block_sums_, sizeof(uint32_t) * (k / 4096 + 1))) != hipSuccess)
            return e;
        if ((e = hipMemset(count_, 0, sizeof(uint32_t) * k)) != hipSuccess) return e;
        cap_ = (uint32_t)c;
        key_cap_ = k;
    }
    kg_ = kg;
    return hipSuccess;
}

hipError_t FusedPath::launch_substep(DeviceScene &s, hipStream_t st, bool sort,
                                     hipEvent_t *ev) {
    hipError_t e;
    int k = 0;
    auto mark = [&]() -> hipError_t { return ev ? hipEventRecord(ev[k++], st) : hipSuccess; };
    if (tun_.path == 1) {
        for (int stage = 0; stage < 6; ++stage) {
            if ((e = mark()) != hipSuccess) return e;
            if ((e = launch_stage(stage, s, st)) != hipSuccess) return e;
        }
        return mark();
    }
    FusedArgs a = args(s);
    if ((e = mark()) != hipSuccess) return e;
    if (sort) {
        BinArgs b{kg_, key_, rank_, count_, offset_, block_sums_, spare_state_, spare_id_};
        if ((e = launch_bin(a, b, st)) != hipSuccess) return e;
        a.state = spare_state_;
        a.id = spare_id_;
    }
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_stage(0, s, st)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_p2g1(a, st)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_p2g2(a, st)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_g2p(a, st)) != hipSuccess) return e;
    return mark();
}

hipError_t FusedPath::substep(DeviceScene &s, hipStream_t st) {
    if (ngraphs_ && !same_shape(s, captured_)) drop_graphs();
    hipError_t e = ensure_buffers(s);
    if (e != hipSuccess) return e;
    const bool sort = tun_.path == 0 && s.n > 0 &&
                      (since_sort_ % (unsigned)tun_.sort_period) == 0;
    GraphEntry *g = nullptr;
    for (int i = 0; i < ngraphs_; ++i)
        if (graphs_[i].sort == sort && graphs_[i].state == s.state) g = &graphs_[i];
    if (!g) {
        if (ngraphs_ == 4) drop_graphs();
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        DeviceScene tmp = s;
        hipError_t le = launch_substep(tmp, st, sort, nullptr);
        e = hipStreamEndCapture(st, &graph);
        if (le != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return le;
        }
        if (e != hipSuccess) return e;
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return e;
        graphs_[ngraphs_] = GraphEntry{sort, s.state, exec};
        g = &graphs_[ngraphs_++];
        captured_ = s;
    }
    e = hipGraphLaunch(g->exec, st);
    if (e != hipSuccess) return e;
    if (sort) {
        std::swap(s.state, spare_state_);
        std::swap(s.id, spare_id_);
        since_sort_ = 0;
    }
    ++since_sort_;
    return hipSuccess;
}

hipError_t FusedPath::time_stages(DeviceScene &s, hipStream_t st, unsigned substeps, float *ms,
                                  int *calls, const char **names, int capacity, int *count) {
    const bool fused = tun_.path == 0;
    const int nslots = fused ? 5 : 6;
    if (capacity < nslots) return hipErrorInvalidValue;
    hipError_t e = ensure_buffers(s);
    if (e != hipSuccess) return e;
    hipEvent_t ev[7];
    for (int i = 0; i <= nslots; ++i)
        if ((e = hipEventCreate(&ev[i])) != hipSuccess) return e;
    for (int i = 0; i < nslots; ++i) {
        ms[i] = 0.0f;
        if (calls) calls[i] = 0;
        if (names) names[i] = fused ? kFusedNames[i] : kStageNames[i];
    }
    for (unsigned k = 0; k < substeps && e == hipSuccess; ++k) {
        const bool sort = fused && s.n > 0 && (since_sort_ % (unsigned)tun_.sort_period) == 0;
        e = launch_substep(s, st, sort, ev);
        if (e == hipSuccess) e = hipEventSynchronize(ev[nslots]);
        for (int i = 0; i < nslots && e == hipSuccess; ++i) {
            if (fused && i == 0 && !sort) continue;
            float t = 0.0f;
            e = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            ms[i] += t;
            if (calls) calls[i] += 1;
        }
        if (sort) {
            std::swap(s.state, spare_state_);
            std::swap(s.id, spare_id_);
            since_sort_ = 0;
        }
        ++since_sort_;
    }
    for (int i = 0; i <= nslots; ++i) (void)hipEventDestroy(ev[i]);
    *count = nslots;
    return e;
}

hipError_t FusedPath::set_tuning(const Tuning &t) {
    if (t.path < 0 || t.path > 1) return hipErrorInvalidValue;
    if (t.ppt != 1 && t.ppt != 2 && t.ppt != 4) return hipErrorInvalidValue;
    if (t.wmax < 27 || t.wmax > 8192) return hipErrorInvalidValue;
    if (t.sort_period < 1) return hipErrorInvalidValue;
    drop_graphs();
    tun_ = t;
    return hipSuccess;
}

void FusedPath::drop_graphs() {
    for (int i = 0; i < ngraphs_; ++i) (void)hipGraphExecDestroy(graphs_[i].exec);
    ngraphs_ = 0;
    std::memset(&captured_, 0, sizeof(captured_));
}

void FusedPath::invalidate() {
    drop_graphs();
    since_sort_ = 0;
}

void FusedPath::release() {
    drop_graphs();
    if (spare_state_) (void)hipFree(spare_state_);
    if (spare_id_) (void)hipFree(spare_id_);
    if (key_) (void)hipFree(key_);
    if (rank_) (void)hipFree(rank_);
    if (count_) (void)hipFree(count_);
    if (offset_) (void)hipFree(offset_);
    if (block_sums_) (void)hipFree(block_sums_);
    spare_state_ = nullptr;
    spare_id_ = key_ = rank_ = count_ = offset_ = block_sums_ = nullptr;
    cap_ = 0;
    key_cap_ = 0;
}

}  // namespace mpm
