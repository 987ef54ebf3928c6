// mpm_batch.hip -- buffers, chunk table, hipGraph replay and autotuner of a
// batch of scenes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>

#include "mpm_batch.h"

namespace mpm {
namespace {

constexpr int kFusedSlots = 8;
const char *const kFusedNames[kFusedSlots] = {"bin_gather", "bin_scan", "bin_scatter",
                                              "p2g_1",      "reduce_1", "p2g_2",
                                              "reduce_2",   "g2p"};
const char *const kStageNames[6] = {"clear_grid", "p2g_1", "p2g_2",
                                    "update_grid", "g2p", "copy_position"};

int32_t f2i_sat_host(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

template <typename T>
void free_ptr(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

#define TRY(call)                          \
    do {                                   \
        hipError_t e_ = (call);            \
        if (e_ != hipSuccess) return e_;   \
    } while (0)

}  // namespace

// Uniforms derived from the two boxes (mls-mpm/p2g_1.wgsl:53-56,
// updateGrid.wgsl:37-43).
Uniforms derive_uniforms(const float init_box[3], const float real_box[3]) {
    Uniforms u{};
    for (int a = 0; a < 3; ++a) {
        u.init_box[a] = init_box[a];
        u.real_box[a] = real_box[a];
        u.wall_hi[a] = f2i_sat_host(std::ceil(real_box[a]) - 3.0f);
    }
    u.Y = f2i_sat_host(init_box[1]);
    u.Z = f2i_sat_host(init_box[2]);
    u.pad = 0;
    return u;
}

namespace {

// Device allocation of the batch buffers.  MPM355_FAIL_ALLOC_AFTER=k (a
// fault-injection knob for the failure-handling tests) makes the allocation
// after the first k of one (re)allocation pass fail as out of memory.
struct Allocator {
    long budget = -1;
    Allocator() {
        if (const char *e = std::getenv("MPM355_FAIL_ALLOC_AFTER")) budget = std::atol(e);
    }
    template <typename T>
    hipError_t operator()(T **p, size_t bytes) {
        if (budget == 0) return hipErrorOutOfMemory;
        if (budget > 0) --budget;
        return hipMalloc(p, bytes);
    }
};

}  // namespace

hipError_t Batch::ensure_capacity(uint32_t total_n, uint64_t cells, uint64_t keys,
                                  uint64_t tchunks, uint64_t partial_cells) {
    const uint32_t cap = std::max<uint32_t>(64u, (total_n + 63u) & ~63u);
    if (cap <= cap_n_ && cells <= cap_cells_ && keys <= cap_keys_ && tchunks <= cap_tchunks_ &&
        partial_cells <= cap_partial_)
        return hipSuccess;
    Allocator alloc;
    drop_graphs();
    free_ptr(state_);
    free_ptr(spare_state_);
    free_ptr(id_);
    free_ptr(spare_id_);
    free_ptr(grid_);
    free_ptr(vel_);
    free_ptr(wild_);
    free_ptr(partial_);
    free_ptr(posvel_);
    free_ptr(staging_);
    free_ptr(key_);
    free_ptr(rank_);
    free_ptr(src_);
    free_ptr(mig_);
    free_ptr(order_);
    free_ptr(perm_);
    free_ptr(count_);
    free_ptr(offset_);
    free_ptr(sup_);
    free_ptr(tchunks_);
    free_ptr(ntchunks_);
    free_ptr(tile_chunk_);
    cap_n_ = 0;
    cap_cells_ = cap_keys_ = cap_tchunks_ = cap_partial_ = 0;
    const size_t c = cap, g = std::max<uint64_t>(cells, 1), k = std::max<uint64_t>(keys, 1),
                 tc = std::max<uint64_t>(tchunks, 1),
                 pc = std::max<uint64_t>(partial_cells, (uint64_t)kMaxWinCells);
    TRY(alloc(&state_, sizeof(float) * NCOMP * c));
    state_buf0_ = state_;
    TRY(alloc(&spare_state_, sizeof(float) * NCOMP * c));
    TRY(alloc(&id_, sizeof(uint32_t) * c));
    TRY(alloc(&spare_id_, sizeof(uint32_t) * c));
    TRY(alloc(&grid_, sizeof(int4) * g));
    TRY(alloc(&vel_, sizeof(float4) * g));
    TRY(alloc(&wild_, sizeof(int4) * g));
    TRY(hipMemset(wild_, 0, sizeof(int4) * g));
    TRY(alloc(&partial_, sizeof(int4) * pc));
    TRY(alloc(&posvel_, 2 * sizeof(float4) * c));
    TRY(alloc(&staging_, 80 * c));
    TRY(alloc(&key_, sizeof(uint32_t) * c));
    TRY(alloc(&rank_, sizeof(uint32_t) * c));
    TRY(alloc(&src_, sizeof(uint32_t) * c));
    TRY(alloc(&mig_, sizeof(uint32_t) * 27 * tc));
    TRY(alloc(&order_, sizeof(uint32_t) * c));
    TRY(alloc(&perm_, sizeof(uint32_t) * c));
    TRY(hipMemset(perm_, 0, sizeof(uint32_t) * c));  // every entry a valid slot
    TRY(alloc(&count_, sizeof(uint32_t) * k));
    TRY(alloc(&offset_, sizeof(uint32_t) * k));
    const size_t nsup = (k + (1u << kSupBits) - 1) >> kSupBits;
    TRY(alloc(&sup_, sizeof(uint2) * 2 * nsup));  // two substep parities
    TRY(hipMemset(sup_, 0, sizeof(uint2) * 2 * nsup));
    sup_cap_ = nsup;
    TRY(alloc(&tchunks_, sizeof(TileChunk) * tc));
    // chunk count, wild chunks of the last key, then the kXcds + 1 XCD range
    // boundaries of the chunk walk
    TRY(alloc(&ntchunks_, sizeof(uint32_t) * kWalkHeadWords));
    TRY(hipMemset(ntchunks_, 0, sizeof(uint32_t) * kWalkHeadWords));
    TRY(alloc(&tile_chunk_, sizeof(uint2) * k));
    TRY(hipMemset(count_, 0, sizeof(uint32_t) * k));
    TRY(hipMemset(ntchunks_, 0, sizeof(uint32_t)));
    cap_n_ = cap;
    cap_cells_ = g;
    cap_keys_ = k;
    cap_tchunks_ = tc;
    cap_partial_ = pc;
    return hipSuccess;
}

// A failed (re)initialisation releases everything: the batch then holds no
// scene (ready() is false) rather than scene records over missing buffers.
hipError_t Batch::init(const SceneInit *sc, uint32_t count, hipStream_t st) {
    const hipError_t e = init_impl(sc, count, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        release();
    }
    return e;
}

void Batch::layout_keys(int tb, uint64_t *keys, uint64_t *tiles) {
    uint64_t koff = 0, soff = 0;
    for (SceneDesc &d : h_scenes_) {
        scene_key_geometry(d, tb);
        d.koff = (uint32_t)koff;
        d.toff = (uint32_t)koff;
        d.soff = (uint32_t)soff;
        d.rank_off = 0;  // set by apply_tile_geometry
        koff += d.nkeys;
        soff += d.ntiles;
    }
    *keys = koff;
    *tiles = soff;
}

hipError_t Batch::apply_tile_geometry(int tb) {
    uint64_t keys = 0, tiles = 0;
    layout_keys(tb, &keys, &tiles);
    total_keys_ = keys;
    total_tiles_ = (uint32_t)tiles;
    const uint32_t count = (uint32_t)h_scenes_.size();
    std::vector<uint32_t> ranks, xmaj, nbr;
    std::vector<uint4> info;
    for (uint32_t s = 0; s < count; ++s) {
        SceneDesc &d = h_scenes_[s];
        d.rank_off = (uint32_t)ranks.size();
        const size_t x0 = xmaj.size();
        append_hilbert_ranks(d, ranks, xmaj);
        append_tile_tables(d, s, ranks.data() + d.rank_off, xmaj.data() + x0, nbr, info);
    }
    if (info.size() > tile_tab_cap_) {
        free_ptr(d_nbr_);
        free_ptr(d_tile_info_);
        free_ptr(d_dirty_);
        free_ptr(d_near_);
        tile_tab_cap_ = 0;
        TRY(hipMalloc(&d_nbr_, sizeof(uint32_t) * nbr.size()));
        TRY(hipMalloc(&d_tile_info_, sizeof(uint4) * info.size()));
        TRY(hipMalloc(&d_dirty_, info.size()));
        TRY(hipMalloc(&d_near_, 2 * info.size()));
        tile_tab_cap_ = info.size();
    }
    if (!info.empty()) {
        TRY(hipMemset(d_dirty_, 1, info.size()));
        TRY(hipMemset(d_near_, 0, 2 * tile_tab_cap_));
    }
    // per key: its scene (keys of scene s are [koff_s, koff_s + nkeys_s))
    std::vector<uint32_t> key_scene(xmaj.size());
    for (uint32_t s = 0; s < count; ++s)
        for (uint32_t k = 0; k < h_scenes_[s].nkeys; ++k) key_scene[h_scenes_[s].koff + k] = s;
    if (key_scene.size() > key_scene_cap_) {
        free_ptr(d_key_scene_);
        key_scene_cap_ = 0;
        TRY(hipMalloc(&d_key_scene_, sizeof(uint32_t) * key_scene.size()));
        key_scene_cap_ = key_scene.size();
    }
    TRY(hipMemcpy(d_key_scene_, key_scene.data(), sizeof(uint32_t) * key_scene.size(),
                  hipMemcpyHostToDevice));
    if (!info.empty()) {
        TRY(hipMemcpy(d_nbr_, nbr.data(), sizeof(uint32_t) * nbr.size(), hipMemcpyHostToDevice));
        TRY(hipMemcpy(d_tile_info_, info.data(), sizeof(uint4) * info.size(),
                      hipMemcpyHostToDevice));
    }
    if (ranks.size() > tile_rank_cap_) {
        free_ptr(d_tile_rank_);
        tile_rank_cap_ = 0;
        TRY(hipMalloc(&d_tile_rank_, sizeof(uint32_t) * ranks.size()));
        tile_rank_cap_ = ranks.size();
    }
    if (xmaj.size() > tile_xmaj_cap_) {
        free_ptr(d_tile_xmaj_);
        tile_xmaj_cap_ = 0;
        TRY(hipMalloc(&d_tile_xmaj_, sizeof(uint32_t) * xmaj.size()));
        tile_xmaj_cap_ = xmaj.size();
    }
    if (!ranks.empty())
        TRY(hipMemcpy(d_tile_rank_, ranks.data(), sizeof(uint32_t) * ranks.size(),
                      hipMemcpyHostToDevice));
    TRY(hipMemcpy(d_tile_xmaj_, xmaj.data(), sizeof(uint32_t) * xmaj.size(),
                  hipMemcpyHostToDevice));
    TRY(hipMemcpy(d_scenes_, h_scenes_.data(), sizeof(SceneDesc) * count,
                  hipMemcpyHostToDevice));
    keys_ready_ = false;  // keys and counts were taken for the other tiles
    return hipSuccess;
}

hipError_t Batch::init_impl(const SceneInit *sc, uint32_t count, hipStream_t st) {
    TRY(hipStreamSynchronize(st));  // (graphs stay: they match only their own arguments)
    h_scenes_.assign(count, SceneDesc{});
    uint64_t poff = 0, goff = 0;
    for (uint32_t s = 0; s < count; ++s) {
        SceneDesc &d = h_scenes_[s];
        d.uni = derive_uniforms(sc[s].init_box, sc[s].init_box);
        d.prm = sc[s].prm;
        d.poff = (uint32_t)poff;
        d.n = sc[s].n;
        d.goff = (uint32_t)goff;
        d.G = (uint32_t)((uint64_t)std::ceil(sc[s].init_box[0]) *
                         (uint64_t)std::ceil(sc[s].init_box[1]) *
                         (uint64_t)std::ceil(sc[s].init_box[2]));
        poff += d.n;
        goff += d.G;
    }
    total_n_ = (uint32_t)poff;
    total_cells_ = goff;
    // Buffers for both tile sizes, so that the autotuner can switch without
    // reallocating.  Tile chunks of one binning: sum over keys of
    // ceil(count / 256) <= n / 256 + (non-empty keys).
    uint64_t keys2 = 0, tiles2 = 0, keys3 = 0, tiles3 = 0;
    layout_keys(3, &keys3, &tiles3);
    layout_keys(2, &keys2, &tiles2);
    const uint64_t tc2 = (poff + kBlock - 1) / kBlock + std::min<uint64_t>(poff, keys2);
    const uint64_t tc3 = (poff + kBlock - 1) / kBlock + std::min<uint64_t>(poff, keys3);
    TRY(ensure_capacity(total_n_, total_cells_, std::max(keys2, keys3), std::max(tc2, tc3),
                        std::max(tc2 * (uint64_t)TileGeo<2>::win_cells,
                                 tc3 * (uint64_t)TileGeo<3>::win_cells)));
    stride_ = cap_n_;
    if (count > scene_cap_) {
        free_ptr(d_scenes_);
        scene_cap_ = 0;
        TRY(hipMalloc(&d_scenes_, sizeof(SceneDesc) * count));
        scene_cap_ = count;
    }
    TRY(apply_tile_geometry(tun_.tile_bits));
    TRY(hipMemsetAsync(grid_, 0, sizeof(int4) * std::max<uint64_t>(total_cells_, 1), st));
    TRY(hipMemsetAsync(posvel_, 0, 2 * sizeof(float4) * (size_t)stride_, st));
    for (uint32_t s = 0; s < count; ++s)
        if (sc[s].n)
            TRY(hipMemcpyAsync(staging_ + 20 * (size_t)h_scenes_[s].poff, sc[s].particles,
                               80 * (size_t)sc[s].n, hipMemcpyHostToDevice, st));
    DeviceScene all = stage_view();
    all.n = total_n_;
    TRY(launch_unpack_particles(staging_, all, st));
    keys_ready_ = false;
    TRY(upload_chunks());
    return hipStreamSynchronize(st);
}

hipError_t Batch::upload_chunks() {
    const uint32_t per = (uint32_t)kBlock;
    h_chunks_.clear();
    for (uint32_t s = 0; s < h_scenes_.size(); ++s) {
        const SceneDesc &d = h_scenes_[s];
        for (uint32_t b = 0; b < d.n; b += per)
            h_chunks_.push_back(Chunk{s, d.poff + b, std::min(per, d.n - b), 0});
    }
    if (h_chunks_.size() > chunk_cap_) {
        free_ptr(d_chunks_);
        TRY(hipMalloc(&d_chunks_, sizeof(Chunk) * h_chunks_.size()));
        chunk_cap_ = h_chunks_.size();
    }
    if (!h_chunks_.empty())
        TRY(hipMemcpy(d_chunks_, h_chunks_.data(), sizeof(Chunk) * h_chunks_.size(),
                      hipMemcpyHostToDevice));
    return hipSuccess;
}

DeviceScene Batch::stage_view() const {
    DeviceScene v{};
    v.state = state_;
    v.id = id_;
    v.stride = stride_;
    v.grid = grid_;
    v.posvel = posvel_;
    if (!h_scenes_.empty()) {
        const SceneDesc &d = h_scenes_[0];
        v.n = d.n;
        v.G = d.G;
        v.prm = d.prm;
        v.uni = reinterpret_cast<Uniforms *>(reinterpret_cast<char *>(d_scenes_) +
                                             offsetof(SceneDesc, uni));
        for (int a = 0; a < 3; ++a) v.init_box[a] = d.uni.init_box[a];
    }
    return v;
}

// Tile chunks one binning may emit: the chunk buffers' capacity, bounded by
// the partial windows the partial buffer holds at the current tile size (a
// binning emits at most n / 256 + (non-empty keys) chunks, which both bounds
// cover for either tile size).
uint32_t Batch::tchunk_cap() const {
    const uint64_t wc = tun_.tile_bits == 3 ? TileGeo<3>::win_cells : TileGeo<2>::win_cells;
    return (uint32_t)std::min<uint64_t>(cap_tchunks_, cap_partial_ / wc);
}

BatchArgs Batch::args() const {
    BatchArgs a{};
    a.scenes = d_scenes_;
    a.nscenes = (uint32_t)h_scenes_.size();
    a.tile_rank = d_tile_rank_;
    a.tile_xmaj = d_tile_xmaj_;
    a.chunks = d_chunks_;
    a.nchunks = (uint32_t)h_chunks_.size();
    a.tchunks = tchunks_;
    a.ntchunks = ntchunks_;
    a.xcd_first = ntchunks_ + 2;
    a.total_slots = total_n_;
    a.tchunk_cap = tchunk_cap();
    a.tile_chunk = tile_chunk_;
    a.stiles = total_tiles_;
    a.nbr = d_nbr_;
    a.tile_info = d_tile_info_;
    a.dirty = d_dirty_;
    {
        const size_t par = state_ == state_buf0_ ? 0 : 1;  // substep parity
        a.near = d_near_ + par * tile_tab_cap_;
        a.near_prev = d_near_ + (1 - par) * tile_tab_cap_;
    }
    a.key_scene = d_key_scene_;
    a.state = state_;
    a.id = id_;
    a.perm = perm_;
    a.state_out = spare_state_;
    a.id_out = spare_id_;
    a.stride = stride_;
    a.grid = grid_;
    a.vel = vel_;
    a.wild = wild_;
    a.key = key_;
    a.rank = rank_;
    a.src = src_;
    a.mig = mig_;
    a.order = order_;
    a.count = count_;
    a.count_next = 1;
    a.partial = partial_;
    a.posvel = posvel_;
    a.ppt = tun_.ppt;
    a.grid_blocks = tun_.grid_blocks;
    a.xcd_remap = tun_.xcd_remap;
    a.tile_bits = tun_.tile_bits;
    return a;
}

BinBuffers Batch::bin_buffers() const {
    BinBuffers b{};
    b.count = count_;
    b.offset = offset_;
    // totals of this substep's parity (state buffer 0 or 1), the other one next
    const size_t par = state_ == state_buf0_ ? 0 : 1;
    b.nsup = (uint32_t)((total_keys_ + (1u << kSupBits) - 1) >> kSupBits);
    b.sup = sup_ + par * sup_cap_;
    b.sup_next = sup_ + (1 - par) * sup_cap_;
    b.nkeys = (uint32_t)total_keys_;
    b.perm_out = perm_;
    b.tchunks = tchunks_;
    b.tchunk_cap = tchunk_cap();
    b.ntchunks = ntchunks_;
    b.xcd_first = ntchunks_ + 2;
    b.tile_chunk = tile_chunk_;
    return b;
}

hipError_t Batch::set_box(uint32_t s, const float real_box[3], hipStream_t st) {
    SceneDesc &d = h_scenes_[s];
    d.uni = derive_uniforms(d.uni.init_box, real_box);
    return launch_set_scene_uniforms(d_scenes_ + s, d.uni, st);
}

// The binned path re-bins every substep: the tile chunks are only valid for
// the positions they were built from.  Its g2p leaves the keys and counts of
// the new state for the next binning.  The reference-shaped kernels change
// particles and cells behind the binned path's back.
void Batch::after_substep(hipStream_t st) {
    if (tun_.path == 0) {
        std::swap(state_, spare_state_);
        std::swap(id_, spare_id_);
        keys_ready_ = total_n_ > 0;
    } else {
        (void)external_change(st);
    }
}

hipError_t Batch::external_change(hipStream_t st) {
    keys_ready_ = false;
    if (total_tiles_ == 0) return hipSuccess;
    return hipMemsetAsync(d_dirty_, 1, total_tiles_, st);
}

// With `t`, each slot of the production path arms kernel-exact events for its
// kernel (hipExtLaunchKernel timestamps); the reference-shaped path records
// events around each stage.
hipError_t Batch::launch_substep(hipStream_t st, SlotTimes *t) {
    int k = 0;
    if (!t) (void)disarm_kernel_timer();
    auto before = [&]() -> hipError_t {
        if (!t) return hipSuccess;
        if (tun_.path == 1) return hipEventRecord(t->start[k], st);
        arm_kernel_timer(t->start[k], t->stop[k]);
        return hipSuccess;
    };
    auto after = [&]() -> hipError_t {
        if (!t) return hipSuccess;
        hipError_t e = hipSuccess;
        if (tun_.path == 1) {
            e = hipEventRecord(t->stop[k], st);
            t->used[k] = true;
        } else {
            t->used[k] = !disarm_kernel_timer();  // false: the slot launched nothing
        }
        ++k;
        return e;
    };
    if (tun_.path == 1) {
        const DeviceScene v = stage_view();
        for (int stage = 0; stage < 6; ++stage) {
            TRY(before());
            TRY(launch_stage(stage, v, st));
            TRY(after());
        }
        return hipSuccess;
    }
    BatchArgs a = args();
    const BinBuffers b = bin_buffers();
    TRY(before());
    TRY(launch_bin_gather(a, b, !keys_ready_, st));
    TRY(after());
    TRY(before());
    TRY(launch_bin_scan(a, b, st));
    TRY(after());
    TRY(before());
    TRY(launch_bin_scatter(a, b, st));
    TRY(after());
    TRY(before());
    TRY(launch_p2g1(a, st));
    TRY(after());
    TRY(before());
    TRY(launch_reduce1(a, st));
    TRY(after());
    TRY(before());
    TRY(launch_p2g2(a, st));
    TRY(after());
    TRY(before());
    TRY(launch_reduce2(a, st));
    TRY(after());
    TRY(before());
    TRY(launch_g2p(a, st));
    TRY(after());
    return hipSuccess;
}

hipError_t Batch::substep(hipStream_t st) {
    GraphEntry *g = nullptr;
    const BatchArgs a = args();
    const BinBuffers b = bin_buffers();
    const DeviceScene v = stage_view();
    for (int i = 0; i < ngraphs_; ++i)
        if (graphs_[i].keys_ready == keys_ready_ && graphs_[i].path == tun_.path &&
            std::memcmp(&graphs_[i].args, &a, sizeof a) == 0 &&
            std::memcmp(&graphs_[i].bins, &b, sizeof b) == 0 &&
            std::memcmp(&graphs_[i].stage, &v, sizeof v) == 0)
            g = &graphs_[i];
    if (!g) {
        if (ngraphs_ == kMaxGraphs) drop_graphs();
        hipGraph_t graph = nullptr;
        TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_substep(st, nullptr);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (le != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return le;
        }
        if (e != hipSuccess) return e;
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return e;
        graphs_[ngraphs_] = GraphEntry{a, b, v, tun_.path, keys_ready_, exec};
        g = &graphs_[ngraphs_++];
    }
    graph_stream_ = st;
    graph_launched_ = true;
    TRY(hipGraphLaunch(g->exec, st));
    after_substep(st);
    return hipSuccess;
}

hipError_t Batch::time_stages(hipStream_t st, unsigned substeps, float *ms, int *calls,
                              const char **names, int capacity, int *count) {
    const bool fused = tun_.path == 0;
    const int nslots = fused ? kFusedSlots : 6;
    static_assert(kFusedSlots <= SlotTimes::kMax, "slot events");
    if (capacity < nslots) return hipErrorInvalidValue;
    SlotTimes t{};
    hipError_t e = hipSuccess;
    int made = 0;  // slots whose two events exist
    while (made < nslots) {
        e = hipEventCreate(&t.start[made]);
        if (e != hipSuccess) break;
        e = hipEventCreate(&t.stop[made]);
        if (e != hipSuccess) {
            (void)hipEventDestroy(t.start[made]);
            break;
        }
        ++made;
    }
    for (int i = 0; i < nslots; ++i) {
        ms[i] = 0.0f;
        if (calls) calls[i] = 0;
        if (names) names[i] = fused ? kFusedNames[i] : kStageNames[i];
    }
    for (unsigned k = 0; k < substeps && e == hipSuccess; ++k) {
        e = launch_substep(st, &t);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        for (int i = 0; i < nslots && e == hipSuccess; ++i) {
            if (!t.used[i]) continue;
            float dt = 0.0f;
            e = hipEventElapsedTime(&dt, t.start[i], t.stop[i]);
            ms[i] += dt;
            if (calls) calls[i] += 1;
        }
        after_substep(st);
    }
    (void)disarm_kernel_timer();
    for (int i = 0; i < made; ++i) {
        (void)hipEventDestroy(t.start[i]);
        (void)hipEventDestroy(t.stop[i]);
    }
    *count = nslots;
    return e;
}

hipError_t Batch::read_posvel(uint32_t s, void *dst, uint32_t n, hipStream_t st) {
    if (n)
        TRY(hipMemcpyAsync(dst, posvel_ptr(s), 32 * (size_t)n, hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

hipError_t Batch::read_particles(uint32_t s, void *dst, uint32_t n, hipStream_t st) {
    DeviceScene all = stage_view();
    all.n = total_n_;
    TRY(launch_pack_particles(staging_, all, st));
    if (n)
        TRY(hipMemcpyAsync(dst, staging_ + 20 * (size_t)h_scenes_[s].poff, 80 * (size_t)n,
                           hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

hipError_t Batch::write_particles(uint32_t s, const void *src, uint32_t n, hipStream_t st) {
    // Every scene returns to caller order: pack all, replace one scene, unpack all.
    DeviceScene all = stage_view();
    all.n = total_n_;
    TRY(launch_pack_particles(staging_, all, st));
    if (n)
        TRY(hipMemcpyAsync(staging_ + 20 * (size_t)h_scenes_[s].poff, src, 80 * (size_t)n,
                           hipMemcpyHostToDevice, st));
    TRY(launch_unpack_particles(staging_, all, st));
    keys_ready_ = false;
    return hipStreamSynchronize(st);
}

hipError_t Batch::read_grid(uint32_t s, int32_t *dst, uint64_t ncells, hipStream_t st) {
    if (ncells)
        TRY(hipMemcpyAsync(dst, grid_ + h_scenes_[s].goff, 16 * (size_t)ncells,
                           hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

hipError_t Batch::read_vel(uint32_t s, float *dst, uint64_t ncells, hipStream_t st) {
    if (ncells)
        TRY(hipMemcpyAsync(dst, vel_ + h_scenes_[s].goff, 16 * (size_t)ncells,
                           hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

hipError_t Batch::write_grid(uint32_t s, const int32_t *src, uint64_t ncells, hipStream_t st) {
    if (ncells)
        TRY(hipMemcpyAsync(grid_ + h_scenes_[s].goff, src, 16 * (size_t)ncells,
                           hipMemcpyHostToDevice, st));
    TRY(external_change(st));
    return hipStreamSynchronize(st);
}

hipError_t Batch::set_tuning(const Tuning &t, hipStream_t st) {
    if (t.path < 0 || t.path > 1) return hipErrorInvalidValue;
    if (t.ppt != 1 && t.ppt != 2 && t.ppt != 4) return hipErrorInvalidValue;
    if (t.grid_blocks < 0 || t.grid_blocks > (1 << 20) || t.xcd_remap < 0 || t.xcd_remap > 1)
        return hipErrorInvalidValue;
    if (t.tile_bits != 2 && t.tile_bits != 3) return hipErrorInvalidValue;
    if (t.path == 1 && h_scenes_.size() > 1) return hipErrorInvalidValue;
    drop_graphs();
    if (t.tile_bits != tun_.tile_bits && ready()) {
        TRY(hipStreamSynchronize(st));  // the tables change under no running kernel
        tun_.tile_bits = t.tile_bits;
        const hipError_t e = apply_tile_geometry(t.tile_bits);
        if (e != hipSuccess) return e;
    }
    tun_ = t;
    keys_ready_ = false;  // counts were taken for the previous chunk size
    return hipSuccess;
}

hipError_t Batch::autotune(hipStream_t st, unsigned substeps_per_trial, Tuning *chosen) {
    TRY(hipStreamSynchronize(st));
    // Snapshot of everything a substep changes.
    const size_t sb = sizeof(float) * NCOMP * stride_, ib = sizeof(uint32_t) * stride_,
                 gb = sizeof(int4) * std::max<uint64_t>(total_cells_, 1),
                 pb = 2 * sizeof(float4) * (size_t)stride_;
    char *snap = nullptr;
    TRY(hipMalloc(&snap, sb + ib + gb + pb));
    auto save = [&]() -> hipError_t {
        TRY(hipMemcpyAsync(snap, state_, sb, hipMemcpyDeviceToDevice, st));
        TRY(hipMemcpyAsync(snap + sb, id_, ib, hipMemcpyDeviceToDevice, st));
        TRY(hipMemcpyAsync(snap + sb + ib, grid_, gb, hipMemcpyDeviceToDevice, st));
        return hipMemcpyAsync(snap + sb + ib + gb, posvel_, pb, hipMemcpyDeviceToDevice, st);
    };
    const Tuning orig = tun_;
    float *const orig_state = state_;
    hipError_t e = save();
    Tuning best = orig;
    float best_ms = 1e30f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int tbs[2] = {2, 3};
    const int ppts[3] = {1, 2, 4};
    const int blocks[3] = {1024, 2048, 4096};
    const int remaps[2] = {1, 0};
    for (int ti = 0; ti < 2 && e == hipSuccess; ++ti)
    for (int pi = 0; pi < 3 && e == hipSuccess; ++pi)
        for (int bi = 0; bi < 3 && e == hipSuccess; ++bi)
            for (int ri = 0; ri < 2 && e == hipSuccess; ++ri) {
                Tuning t = orig;
                t.path = 0;
                t.tile_bits = tbs[ti];
                t.ppt = ppts[pi];
                t.grid_blocks = blocks[bi];
                t.xcd_remap = remaps[ri];
                if ((e = set_tuning(t, st)) != hipSuccess) break;
                // restore the snapshot into the current buffers
                if (state_ != orig_state) {
                    std::swap(state_, spare_state_);
                    std::swap(id_, spare_id_);
                }
                e = hipMemcpyAsync(state_, snap, sb, hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = hipMemcpyAsync(id_, snap + sb, ib, hipMemcpyDeviceToDevice, st);
                keys_ready_ = false;
                // timed as production runs: graph-replayed substeps (the
                // warm-up substeps capture this candidate's graphs)
                for (unsigned k = 0; k < 3 && e == hipSuccess; ++k) e = substep(st);
                if (e == hipSuccess) e = hipEventRecord(e0, st);
                for (unsigned k = 0; k < substeps_per_trial && e == hipSuccess; ++k)
                    e = substep(st);
                if (e == hipSuccess) e = hipEventRecord(e1, st);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess && ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
    // restore state, tuning and binning phase
    if (state_ != orig_state) {
        std::swap(state_, spare_state_);
        std::swap(id_, spare_id_);
    }
    hipError_t r = hipMemcpyAsync(state_, snap, sb, hipMemcpyDeviceToDevice, st);
    if (r == hipSuccess) r = hipMemcpyAsync(id_, snap + sb, ib, hipMemcpyDeviceToDevice, st);
    keys_ready_ = false;
    if (r == hipSuccess) r = hipMemcpyAsync(grid_, snap + sb + ib, gb, hipMemcpyDeviceToDevice, st);
    if (r == hipSuccess)
        r = hipMemcpyAsync(posvel_, snap + sb + ib + gb, pb, hipMemcpyDeviceToDevice, st);
    if (r == hipSuccess) r = external_change(st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    (void)hipFree(snap);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e == hipSuccess) e = r;
    if (e != hipSuccess) {
        (void)set_tuning(orig, st);
        return e;
    }
    e = set_tuning(best_ms < 1e30f ? best : orig, st);
    if (chosen) *chosen = tun_;
    return e;
}

void Batch::drop_graphs() {
    if (ngraphs_ && graph_launched_) (void)hipStreamSynchronize(graph_stream_);
    graph_launched_ = false;
    for (int i = 0; i < ngraphs_; ++i) (void)hipGraphExecDestroy(graphs_[i].exec);
    ngraphs_ = 0;
}

void Batch::release() {
    drop_graphs();
    free_ptr(adj_ws_);
    adj_ws_cap_ = 0;
    free_ptr(state_);
    free_ptr(spare_state_);
    free_ptr(id_);
    free_ptr(spare_id_);
    free_ptr(grid_);
    free_ptr(vel_);
    free_ptr(wild_);
    free_ptr(partial_);
    free_ptr(posvel_);
    free_ptr(staging_);
    free_ptr(key_);
    free_ptr(rank_);
    free_ptr(src_);
    free_ptr(mig_);
    free_ptr(order_);
    free_ptr(perm_);
    free_ptr(count_);
    free_ptr(offset_);
    free_ptr(sup_);
    free_ptr(tchunks_);
    free_ptr(ntchunks_);
    free_ptr(tile_chunk_);
    free_ptr(d_scenes_);
    free_ptr(d_chunks_);
    free_ptr(d_tile_rank_);
    free_ptr(d_tile_xmaj_);
    free_ptr(d_nbr_);
    free_ptr(d_tile_info_);
    free_ptr(d_dirty_);
    free_ptr(d_near_);
    free_ptr(d_key_scene_);
    key_scene_cap_ = 0;
    tile_rank_cap_ = tile_xmaj_cap_ = tile_tab_cap_ = 0;
    cap_n_ = 0;
    cap_cells_ = cap_keys_ = cap_tchunks_ = cap_partial_ = 0;
    scene_cap_ = 0;
    chunk_cap_ = 0;
    h_scenes_.clear();
    h_chunks_.clear();
}

}  // namespace mpm
