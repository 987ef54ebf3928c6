// chunk_index_emu.cpp -- host emulation of the production substep's index
// arithmetic (test infrastructure; built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_chunk_index_emu.py).
//
// The CPU oracle (oracle/mpm_oracle.c) advances a scene substep by substep;
// its positions are bit-identical to the library's, so every binning
// decision the gfx950 kernels take is reproduced here.  At every substep the
// program rebuilds on the host the state the kernels build -- keys, g2p's
// migration counts and ranks, k_bin_gather's running offsets, the key scan's
// tile chunks and XCD ranges, the permutation -- with the kernels' own index
// functions (webgpu-ocean_amd/csrc/mpm_index.h) and tables (mpm_tables.cpp),
// and checks, against the capacities Batch::init allocates, that:
//   - every index a kernel forms lies inside the buffer it addresses;
//   - the chunk walks of every launch shape visit every tile chunk once, and
//     the reduce launch visits every tile once;
//   - the scan's chunks tile the particle slots in order, each chunk holds
//     only particles of its tile, and the permutation is a bijection;
//   - each per-tile reduce plan names partial windows that contain the cell.
// Any violation aborts with a message (exit 2); sanitizer reports abort too.
//
// usage: chunk_index_emu particles.bin n bx by bz substeps ppt threads
//            [substep:real_x ...]   (box x changes, like mpm_set_box)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mpm_index.h"
#include "mpm_oracle.h"

using namespace mpm;

namespace {

constexpr uint32_t kFinal = 0xffffffffu;
constexpr uint32_t kBlockThreads = 256;
constexpr uint32_t kUnset = 0xffffffffu;

long g_checks = 0;

#define CHECK(cond, ...)                                             \
    do {                                                             \
        ++g_checks;                                                  \
        if (!(cond)) {                                               \
            std::fprintf(stderr, "index check failed: %s: ", #cond); \
            std::fprintf(stderr, __VA_ARGS__);                       \
            std::fprintf(stderr, "\n");                              \
            std::exit(2);                                            \
        }                                                            \
    } while (0)

// The P2G kernels' level order of a chunk with cnt[c] particles in tile cell
// c (k_p2g1's wave-0 computation and level_position): every particle's
// position lies in [0, total) and no two particles share one.
void check_level_order(const uint32_t cnt[kTileCells]) {
    uint32_t cmax = 0, total = 0;
    for (uint32_t c = 0; c < kTileCells; ++c) {
        cmax = std::max(cmax, cnt[c]);
        total += cnt[c];
    }
    const uint32_t levels = std::min(cmax, kLevelCap);
    uint32_t base[kLevelCap] = {0}, exoff[kTileCells] = {0};
    uint64_t mask[kLevelCap] = {0};
    uint32_t run = 0;
    for (uint32_t k = 0; k < levels; ++k) {
        uint64_t m = 0;
        for (uint32_t c = 0; c < kTileCells; ++c)
            if (cnt[c] > k) m |= 1ull << c;
        mask[k] = m;
        base[k] = run;
        run += (uint32_t)__builtin_popcountll(m);
    }
    uint32_t ex_run = run;  // exclusive scan of the excess, cell order
    for (uint32_t c = 0; c < kTileCells; ++c) {
        exoff[c] = ex_run;
        ex_run += cnt[c] > kLevelCap ? cnt[c] - kLevelCap : 0u;
    }
    CHECK(ex_run == total, "level order covers %u of %u positions", ex_run, total);
    std::vector<uint8_t> used(total, 0);
    for (uint32_t c = 0; c < kTileCells; ++c)
        for (uint32_t lev = 0; lev < cnt[c]; ++lev) {
            const uint32_t q = level_position(lev, c, base, mask, exoff);
            CHECK(q < total, "level position %u of %u (cell %u level %u)", q, total, c, lev);
            CHECK(!used[q], "level position %u taken twice", q);
            used[q] = 1;
        }
}

// Synthetic count patterns: all cells below the level cap, one crowded cell,
// several crowded cells, every particle in one cell.
void check_level_order_patterns(uint32_t cap_particles) {
    uint32_t cnt[kTileCells];
    for (uint32_t pat = 0; pat < 6; ++pat) {
        uint32_t left = cap_particles;
        for (uint32_t c = 0; c < kTileCells; ++c) {
            uint32_t want = 0;
            switch (pat) {
            case 0: want = 8; break;
            case 1: want = c == 5 ? cap_particles : 3; break;
            case 2: want = (c % 9 == 0) ? 60 : 2; break;
            case 3: want = c == 63 ? cap_particles : 0; break;
            case 4: want = c == 0 ? kLevelCap + 1 : (c == 1 ? kLevelCap : 1); break;
            default: want = (c * 7) % 41; break;
            }
            cnt[c] = std::min(want, left);
            left -= cnt[c];
        }
        check_level_order(cnt);
    }
}

struct Emu {
    SceneDesc d{};
    std::vector<uint32_t> rank_of, xmaj, nbr;
    std::vector<uint4> info;
    uint32_t n = 0, cap = 0, cap_n = 0, nkeys = 0, stiles = 0;
    uint64_t cap_tchunks = 0;
    // binning state
    std::vector<uint32_t> key, rnk, src;      // per storage slot (left by g2p)
    std::vector<uint32_t> count, offset;      // per key
    std::vector<uint32_t> mig;                // 27 per tile chunk
    std::vector<uint4> tchunks;               // (scene, x-major tile, first slot, count)
    std::vector<uint2> tile_chunk;            // per key: (first chunk, chunks)
    std::vector<uint8_t> near_prev, near_now;  // per tile
    std::vector<uint32_t> xcd_first;          // kXcds + 1
    uint32_t ntchunks = 0;
    std::vector<uint32_t> perm;               // bin slot -> storage slot
    std::vector<uint32_t> stor;               // storage slot -> caller index
    bool keys_ready = false;

    void init(uint32_t n_, const float box[3], int ppt) {
        n = n_;
        cap = kBlockThreads * (uint32_t)ppt;
        for (int a = 0; a < 3; ++a) {
            d.uni.init_box[a] = d.uni.real_box[a] = box[a];
        }
        d.uni.Y = (int32_t)box[1];
        d.uni.Z = (int32_t)box[2];
        d.n = n;
        d.G = (uint32_t)((uint64_t)std::ceil(box[0]) * (uint64_t)std::ceil(box[1]) *
                         (uint64_t)std::ceil(box[2]));
        scene_key_geometry(d, kTileBits);
        append_hilbert_ranks(d, rank_of, xmaj);
        append_tile_tables(d, 0, rank_of.data(), xmaj.data(), nbr, info);
        nkeys = d.nkeys;
        stiles = d.ntiles;
        // capacities as Batch::init allocates them
        cap_n = std::max<uint32_t>(64u, (n + 63u) & ~63u);
        cap_tchunks = (n + kBlockThreads - 1) / kBlockThreads + std::min<uint64_t>(n, nkeys);
        key.assign(n, 0);
        rnk.assign(n, 0);
        src.assign(n, kFinal);
        count.assign(nkeys, 0);
        offset.assign(nkeys, 0);
        mig.assign(27 * cap_tchunks, 0);
        tchunks.assign(cap_tchunks, make_uint4(0, 0, 0, 0));
        tile_chunk.assign(nkeys, make_uint2(0, 0));
        near_prev.assign(stiles, 0);
        near_now.assign(stiles, 0);
        xcd_first.assign(kXcds + 1, 0);
        perm.assign(n, 0);
        stor.resize(n);
        for (uint32_t i = 0; i < n; ++i) stor[i] = i;
        CHECK(d.fast, "the emulation covers integral boxes");
    }

    const float *pos(const std::vector<oracle_particle> &p, uint32_t slot) const {
        return p[stor[slot]].position;
    }

    // k_bin_count (state replaced from outside): keys and final ranks.
    void count_from_scratch(const std::vector<oracle_particle> &p) {
        std::fill(count.begin(), count.end(), 0u);
        const int nowhere[3] = {-4, -4, -4};
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t dir;
            key[i] = bin_key_of(rank_of.data(), &d, pos(p, i), nowhere, &dir);
            CHECK(key[i] < nkeys, "slot %u key %u", i, key[i]);
            rnk[i] = count[key[i]]++;
            src[i] = kFinal;
        }
    }

    // k_bin_gather: per key, the counts of the 27 neighbour tiles' source
    // chunks turned into running offsets, in the kernel's order (rounds over
    // a tile's chunks, lanes = neighbours within a round).
    void gather(bool final_counts) {
        for (uint32_t k = 0; k < nkeys; ++k) {
            uint32_t total = count[k];
            const uint32_t rank = k - d.koff;
            if (!final_counts && rank < d.ntiles && near_prev[d.soff + rank]) {
                const uint32_t b = d.soff + rank;
                uint32_t first[27], cnt[27], rounds = 0;
                for (uint32_t lane = 0; lane < 27; ++lane) {
                    CHECK((size_t)b * 27 + lane < nbr.size(), "nbr index");
                    const uint32_t s = nbr[(size_t)b * 27 + lane];
                    first[lane] = cnt[lane] = 0;
                    if (s != kNoTile) {
                        CHECK(s < nkeys, "neighbour key %u", s);
                        const uint2 tc = tile_chunk[s];
                        first[lane] = tc.x;
                        cnt[lane] = tc.x <= cap_tchunks && tc.y <= cap_tchunks - tc.x ? tc.y : 0;
                        CHECK(cnt[lane] == tc.y, "tile_chunk guard dropped a range (%u,%u)", tc.x,
                              tc.y);
                    }
                    rounds = std::max(rounds, cnt[lane]);
                }
                for (uint32_t j = 0; j < rounds; ++j)
                    for (uint32_t lane = 0; lane < 27; ++lane) {
                        if (j >= cnt[lane]) continue;
                        const size_t e = (size_t)(first[lane] + j) * 27 + (26 - lane);
                        CHECK(e < mig.size(), "mig index %zu of %zu", e, mig.size());
                        const uint32_t v = mig[e];
                        mig[e] = total;
                        total += v;
                    }
            }
            count[k] = total;
        }
    }

    // k_scan_keys: offsets, tile chunks, XCD ranges, reach flags.
    void scan() {
        uint32_t off = 0, cb = 0;
        std::fill(xcd_first.begin(), xcd_first.end(), 0u);
        for (uint32_t k = 0; k < nkeys; ++k) {
            const uint32_t c = count[k], nch = chunks_of(c, cap);
            offset[k] = off;
            tile_chunk[k] = make_uint2(cb, nch);
            for (uint32_t i = 0; i < nch; ++i) {
                const uint32_t z = off + i * cap, zn = std::min(cap, c - i * cap);
                CHECK(cb + i < cap_tchunks, "tile chunk %u beyond capacity %llu", cb + i,
                      (unsigned long long)cap_tchunks);
                CHECK(z + zn <= n, "chunk slots [%u, %u) beyond %u", z, z + zn, n);
                tchunks[cb + i] = make_uint4(0u, xmaj[k - d.koff], z, zn);
                for (uint32_t x = 1; x < kXcds; ++x) {
                    const uint32_t sx = xcd_split_slot(x, n);
                    if (z < sx && sx <= z + zn) xcd_first[x] = cb + i + 1;
                }
            }
            off += c;
            cb += nch;
        }
        CHECK(off == n, "counts sum to %u, not %u", off, n);
        ntchunks = cb;
        xcd_first[0] = 0;
        xcd_first[kXcds] = cb;
        for (uint32_t x = 0; x < kXcds; ++x)
            CHECK(xcd_first[x] <= xcd_first[x + 1], "XCD ranges not ordered at %u", x);
        // reach flags (every tile key)
        const bool wild = count[d.koff + d.ntiles] != 0;
        for (uint32_t r = 0; r < d.ntiles; ++r) {
            bool any = count[d.koff + r] != 0 || wild;
            for (uint32_t q = 0; q < 27 && !any; ++q) {
                const uint32_t s = nbr[(size_t)(d.soff + r) * 27 + q];
                any = s != kNoTile && count[s] != 0;
            }
            near_now[d.soff + r] = any;
        }
    }

    // k_bin_scatter: the permutation.
    void scatter() {
        std::fill(perm.begin(), perm.end(), kUnset);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t r = rnk[i];
            if (src[i] != kFinal) {
                CHECK(src[i] < mig.size(), "src %u", src[i]);
                r += mig[src[i]];
            }
            CHECK(key[i] < nkeys, "key %u", key[i]);
            const uint32_t dd = offset[key[i]] + r;
            CHECK(dd < n, "slot %u goes to %u of %u", i, dd, n);
            CHECK(perm[dd] == kUnset, "bin slot %u written twice", dd);
            perm[dd] = i;
        }
    }

    // Walks of the chunk kernels: every launch shape visits each chunk once.
    void check_walks() const {
        const uint32_t grids[] = {8, 24, 128, 1024, 2048, 4096, 8192};
        for (uint32_t grid : grids)
            for (int remap = 0; remap < 2; ++remap) {
                std::vector<uint32_t> visits(ntchunks, 0);
                const uint32_t nn = std::min<uint32_t>(ntchunks, (uint32_t)cap_tchunks);
                for (uint32_t block = 0; block < grid; ++block) {
                    // the descriptor a block loads before it knows the count
                    const uint32_t first = walk_first_of(block, grid);
                    const uint32_t cap = (uint32_t)cap_tchunks;
                    CHECK(spec_chunk_of(first, cap) < cap_tchunks,
                          "speculative descriptor %u of %llu", spec_chunk_of(first, cap),
                          (unsigned long long)cap_tchunks);
                    CHECK(spec_chunk_of(first, 0) == 0, "speculative descriptor at capacity 0");
                    const TileWalk w = tile_walk_of(block, grid, nn, xcd_first.data(), remap);
                    if (!remap) CHECK(w.c == first, "first chunk of block %u", block);
                    CHECK(w.end <= nn, "walk end %u beyond %u", w.end, nn);
                    for (uint32_t c = w.c; c < w.end; c += w.step) {
                        CHECK(c < ntchunks, "walk reaches chunk %u of %u", c, ntchunks);
                        ++visits[c];
                    }
                }
                for (uint32_t c = 0; c < ntchunks; ++c)
                    CHECK(visits[c] == 1, "grid %u remap %d visits chunk %u %u times", grid,
                          remap, c, visits[c]);
            }
        // the per-tile kernels: every tile once, the rest padding
        for (int remap = 0; remap < 2; ++remap) {
            std::vector<uint32_t> seen(stiles, 0);
            const uint32_t nb = tile_blocks(stiles);
            for (uint32_t block = 0; block < nb; ++block)
                for (uint32_t wave = 0; wave < kTileWaves; ++wave) {
                    const uint32_t b = reduce_tile_of(block, wave, stiles, remap);
                    CHECK(b < nb * kTileWaves, "reduce tile %u beyond the launch", b);
                    if (b < stiles) ++seen[b];
                }
            for (uint32_t b = 0; b < stiles; ++b)
                CHECK(seen[b] == 1, "reduce remap %d visits tile %u %u times", remap, b, seen[b]);
        }
    }

    // p2g_1 / p2g_2 / g2p index arithmetic of every tile chunk; per chunk
    // also the level order's bounds.
    void check_chunks(const std::vector<oracle_particle> &p) const {
        for (uint32_t c = 0; c < ntchunks; ++c) {
            const uint4 ch = tchunks[c];
            CHECK(ch.w >= 1 && ch.w <= cap, "chunk %u holds %u particles (cap %u)", c, ch.w, cap);
            CHECK((size_t)c * kWinCells + kWinCells <= (size_t)cap_tchunks * kWinCells,
                  "partial window of chunk %u", c);
            if (ch.y == kWildTile) {
                for (uint32_t j = 0; j < ch.w; ++j) CHECK(ch.z + j < n, "wild slot");
                continue;
            }
            int t[3];
            CHECK(ch.y < d.ntiles, "chunk tile %u", ch.y);
            tile_coords(ch.y, &d, t);
            uint32_t cells[kTileCells] = {0};
            for (uint32_t j = 0; j < ch.w; ++j) {
                CHECK(ch.z + j < n, "slot %u of chunk %u", ch.z + j, c);
                const uint32_t i = perm[ch.z + j];
                const float *x = pos(p, i);
                int cx[3];
                for (int a = 0; a < 3; ++a) {
                    cx[a] = (int)std::floor(x[a]);
                    CHECK((cx[a] >> kTileBits) == t[a], "particle %u (x %g) in chunk of tile %d",
                          stor[i], (double)x[a], t[a]);
                }
                const int l0 = win_l0(x);
                const int far = (2 * kWinEdge + 2) * kWinEdge + 2;  // node (2,2,2)
                CHECK(l0 >= 0 && l0 + far < kWinCells, "window node %d", l0 + far);
                const int cell = (((cx[0] & 3) * kTileEdge) + (cx[1] & 3)) * kTileEdge + (cx[2] & 3);
                CHECK(cell >= 0 && cell < kTileCells, "level-order cell %d", cell);
                ++cells[cell];
            }
            uint32_t positions = 0;
            for (uint32_t q = 0; q < kTileCells; ++q) positions += cells[q];
            CHECK(positions == ch.w && positions <= cap, "level order positions");
            check_level_order(cells);
            for (int l = 0; l < kWinCells; ++l) {
                const int64_t gi = win_cell_index(l, t, &d);
                CHECK(gi >= -1 && gi < (int64_t)d.G, "window cell %d -> %lld", l, (long long)gi);
            }
        }
    }

    // Per-tile reduce plans: each (neighbour, chunk) slot is a partial window
    // of that neighbour's tile that contains the cell at the index used.
    void check_reduce() const {
        for (uint32_t b = 0; b < stiles; ++b) {
            const uint4 ti = info[b];
            for (int lane = 0; lane < kTileCells; ++lane) {
                const int l[3] = {lane >> 4, (lane >> 2) & 3, lane & 3};
                const int g[3] = {(int)ti.x * kTileEdge + l[0], (int)ti.y * kTileEdge + l[1],
                                  (int)ti.z * kTileEdge + l[2]};
                const bool cell_ok = in_dims(g[0], g[1], g[2], &d);
                for (int s = 0; s < 8; ++s) {
                    int q;
                    bool valid;
                    const int widx = cover_slot(l, s, &q, &valid);
                    CHECK(q >= 0 && q < 27 && widx >= 0 && widx < kWinCells, "slot %d", s);
                    if (!valid || !cell_ok) continue;
                    const uint32_t kt = nbr[(size_t)b * 27 + q];
                    if (kt == kNoTile) continue;
                    const uint2 tc = tile_chunk[kt];
                    for (uint32_t j = 0; j < tc.y; ++j) {
                        const uint32_t c = tc.x + j;
                        CHECK(c < ntchunks, "partial %u of %u", c, ntchunks);
                        const uint4 ch = tchunks[c];
                        CHECK(ch.y == xmaj[kt - d.koff], "partial %u is of another tile", c);
                        int tt[3];
                        tile_coords(ch.y, &d, tt);
                        const int w[3] = {g[0] - (tt[0] * kTileEdge - 1),
                                          g[1] - (tt[1] * kTileEdge - 1),
                                          g[2] - (tt[2] * kTileEdge - 1)};
                        for (int a = 0; a < 3; ++a)
                            CHECK(w[a] >= 0 && w[a] < kWinEdge, "cell outside window");
                        CHECK((w[0] * kWinEdge + w[1]) * kWinEdge + w[2] == widx,
                              "window index %d", widx);
                    }
                }
            }
        }
    }

    // g2p: the next binning's keys, local ranks, migration counts, far ranks;
    // the state moves into bin order.
    void g2p(const std::vector<oracle_particle> &next) {
        std::fill(count.begin(), count.end(), 0u);  // cleared by reduce_2 before g2p
        std::vector<uint32_t> nkey(n), nrank(n), nsrc(n), nstor(n);
        for (uint32_t c = 0; c < ntchunks; ++c) {
            const uint4 ch = tchunks[c];
            int t[3] = {-4, -4, -4};
            const bool wild_chunk = ch.y == kWildTile;
            if (!wild_chunk) tile_coords(ch.y, &d, t);
            uint32_t lcnt[kDirWild + 2] = {0};
            std::vector<uint32_t> wild_slots;
            for (uint32_t j = 0; j < ch.w; ++j) {
                const uint32_t i = ch.z + j, from = perm[i];
                nstor[i] = stor[from];
                uint32_t dir;
                const uint32_t k = bin_key_of(rank_of.data(), &d, next[stor[from]].position, t,
                                              &dir);
                CHECK(k < nkeys, "key %u", k);
                if (wild_chunk) dir = kDirFar;
                CHECK(dir <= kDirFar, "direction %u", dir);
                nkey[i] = k;
                if (dir <= kDirWild) {
                    nrank[i] = lcnt[dir]++;
                    if (dir == kDirWild) wild_slots.push_back(i);
                } else {
                    nrank[i] = count[k]++;
                }
                nsrc[i] = dir < kDirs ? c * kDirs + dir : kFinal;
                CHECK(nsrc[i] == kFinal || nsrc[i] < mig.size(), "source entry %u", nsrc[i]);
            }
            if (!wild_chunk) {
                for (uint32_t q = 0; q < kDirs; ++q) {
                    CHECK((size_t)c * kDirs + q < mig.size(), "mig of chunk %u", c);
                    mig[(size_t)c * kDirs + q] = lcnt[q];
                }
                const uint32_t base = count[d.koff + d.ntiles];
                count[d.koff + d.ntiles] += lcnt[kDirWild];
                for (uint32_t i : wild_slots) nrank[i] += base;
            }
        }
        key.swap(nkey);
        rnk.swap(nrank);
        src.swap(nsrc);
        stor.swap(nstor);
    }

    void substep(const std::vector<oracle_particle> &now,
                 const std::vector<oracle_particle> &next) {
        if (!keys_ready) count_from_scratch(now);
        gather(!keys_ready);
        scan();
        scatter();
        check_walks();
        check_chunks(now);
        check_reduce();
        g2p(next);
        near_prev.swap(near_now);
        keys_ready = true;
    }
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s particles.bin n bx by bz substeps ppt threads [s:x ...]\n",
                     argv[0]);
        return 1;
    }
    const uint32_t n = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    float box[3] = {std::strtof(argv[3], nullptr), std::strtof(argv[4], nullptr),
                    std::strtof(argv[5], nullptr)};
    const int substeps = std::atoi(argv[6]), ppt = std::atoi(argv[7]), threads = std::atoi(argv[8]);
    std::vector<std::pair<int, float>> resize;
    for (int a = 9; a < argc; ++a) {
        const std::string s(argv[a]);
        const size_t colon = s.find(':');
        resize.emplace_back(std::atoi(s.substr(0, colon).c_str()),
                            std::strtof(s.substr(colon + 1).c_str(), nullptr));
    }
    std::vector<oracle_particle> p(n);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(p.data(), sizeof(oracle_particle), n, f) != n) {
        std::fprintf(stderr, "cannot read %u particles from %s\n", n, argv[1]);
        return 1;
    }
    std::fclose(f);
    for (uint32_t cp : {256u, 512u, 1024u}) check_level_order_patterns(cp);
    Emu e;
    e.init(n, box, ppt);
    const oracle_params prm = {3.0f, 4.0f, 0.1f, 0.2f, 1e7f};
    std::vector<oracle_cell> cells(e.d.G);
    std::vector<oracle_posvel> posvel(n);
    float real[3] = {box[0], box[1], box[2]};
    uint32_t max_chunks = 0, wild_substeps = 0, far_moves = 0;
    for (int s = 0; s < substeps; ++s) {
        for (auto &r : resize)
            if (r.first == s) real[0] = r.second;
        std::vector<oracle_particle> next = p;
        oracle_substep_omp(next.data(), n, cells.data(), e.d.G, posvel.data(), real, box, &prm,
                           threads);
        e.substep(p, next);
        max_chunks = std::max(max_chunks, e.ntchunks);
        for (uint32_t i = 0; i < n; ++i) far_moves += e.src[i] == kFinal;
        wild_substeps += e.count[e.d.koff + e.d.ntiles] != 0;
        p.swap(next);
        if ((s + 1) % 50 == 0) {
            std::fprintf(stderr, "substep %d checks=%ld\n", s + 1, g_checks);
            std::fflush(stderr);
        }
    }
    std::printf("ok substeps=%d particles=%u tiles=%u max_chunks=%u cap_chunks=%llu "
                "final_ranks=%u wild_substeps=%u checks=%ld\n",
                substeps, n, e.stiles, max_chunks, (unsigned long long)e.cap_tchunks, far_moves,
                wild_substeps, g_checks);
    return 0;
}
