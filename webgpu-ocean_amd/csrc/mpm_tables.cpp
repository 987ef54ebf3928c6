// mpm_tables.cpp -- host-built per-scene tables of the production substep:
// bin-key geometry, Hilbert order of the tiles and each tile's neighbours.
// Host code only (also linked by the host-side index emulation).
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "mpm_fused.h"

namespace mpm {

// 3D Hilbert index of (x, y, z) in a 2^bits cube: J. Skilling, "Programming
// the Hilbert curve", AIP Conf. Proc. 707, 381 (2004) -- axes to transposed
// form, then the transposed bits interleaved most significant first.
static uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    if (bits <= 0) return 0;
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t h = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

void append_hilbert_ranks(const SceneDesc &d, std::vector<uint32_t> &ranks,
                          std::vector<uint32_t> &xmaj) {
    const uint32_t n0 = d.knt[0], n1 = d.knt[1], n2 = d.knt[2];
    const uint32_t nt = n0 * n1 * n2;
    int bits = 0;
    while ((1u << bits) < std::max(n0, std::max(n1, n2))) ++bits;
    std::vector<std::pair<uint64_t, uint32_t>> order(nt);
    for (uint32_t x = 0; x < n0; ++x)
        for (uint32_t y = 0; y < n1; ++y)
            for (uint32_t z = 0; z < n2; ++z) {
                const uint32_t t = (x * n1 + y) * n2 + z;
                order[t] = {hilbert3(x, y, z, bits), t};
            }
    std::sort(order.begin(), order.end());
    const size_t base = ranks.size();
    ranks.resize(base + nt);
    for (uint32_t r = 0; r < nt; ++r) {
        ranks[base + order[r].second] = r;
        xmaj.push_back(order[r].second);
    }
    xmaj.push_back(kWildTile);
}

void append_tile_tables(const SceneDesc &d, uint32_t s, const uint32_t *ranks,
                        const uint32_t *xmaj, std::vector<uint32_t> &nbr,
                        std::vector<uint4> &info) {
    const uint32_t n1 = d.knt[1], n2 = d.knt[2];
    for (uint32_t r = 0; r < d.ntiles; ++r) {
        const uint32_t xm = xmaj[r];
        const uint32_t t[3] = {xm / (n1 * n2), (xm / n2) % n1, xm % n2};
        info.push_back(make_uint4(t[0], t[1], t[2], s));
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int64_t nx = (int64_t)t[0] + dx, ny = (int64_t)t[1] + dy,
                                  nz = (int64_t)t[2] + dz;
                    if (nx < 0 || ny < 0 || nz < 0 || nx >= d.knt[0] || ny >= n1 || nz >= n2) {
                        nbr.push_back(kNoTile);
                        continue;
                    }
                    nbr.push_back(d.toff + ranks[((uint32_t)nx * n1 + (uint32_t)ny) * n2 +
                                                 (uint32_t)nz]);
                }
    }
}

void scene_key_geometry(SceneDesc &d, int tile_bits) {
    const int edge = 1 << tile_bits;
    uint64_t nt = 1;
    bool integral = true;
    for (int a = 0; a < 3; ++a) {
        const float b = d.uni.init_box[a];
        const int32_t dim = (int32_t)std::ceil(b);
        d.kdim[a] = dim < 1 ? 1 : dim;
        d.knt[a] = (uint32_t)((d.kdim[a] + edge - 1) >> tile_bits);
        nt *= d.knt[a];
        integral = integral && std::floor(b) == b;
    }
    d.ntiles = (uint32_t)nt;
    d.nkeys = (uint32_t)(nt + 1);  // one key per tile plus the wild key
    d.fast = integral ? 1u : 0u;
    d.tile_bits = (uint32_t)tile_bits;
}

}  // namespace mpm
