// sph.hip -- gfx950 kernels and C ABI (include/sph.h) of the SPH step of
// WebGPU-Ocean's SPHSimulator (sph/sph.ts:301-336).
//
// One substep is seven kernels, graph-replayed:
//   count        cell id of every particle + its rank among the cell's
//                particles (atomic; gridBuild.wgsl:51-60)
//   scan_reduce  per-tile sums of the cell counts  } exclusive prefix sum
//   scan         tile prefix + in-tile scan        } (sph.ts:310-313); also
//                clears the counts for the next substep (gridClear.wgsl)
//   scatter      slot of every particle in its cell's range
//                (reorderParticles.wgsl:52-62)
//   sort_gather  puts each cell's particles in decreasing index order -- the
//                order an in-order run of the reference's atomics gives --
//                and gathers the sorted position/velocity copies
//   density      density.wgsl over the sorted slots; writes the particle's
//                densities AND the sorted copy the reference's second
//                reorderParticles pass makes (sph.ts:321-323)
//   force        force.wgsl + integrate.wgsl + copyPosition.wgsl fused: a
//                particle's integration needs only its own force, and the
//                neighbour data come from the sorted copies
//
// Device state is structure-of-arrays (float4 position, velocity, force; a
// float2 of densities), sorted copies likewise; the 64-B records of the ABI
// are assembled on the host.  Every float operation is one IEEE binary32
// operation in the reference's order (-ffp-contract=off and the pragma of
// mpm_device.h); sums over neighbours run in the sorted order, so results are
// deterministic.
#include <hip/hip_ext.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mpm_device.h"
#include "sph.h"

#pragma clang fp contract(off)

namespace sph {

// Environment + SPHParams uniforms (sph/sph.ts:129-174) and the real box.
struct Uniforms {
    int32_t X, Y, Z;       // xGrids, yGrids, zGrids
    float cell, hx, hy, hz, offset;
    float mass, h, h2, h5, h6, h9, dt, stiff, near_stiff, rest, visc;
    float real_half[3];
    uint32_t G;            // gridCount
};

constexpr unsigned kBlock = 256;
constexpr unsigned kScanPer = 16;                     // counts per thread of the scan
constexpr unsigned kScanTile = kBlock * kScanPer;     // 4096 counts per scan block
constexpr double kMaxCells = 268435456.0;             // 2^28

struct Bufs {
    float4 *pos, *vel, *frc;  // n
    float2 *dn;               // n: density, nearDensity
    float4 *posvel;           // 2n: position, v (32-B PosVel records)
    int32_t *cell;            // n
    uint32_t *rank;           // n
    uint32_t *counts;         // gpad
    uint32_t *start;          // gpad: exclusive prefix, start[G] = particles binned
    uint32_t *tsum;           // ntiles
    uint32_t *order, *order2; // n: particle of each sorted slot (scattered, sorted)
    float4 *spos, *svel;      // n: sorted copies
    float2 *sdn;              // n
    const Uniforms *u;
    uint32_t n;
};

// cellPosition -- sph/density.wgsl:53-58
__device__ __forceinline__ void cell_position(const Uniforms &u, float4 p, int32_t c[3]) {
    c[0] = mpm::f2i_sat(floorf((p.x + u.hx + u.offset) / u.cell));
    c[1] = mpm::f2i_sat(floorf((p.y + u.hy + u.offset) / u.cell));
    c[2] = mpm::f2i_sat(floorf((p.z + u.hz + u.offset) / u.cell));
}

// cellId -- sph/grid/gridBuild.wgsl:41-48 (i32 wrap-around arithmetic)
__device__ __forceinline__ int32_t cell_id(const Uniforms &u, float4 p) {
    int32_t c[3];
    cell_position(u, p, c);
    return (int32_t)((uint32_t)c[0] + (uint32_t)c[1] * (uint32_t)u.X +
                     (uint32_t)c[2] * (uint32_t)u.X * (uint32_t)u.Y);
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

// gridBuild: ranks by atomicAdd.  Particles whose linear cell id is outside
// [0, G) are not binned; their substep is final here (density 0, force 0,
// integrate skipped, posvel = current state).
__global__ void __launch_bounds__(kBlock) k_count(Bufs b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= b.n) return;
    const Uniforms &u = *b.u;
    const float4 p = b.pos[i];
    const int32_t c = cell_id(u, p);
    b.cell[i] = c;
    if (c >= 0 && (uint32_t)c < u.G) {
        b.rank[i] = atomicAdd(&b.counts[c], 1u);
    } else {
        b.dn[i] = make_float2(0.0f, 0.0f);
        b.frc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.posvel[2 * (size_t)i] = p;
        b.posvel[2 * (size_t)i + 1] = b.vel[i];
    }
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (unsigned w = 0; w < kBlock / 64; ++w) s += lds[w];
    __syncthreads();
    return s;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(Bufs b) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint4 *src = reinterpret_cast<const uint4 *>(b.counts + (size_t)blockIdx.x * kScanTile) +
                       threadIdx.x * (kScanPer / 4);
    uint32_t s = 0;
#pragma unroll
    for (unsigned q = 0; q < kScanPer / 4; ++q) {
        const uint4 v = src[q];
        s += v.x + v.y + v.z + v.w;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) b.tsum[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_scan(Bufs b) {
    __shared__ uint32_t lds[kBlock / 64];
    uint32_t base = 0;
    for (unsigned t = threadIdx.x; t < blockIdx.x; t += kBlock) base += b.tsum[t];
    base = block_sum(base, lds);
    uint4 *cnt = reinterpret_cast<uint4 *>(b.counts + (size_t)blockIdx.x * kScanTile) +
                 threadIdx.x * (kScanPer / 4);
    uint4 *dst = reinterpret_cast<uint4 *>(b.start + (size_t)blockIdx.x * kScanTile) +
                 threadIdx.x * (kScanPer / 4);
    uint4 v[kScanPer / 4];
    uint32_t s = 0;
#pragma unroll
    for (unsigned q = 0; q < kScanPer / 4; ++q) {
        v[q] = cnt[q];
        s += v[q].x + v[q].y + v[q].z + v[q].w;
    }
    const uint32_t incl = wave_incl_scan(s);
    if ((threadIdx.x & 63) == 63) lds[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t run = base + incl - s;
    for (unsigned w = 0; w < (threadIdx.x >> 6); ++w) run += lds[w];
    const uint4 zero = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (unsigned q = 0; q < kScanPer / 4; ++q) {
        uint4 o;
        o.x = run; run += v[q].x;
        o.y = run; run += v[q].y;
        o.z = run; run += v[q].z;
        o.w = run; run += v[q].w;
        dst[q] = o;
        cnt[q] = zero;  // gridClear for the next substep
    }
}

// reorderParticles: targetIndex = end of the cell - rank - 1
// (sph/grid/reorderParticles.wgsl:57).
__global__ void __launch_bounds__(kBlock) k_scatter(Bufs b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= b.n) return;
    const int32_t c = b.cell[i];
    if (c < 0 || (uint32_t)c >= b.u->G) return;
    // the slot is below start[G] <= n whenever the counts were cleared before
    // count ran (stages in order); the test mirrors `targetIndex < params.n`
    const uint32_t t = b.start[c + 1] - b.rank[i] - 1;
    if (t < b.n) b.order[t] = i;
}

// Slot d of particle i = its cell's start + the number of the cell's
// particles with a larger index (decreasing index order).
__global__ void __launch_bounds__(kBlock) k_sort_gather(Bufs b) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= b.n || k >= b.start[b.u->G]) return;
    const uint32_t i = b.order[k];
    const int32_t c = b.cell[i];
    const uint32_t s = b.start[c], e = min(b.start[c + 1], b.n);
    uint32_t d = s;
    for (uint32_t j = s; j < e; ++j) d += b.order[j] > i ? 1u : 0u;
    if (d >= b.n) return;
    b.order2[d] = i;
    b.spos[d] = b.pos[i];
    b.svel[d] = b.vel[i];
}

// Neighbour rows of density.wgsl:72-85 / force.wgsl:82-96: for dz, dy in
// [-1, 1] inside the grid, the sorted range [lo, hi) spanning the x row.
struct Rows {
    int32_t v[3];
    int dz0, dz1, dy0, dy1, dx0, dx1;
    bool inside;
};

__device__ __forceinline__ Rows rows_of(const Uniforms &u, float4 p) {
    Rows r;
    cell_position(u, p, r.v);
    r.inside = r.v[0] < u.X && 0 <= r.v[0] && r.v[1] < u.Y && 0 <= r.v[1] && r.v[2] < u.Z &&
               0 <= r.v[2];
    r.dz0 = max(-1, -r.v[2]);
    r.dz1 = min(1, u.Z - r.v[2] - 1);
    r.dy0 = max(-1, -r.v[1]);
    r.dy1 = min(1, u.Y - r.v[1] - 1);
    r.dx0 = max(-1, -r.v[0]);
    r.dx1 = min(1, u.X - r.v[0] - 1);
    return r;
}

// The nine neighbour rows of a particle as sorted-slot ranges [lo, hi) in
// the reference's order (dz outer, dy inner); rows outside the grid, and all
// rows of a particle whose cell is outside the grid, are empty.
struct RowRanges {
    uint32_t lo[9], hi[9];
};

__device__ __forceinline__ RowRanges row_ranges(const Bufs &b, const Uniforms &u, const Rows &r) {
    RowRanges rr;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int dz = t / 3 - 1, dy = t % 3 - 1;
        const bool ok = r.inside && dz >= r.dz0 && dz <= r.dz1 && dy >= r.dy0 && dy <= r.dy1;
        uint32_t lo = 0, hi = 0;
        if (ok) {
            const int32_t row = (r.v[1] + dy) * u.X + (r.v[2] + dz) * u.X * u.Y;
            lo = b.start[row + r.v[0] + r.dx0];
            hi = min(b.start[row + r.v[0] + r.dx1 + 1], b.n);
        }
        rr.lo[t] = lo;
        rr.hi[t] = hi;
    }
    return rr;
}

// Neighbour walks: four lanes (a quad) per particle.  The candidates of the
// nine rows form one stream; lane q of the quad evaluates stream entries q,
// q + 4, q + 8, ... (the distance test, the kernels and the IEEE divisions,
// i.e. nearly all of the work), and every lane then adds the quad's four
// terms to its accumulators in stream order, reading them through DPP quad
// broadcasts.  The sequence of float additions is exactly the sequential
// walk's: a candidate that does not interact adds +0.0, which leaves any
// accumulator unchanged (an accumulator starting at +0 never becomes -0
// under round-to-nearest, and x + 0 == x for every other x).  Four times the
// lanes of a particle-per-lane walk, and a quarter of its dependent steps.
constexpr unsigned kNbrBlock = 64;
constexpr unsigned kQuad = 4;

struct Stream {
    uint32_t lo[9], pre[10];  // row starts; pre[t] = candidates in rows < t
};

__device__ __forceinline__ Stream stream_of(const RowRanges &rr) {
    Stream s;
    s.pre[0] = 0;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        s.lo[t] = rr.lo[t];
        s.pre[t + 1] = s.pre[t] + (rr.hi[t] > rr.lo[t] ? rr.hi[t] - rr.lo[t] : 0u);
    }
    return s;
}

// Sorted slot of stream entry x < pre[9]: in the last row t with pre[t] <= x
// (that row holds x: the rows after it start beyond x).
__device__ __forceinline__ uint32_t slot_at(const Stream &s, uint32_t x) {
    uint32_t j = 0;
#pragma unroll
    for (int t = 0; t < 9; ++t)
        if (x >= s.pre[t]) j = s.lo[t] + (x - s.pre[t]);
    return j;
}

// Lane Q's value of the calling lane's quad (DPP quad_perm [Q,Q,Q,Q]).
template <int Q>
__device__ __forceinline__ float quad_lane(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), Q * 0x55, 0xF, 0xF, false));
}

// acc = (((acc + t[0]) + t[1]) + t[2]) + t[3], t[q] = lane q's term.
__device__ __forceinline__ float quad_accumulate(float acc, float t) {
    acc = acc + quad_lane<0>(t);
    acc = acc + quad_lane<1>(t);
    acc = acc + quad_lane<2>(t);
    acc = acc + quad_lane<3>(t);
    return acc;
}

// Value of lane l ^ 4 within the same 8-lane group (DPP row_shl:4 for the
// lower quad, row_shr:4 for the upper; groups of 8 lie inside one 16-lane
// DPP row).
__device__ __forceinline__ float other_quad(float v, bool upper) {
    const int i = __float_as_int(v);
    const int from_above = __builtin_amdgcn_update_dpp(0, i, 0x104, 0xF, 0xF, false);  // row_shl:4
    const int from_below = __builtin_amdgcn_update_dpp(0, i, 0x114, 0xF, 0xF, false);  // row_shr:4
    return __int_as_float(upper ? from_below : from_above);
}

// acc += the 8 terms of the lane's 8-lane group in order (lower quad's lanes
// 0..3, then the upper quad's), bit for bit the sequential sum.
__device__ __forceinline__ float oct_accumulate(float acc, float t, bool upper) {
    const float o = other_quad(t, upper);
    const float lo = upper ? o : t, hi = upper ? t : o;
    acc = quad_accumulate(acc, lo);
    return quad_accumulate(acc, hi);
}

constexpr unsigned kOct = 8;

// Below this many particles a quad per particle leaves fewer than 4 waves per
// SIMD of the chip (256 CUs x 4 SIMDs), and the density walk takes 8 lanes
// per particle instead (40k: 24.1 -> 20.6 us; 200k, where quads already fill
// the chip: 88 -> 99.5 us with 8 lanes).
constexpr uint32_t kOctBelow = 65536;

// computeDensity -- sph/density.wgsl:41-51 (kernels), :65-106, over the sorted
// slots.  L = 4 or 8 lanes per particle: lane l evaluates stream entries
// L m + l, and every lane adds the group's L terms in stream order (8 lanes:
// the lower quad's four, then the upper quad's).  (Listing the interacting
// candidates first, as the force walk does, was slower here: 88 -> 109 us at
// 200k particles; the density terms are too cheap to pay for a second pass.)
template <unsigned L>
__global__ void __launch_bounds__(kNbrBlock) k_density(Bufs b) {
    static_assert(L == kQuad || L == kOct, "4 or 8 lanes per particle");
    const uint32_t g = blockIdx.x * kNbrBlock + threadIdx.x;
    const uint32_t k = g / L, q = g % L;
    const bool upper = q >= kQuad;
    const Uniforms &u = *b.u;
    if (k >= b.n || k >= b.start[u.G]) return;  // the whole group leaves together
    const float4 p = b.spos[k];
    const float scale_d = 315.0f / ((float)(64. * 3.1415926535) * u.h9);
    const float scale_n = 15.0f / ((float)3.1415926535 * u.h6);
    float density = 0.0f, near_density = 0.0f;
    const Stream st = stream_of(row_ranges(b, u, rows_of(u, p)));
    const uint32_t total = st.pre[9];
    for (uint32_t base = 0; base < total; base += L) {
        const uint32_t x = base + q;
        const bool valid = x < total;
        const float4 c = b.spos[valid ? slot_at(st, x) : k];
        const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
        const float r2 = dot3(dx, dy, dz, dx, dy, dz);
        float td = 0.0f, tn = 0.0f;
        if (valid && r2 < u.h2) {
            const float r = sqrtf(r2);
            const float dd = u.h2 - r * r;
            td = u.mass * (((scale_d * dd) * dd) * dd);
            const float d = u.h - r;
            tn = u.mass * (((scale_n * d) * d) * d);
        }
        if constexpr (L == kOct) {
            density = oct_accumulate(density, td, upper);
            near_density = oct_accumulate(near_density, tn, upper);
        } else {
            density = quad_accumulate(density, td);
            near_density = quad_accumulate(near_density, tn);
        }
    }
    if (q != 0) return;
    const uint32_t i = b.order2[k];
    const float2 out = make_float2(density, near_density);
    b.dn[i] = out;
    b.sdn[k] = out;
}

// computeForce (sph/force.wgsl:41-59 kernels, :73-152) + integrate
// (sph/integrate.wgsl:35-65) + copyPosition (sph/copyPosition.wgsl:34-38),
// over the sorted slots.  |pos_j - pos_i| is the r of the pair: the two
// differences are negations of each other, exactly, so dot(dir, dir) == r2
// bit for bit and normalize's length is sqrt(r2).
//
// Two passes per particle's lane group: the cheap test (densities non-zero,
// 0 < r2 < h^2) over the candidate stream appends the interacting slots, in
// stream order, to the group's list in LDS (group ballot + prefix); the pair
// terms (two square roots, twelve IEEE divisions) are then evaluated over the
// list, one pair per lane, and added in list order through DPP broadcasts.
// Only about one candidate in six interacts, so evaluating the terms over
// the whole stream left most lanes of a wave idle in the expensive part.
// The dropped candidates would have added +0.0: the sums are unchanged, bit
// for bit.  When any list of a wave could overflow (kListCap entries), all of
// the wave's groups evaluate their lists before the test continues: a drain
// taken by one group at a time would run the pair terms with most of the
// wave masked off.
constexpr uint32_t kListCap = 128;

// L = 4 or 8 lanes per particle: the group's list is filled L candidates per
// step and evaluated L pairs per step; with 8 lanes the lower quad's pairs
// come before the upper quad's.  The walk runs with 8 lanes at every size
// (dam break, 40k particles: 37.2 -> 28.3 us; 200k: 183 -> 160 us against
// quads).  16 lanes (one DPP row, row_newbcast) measured slower: 30.9 and
// 177 us; every lane adds all the group's terms, and that share grows with
// the group.
template <unsigned L>
__device__ __forceinline__ void lanes_of_pair(float t, bool upper, float &lo, float &hi) {
    if constexpr (L == kOct) {
        const float o = other_quad(t, upper);
        lo = upper ? o : t;
        hi = upper ? t : o;
    } else {
        lo = t;
        hi = 0.0f;
    }
}

// acc += a[0], n[0], a[1], n[1], ... over the group's pairs in order (the
// pressure term, then the near-pressure term, of each pair).
template <unsigned L>
__device__ __forceinline__ float add_pairs(float acc, float a, float n, bool upper) {
    float alo, ahi, nlo, nhi;
    lanes_of_pair<L>(a, upper, alo, ahi);
    lanes_of_pair<L>(n, upper, nlo, nhi);
    acc = acc + quad_lane<0>(alo); acc = acc + quad_lane<0>(nlo);
    acc = acc + quad_lane<1>(alo); acc = acc + quad_lane<1>(nlo);
    acc = acc + quad_lane<2>(alo); acc = acc + quad_lane<2>(nlo);
    acc = acc + quad_lane<3>(alo); acc = acc + quad_lane<3>(nlo);
    if constexpr (L == kOct) {
        acc = acc + quad_lane<0>(ahi); acc = acc + quad_lane<0>(nhi);
        acc = acc + quad_lane<1>(ahi); acc = acc + quad_lane<1>(nhi);
        acc = acc + quad_lane<2>(ahi); acc = acc + quad_lane<2>(nhi);
        acc = acc + quad_lane<3>(ahi); acc = acc + quad_lane<3>(nhi);
    }
    return acc;
}

template <unsigned L>
__device__ __forceinline__ float add_terms(float acc, float t, bool upper) {
    if constexpr (L == kOct) return oct_accumulate(acc, t, upper);
    else return quad_accumulate(acc, t);
}

template <unsigned L>
__global__ void __launch_bounds__(kNbrBlock) k_force(Bufs b) {
    static_assert(L == kQuad || L == kOct, "4 or 8 lanes per particle");
    __shared__ uint32_t s_list[kNbrBlock / L][kListCap];
    const uint32_t g = blockIdx.x * kNbrBlock + threadIdx.x;
    const uint32_t k = g / L, q = g % L;
    const bool upper = q >= kQuad;
    const Uniforms &u = *b.u;
    if (k >= b.n || k >= b.start[u.G]) return;  // the whole group leaves together
    uint32_t *list = s_list[threadIdx.x / L];
    const uint32_t group_shift = (threadIdx.x & 63u) & ~(L - 1u);
    const float4 p = b.spos[k];
    float4 v = b.svel[k];
    const float2 dni = b.sdn[k];
    const float density_i = dni.x, near_density_i = dni.y;
    const float scale_g = 45.0f / ((float)3.1415926535 * u.h6);
    const float scale_ng = 45.0f / ((float)3.1415926535 * u.h5);
    const float scale_l = 45.0f / ((float)3.1415926535 * u.h6);
    const float pressure_i = u.stiff * (density_i - u.rest);
    const float near_pressure_i = u.near_stiff * near_density_i;
    float fpx = 0.0f, fpy = 0.0f, fpz = 0.0f, fvx = 0.0f, fvy = 0.0f, fvz = 0.0f;
    uint32_t cnt = 0;  // entries in the group's list (group-uniform)
    // Evaluates the listed pairs in order and empties the list.
    auto drain = [&]() {
        // the list entries were written by the other lanes of this group (same
        // wave, earlier in program order); keep the compiler from moving the
        // reads above those writes
        __builtin_amdgcn_wave_barrier();
        for (uint32_t base = 0; base < cnt; base += L) {
            const uint32_t x = base + q;
            const bool valid = x < cnt;
            const uint32_t j = valid ? list[x] : k;
            const float4 c = b.spos[j];
            const float2 dnj = b.sdn[j];
            const float4 w = b.svel[j];
            float ax = 0.0f, ay = 0.0f, az = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
            float vx = 0.0f, vy = 0.0f, vz = 0.0f;
            if (valid) {
                const float ex = p.x - c.x, ey = p.y - c.y, ez = p.z - c.z;
                const float r = sqrtf(dot3(ex, ey, ez, ex, ey, ez));
                const float pressure_j = u.stiff * (dnj.x - u.rest);
                const float near_pressure_j = u.near_stiff * dnj.y;
                const float shared = (pressure_i + pressure_j) / 2.0f;
                const float near_shared = (near_pressure_i + near_pressure_j) / 2.0f;
                const float dirx = (c.x - p.x) / r, diry = (c.y - p.y) / r,
                            dirz = (c.z - p.z) / r;
                const float dg = u.h - r;
                const float grad = (scale_g * dg) * dg;
                const float near_grad = (scale_ng * dg) * dg;
                const float lap = scale_l * dg;
                const float sp = -u.mass * shared, nsp = -u.mass * near_shared;
                ax = ((sp * dirx) * grad) / dnj.x;
                ay = ((sp * diry) * grad) / dnj.x;
                az = ((sp * dirz) * grad) / dnj.x;
                nx = ((nsp * dirx) * near_grad) / dnj.y;
                ny = ((nsp * diry) * near_grad) / dnj.y;
                nz = ((nsp * dirz) * near_grad) / dnj.y;
                vx = ((u.mass * (w.x - v.x)) * lap) / dnj.x;
                vy = ((u.mass * (w.y - v.y)) * lap) / dnj.x;
                vz = ((u.mass * (w.z - v.z)) * lap) / dnj.x;
            }
            // per pair, in order: fPress += pressure term, then += near-pressure
            // term (force.wgsl:113-114); fVisc += (:116)
            fpx = add_pairs<L>(fpx, ax, nx, upper);
            fpy = add_pairs<L>(fpy, ay, ny, upper);
            fpz = add_pairs<L>(fpz, az, nz, upper);
            fvx = add_terms<L>(fvx, vx, upper);
            fvy = add_terms<L>(fvy, vy, upper);
            fvz = add_terms<L>(fvz, vz, upper);
        }
        cnt = 0;
    };
    const Stream st = stream_of(row_ranges(b, u, rows_of(u, p)));
    const uint32_t total = st.pre[9];
    for (uint32_t base = 0; base < total; base += L) {
        const uint32_t x = base + q;
        const bool valid = x < total;
        const uint32_t j = valid ? slot_at(st, x) : k;
        const float4 c = b.spos[j];
        const float2 dnj = b.sdn[j];
        const float ex = p.x - c.x, ey = p.y - c.y, ez = p.z - c.z;
        const float r2 = dot3(ex, ey, ez, ex, ey, ez);
        const bool hit = valid && dnj.x != 0.0f && dnj.y != 0.0f && r2 < u.h2 && 0.0f < r2;
        const uint32_t qm = (uint32_t)(__ballot(hit) >> group_shift) & ((1u << L) - 1u);
        if (__any(cnt + L > kListCap)) drain();  // wave-uniform
        if (hit) list[cnt + (uint32_t)__popc(qm & ((1u << q) - 1u))] = j;
        cnt += (uint32_t)__popc(qm);
    }
    drain();
    if (q != 0) return;
    const uint32_t i = b.order2[k];
    fvx = fvx * u.visc;
    fvy = fvy * u.visc;
    fvz = fvz * u.visc;
    float4 f;
    f.x = (fpx + fvx) + density_i * 0.0f;
    f.y = (fpy + fvy) + density_i * -9.8f;
    f.z = (fpz + fvz) + density_i * 0.0f;
    f.w = 0.0f;
    b.frc[i] = f;
    float4 x = p;
    if (density_i != 0.0f) {
        float a[3] = {f.x / density_i, f.y / density_i, f.z / density_i};
        const float pos[3] = {p.x, p.y, p.z};
        float wall[3][3];  // wall[axis][component]: plus + minus wall force of that axis
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float plus = u.real_half[ax] - pos[ax];
            const float minus = u.real_half[ax] + pos[ax];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float up = c == ax ? 1.0f : 0.0f, down = c == ax ? -1.0f : 0.0f;
                wall[ax][c] = (up * 8000.0f) * fminf(plus, 0.0f) +
                              (down * 8000.0f) * fminf(minus, 0.0f);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] = a[c] + ((wall[0][c] + wall[1][c]) + wall[2][c]);
        v.x = v.x + u.dt * a[0];
        v.y = v.y + u.dt * a[1];
        v.z = v.z + u.dt * a[2];
        x.x = x.x + u.dt * v.x;
        x.y = x.y + u.dt * v.y;
        x.z = x.z + u.dt * v.z;
    }
    b.pos[i] = x;
    b.vel[i] = v;
    b.posvel[2 * (size_t)i] = make_float4(x.x, x.y, x.z, 0.0f);
    b.posvel[2 * (size_t)i + 1] = make_float4(v.x, v.y, v.z, 0.0f);
}

__global__ void k_set_uniforms(Uniforms *d, Uniforms u) { *d = u; }

}  // namespace sph

// ---- host side ---------------------------------------------------------------

namespace sph {
namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? SPH_ERR_OUT_OF_MEMORY : SPH_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define SPH_TRY(call)                                      \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #call);  \
    } while (0)

// Uniform values of sph/sph.ts:58-77,129-174: JavaScript doubles, rounded to
// f32 where the reference stores them in its Float32Array views.
int derive(const sph_params &p, Uniforms &u) {
    const double vals[7] = {p.kernel_radius, p.mass,         p.dt,       p.stiffness,
                            p.near_stiffness, p.rest_density, p.viscosity};
    for (double v : vals)
        if (!std::isfinite(v)) return fail(SPH_ERR_INVALID_ARGUMENT, "parameters must be finite");
    const double h = p.kernel_radius;
    if (!(h > 0.0)) return fail(SPH_ERR_INVALID_ARGUMENT, "kernel_radius must be > 0");
    const double cell = 1.0 * h, half = 2.0, len = 2.0 * half, sentinel = 4 * cell;
    const double grids = std::ceil((len + sentinel) / cell);
    if (!(grids >= 1.0) || grids * grids * grids > kMaxCells)
        return fail(SPH_ERR_GRID_TOO_LARGE,
                    "the neighbour grid of kernel_radius " + std::to_string(h) +
                        " exceeds 2^28 cells");
    u.X = u.Y = u.Z = (int32_t)grids;
    u.cell = (float)cell;
    u.hx = u.hy = u.hz = (float)half;
    u.offset = (float)(sentinel / 2);
    u.mass = (float)p.mass;
    u.h = (float)h;
    u.h2 = (float)std::pow(h, 2);
    u.h5 = (float)std::pow(h, 5);
    u.h6 = (float)std::pow(h, 6);
    u.h9 = (float)std::pow(h, 9);
    u.dt = (float)p.dt;
    u.stiff = (float)p.stiffness;
    u.near_stiff = (float)p.near_stiffness;
    u.rest = (float)p.rest_density;
    u.visc = (float)p.viscosity;
    u.G = (uint32_t)(grids * grids * grids);
    return SPH_OK;
}

int check_box(const float b[3], const char *what) {
    if (!b) return fail(SPH_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b[a]))
            return fail(SPH_ERR_INVALID_ARGUMENT, std::string(what) + " must be finite");
    return SPH_OK;
}

inline unsigned blocks_for(uint64_t count) { return (unsigned)((count + kBlock - 1) / kBlock); }

// One kernel of the substep; with events, they take the kernel's own start
// and end timestamps (hipExtLaunchKernel).
hipError_t launch_kernel(void (*kernel)(Bufs), unsigned grid, const Bufs &b, hipStream_t st,
                         hipEvent_t e0, hipEvent_t e1, unsigned block = kBlock) {
    if (grid == 0) return hipSuccess;
    if (e0)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, e0, e1, 0, b);
    else
        kernel<<<grid, block, 0, st>>>(b);
    return hipGetLastError();
}

hipError_t launch_stage(int stage, const Bufs &b, unsigned ntiles, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const unsigned pb = blocks_for(b.n);
    const unsigned nb = (unsigned)(((uint64_t)b.n * kQuad + kNbrBlock - 1) / kNbrBlock);
    const unsigned ob = (unsigned)(((uint64_t)b.n * kOct + kNbrBlock - 1) / kNbrBlock);
    switch (stage) {
    case SPH_STAGE_COUNT: return launch_kernel(k_count, pb, b, st, e0, e1);
    case SPH_STAGE_SCAN_REDUCE: return launch_kernel(k_scan_reduce, ntiles, b, st, e0, e1);
    case SPH_STAGE_SCAN: return launch_kernel(k_scan, ntiles, b, st, e0, e1);
    case SPH_STAGE_SCATTER: return launch_kernel(k_scatter, pb, b, st, e0, e1);
    case SPH_STAGE_SORT_GATHER: return launch_kernel(k_sort_gather, pb, b, st, e0, e1);
    case SPH_STAGE_DENSITY:
        return b.n < kOctBelow ? launch_kernel(k_density<kOct>, ob, b, st, e0, e1, kNbrBlock)
                               : launch_kernel(k_density<kQuad>, nb, b, st, e0, e1, kNbrBlock);
    case SPH_STAGE_FORCE: return launch_kernel(k_force<kOct>, ob, b, st, e0, e1, kNbrBlock);
    default: return hipErrorInvalidValue;
    }
}

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace
}  // namespace sph

struct sph_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    sph_params params{};
    sph::Uniforms uni{};
    sph::Bufs b{};
    void *slab = nullptr;
    uint32_t n = 0;
    unsigned ntiles = 0;
    size_t gpad = 0;
    hipGraphExec_t graph = nullptr;
    bool ready = false;
};

namespace sph {
namespace {

void release(sph_sim *s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    if (s->slab) (void)hipFree(s->slab);
    s->slab = nullptr;
    s->b = Bufs{};
    s->ready = false;
}

int select(const sph_sim *s) {
    if (!s->ready)
        return fail(SPH_ERR_NOT_READY,
                    "the handle holds no scene (its last reset failed); call sph_reset again");
    SPH_TRY(hipSetDevice(s->device));
    return SPH_OK;
}

// Host 64-B records -> device structure-of-arrays (positions, velocities,
// forces, densities); posvel starts as the records' positions/velocities.
int upload(sph_sim *s, const void *src, uint32_t n) {
    const float *r = static_cast<const float *>(src);
    std::vector<float4> pos(n), vel(n), frc(n), pv(2 * (size_t)n);
    std::vector<float2> dn(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float *q = r + 16 * (size_t)i;
        pos[i] = make_float4(q[0], q[1], q[2], 0.0f);
        vel[i] = make_float4(q[4], q[5], q[6], 0.0f);
        frc[i] = make_float4(q[8], q[9], q[10], 0.0f);
        dn[i] = make_float2(q[11], q[12]);
        pv[2 * (size_t)i] = pos[i];
        pv[2 * (size_t)i + 1] = vel[i];
    }
    if (n == 0) return SPH_OK;
    SPH_TRY(hipMemcpyAsync(s->b.pos, pos.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    SPH_TRY(hipMemcpyAsync(s->b.vel, vel.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    SPH_TRY(hipMemcpyAsync(s->b.frc, frc.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    SPH_TRY(hipMemcpyAsync(s->b.dn, dn.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    SPH_TRY(hipMemcpyAsync(s->b.posvel, pv.data(), 32 * (size_t)n, hipMemcpyHostToDevice,
                           s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

int set_uniforms(sph_sim *s) {
    k_set_uniforms<<<1, 1, 0, s->stream>>>(const_cast<Uniforms *>(s->b.u), s->uni);
    SPH_TRY(hipGetLastError());
    return SPH_OK;
}

int reset_impl(sph_sim *s, const void *particles, uint32_t n, const float half[3]) {
    release(s);
    SPH_TRY(hipSetDevice(s->device));
    const uint32_t cap = n ? n : 1;
    const size_t gcount = (size_t)s->uni.G + 1;
    s->ntiles = (unsigned)((gcount + kScanTile - 1) / kScanTile);
    s->gpad = (size_t)s->ntiles * kScanTile;
    size_t off = 0;
    std::vector<std::pair<void **, size_t>> parts;
    Bufs &b = s->b;
    auto add = [&](void **p, size_t bytes) { parts.push_back({p, bytes}); };
    add((void **)&b.pos, 16 * (size_t)cap);
    add((void **)&b.vel, 16 * (size_t)cap);
    add((void **)&b.frc, 16 * (size_t)cap);
    add((void **)&b.dn, 8 * (size_t)cap);
    add((void **)&b.posvel, 32 * (size_t)cap);
    add((void **)&b.cell, 4 * (size_t)cap);
    add((void **)&b.rank, 4 * (size_t)cap);
    add((void **)&b.counts, 4 * s->gpad);
    add((void **)&b.start, 4 * s->gpad);
    add((void **)&b.tsum, 4 * (size_t)s->ntiles);
    add((void **)&b.order, 4 * (size_t)cap);
    add((void **)&b.order2, 4 * (size_t)cap);
    add((void **)&b.spos, 16 * (size_t)cap);
    add((void **)&b.svel, 16 * (size_t)cap);
    add((void **)&b.sdn, 8 * (size_t)cap);
    add((void **)&b.u, sizeof(Uniforms));
    for (auto &p : parts) off += align_up(p.second);
    SPH_TRY(hipMalloc(&s->slab, off));
    off = 0;
    for (auto &p : parts) {
        *p.first = static_cast<char *>(s->slab) + off;
        off += align_up(p.second);
    }
    b.n = n;
    s->n = n;
    // counts start cleared; every substep's scan clears what it consumed
    SPH_TRY(hipMemsetAsync(b.counts, 0, 4 * s->gpad, s->stream));
    SPH_TRY(hipMemsetAsync(b.start, 0, 4 * s->gpad, s->stream));
    for (int a = 0; a < 3; ++a) s->uni.real_half[a] = half[a];
    int rc = set_uniforms(s);
    if (rc) return rc;
    rc = upload(s, particles, n);
    if (rc) return rc;
    // one substep as a graph: the seven kernels, replayed by sph_step
    hipGraph_t g = nullptr;
    SPH_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int st = 0; st < SPH_STAGE_NUM && le == hipSuccess; ++st)
        le = launch_stage(st, b, s->ntiles, s->stream);
    hipError_t ee = hipStreamEndCapture(s->stream, &g);
    if (le != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le != hipSuccess ? le : ee, "capture of the substep");
    }
    hipError_t ie = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
    s->ready = true;
    return SPH_OK;
}

int reset_checked(sph_sim *s, const void *particles, uint32_t n, const float half[3]) {
    int rc = check_box(half, "init_half_box");
    if (rc) return rc;
    if (n && !particles) return fail(SPH_ERR_INVALID_ARGUMENT, "particles is NULL");
    rc = reset_impl(s, particles, n, half);
    if (rc) {
        const std::string msg = g_last_error;
        release(s);
        g_last_error = msg;
    }
    return rc;
}

}  // namespace
}  // namespace sph

using sph::fail;

extern "C" {

const char *sph_last_error(void) { return sph::g_last_error.c_str(); }

int sph_default_params(sph_params *out) {
    if (!out) return fail(SPH_ERR_INVALID_ARGUMENT, "out is NULL");
    out->kernel_radius = 0.07;
    out->mass = 1.0;
    out->dt = 0.006;
    out->stiffness = 20;
    out->near_stiffness = 1.0;
    out->rest_density = 15000;
    out->viscosity = 100;
    return SPH_OK;
}

int sph_create(const sph_params *params, const void *particles, uint32_t n,
               const float init_half_box[3], sph_sim **out) {
    if (!out) return fail(SPH_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    sph_params p;
    if (params)
        p = *params;
    else
        sph_default_params(&p);
    sph::Uniforms u{};
    int rc = sph::derive(p, u);
    if (rc) return rc;
    rc = sph::check_box(init_half_box, "init_half_box");
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SPH_ERR_NO_DEVICE, "no HIP device is visible");
    sph_sim *s = new (std::nothrow) sph_sim();
    if (!s) return fail(SPH_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->params = p;
    s->uni = u;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        rc = sph::hip_fail(e, "stream");
        sph_destroy(s);
        return rc;
    }
    rc = sph::reset_checked(s, particles, n, init_half_box);
    if (rc) {
        const std::string msg = sph::g_last_error;
        sph_destroy(s);
        sph::g_last_error = msg;
        return rc;
    }
    *out = s;
    return SPH_OK;
}

int sph_destroy(sph_sim *s) {
    if (!s) return SPH_OK;
    (void)hipSetDevice(s->device);
    sph::release(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return SPH_OK;
}

int sph_reset(sph_sim *s, const void *particles, uint32_t n, const float init_half_box[3]) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    return sph::reset_checked(s, particles, n, init_half_box);
}

int sph_set_box(sph_sim *s, const float real_half_box[3]) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = sph::check_box(real_half_box, "real_half_box");
    if (rc) return rc;
    rc = sph::select(s);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) s->uni.real_half[a] = real_half_box[a];
    return sph::set_uniforms(s);
}

int sph_step(sph_sim *s, uint32_t substeps) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = sph::select(s);
    if (rc) return rc;
    if (s->n == 0) return SPH_OK;
    for (uint32_t k = 0; k < substeps; ++k) {
        hipError_t e = hipGraphLaunch(s->graph, s->stream);
        if (e != hipSuccess) return sph::hip_fail(e, "hipGraphLaunch");
    }
    return SPH_OK;
}

int sph_run_stage(sph_sim *s, int stage) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    if (stage < 0 || stage >= SPH_STAGE_NUM)
        return fail(SPH_ERR_INVALID_ARGUMENT, "stage out of range");
    int rc = sph::select(s);
    if (rc) return rc;
    if (s->n == 0) return SPH_OK;
    hipError_t e = sph::launch_stage(stage, s->b, s->ntiles, s->stream);
    if (e != hipSuccess) return sph::hip_fail(e, "launch");
    return SPH_OK;
}

int sph_synchronize(sph_sim *s) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = sph::select(s);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sph::hip_fail(e, "hipStreamSynchronize");
    return SPH_OK;
}

static int copy_out(sph_sim *s, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return SPH_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sph::hip_fail(e, "device to host copy");
    return SPH_OK;
}

int sph_read_posvel(sph_sim *s, void *dst, uint32_t n) {
    if (!s || (n && !dst)) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    if (n > s->n) return fail(SPH_ERR_INVALID_ARGUMENT, "n exceeds the particle count");
    return copy_out(s, dst, s->b.posvel, 32 * (size_t)n);
}

int sph_posvel_device_ptr(sph_sim *s, void **out) {
    if (!s || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    *out = s->b.posvel;
    return SPH_OK;
}

int sph_read_particles(sph_sim *s, void *dst, uint32_t n) {
    if (!s || (n && !dst)) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    if (n > s->n) return fail(SPH_ERR_INVALID_ARGUMENT, "n exceeds the particle count");
    std::vector<float4> pos(n), vel(n), frc(n);
    std::vector<float2> dn(n);
    if ((rc = copy_out(s, pos.data(), s->b.pos, 16 * (size_t)n))) return rc;
    if ((rc = copy_out(s, vel.data(), s->b.vel, 16 * (size_t)n))) return rc;
    if ((rc = copy_out(s, frc.data(), s->b.frc, 16 * (size_t)n))) return rc;
    if ((rc = copy_out(s, dn.data(), s->b.dn, 8 * (size_t)n))) return rc;
    float *r = static_cast<float *>(dst);
    for (uint32_t i = 0; i < n; ++i) {
        float *q = r + 16 * (size_t)i;
        const float rec[16] = {pos[i].x, pos[i].y, pos[i].z, 0.0f, vel[i].x, vel[i].y,
                               vel[i].z, 0.0f,     frc[i].x, frc[i].y, frc[i].z, dn[i].x,
                               dn[i].y,  0.0f,     0.0f,     0.0f};
        std::memcpy(q, rec, sizeof rec);
    }
    return SPH_OK;
}

int sph_write_particles(sph_sim *s, const void *src, uint32_t n) {
    if (!s || (n && !src)) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    if (n != s->n) return fail(SPH_ERR_INVALID_ARGUMENT, "n must equal the particle count");
    return sph::upload(s, src, n);
}

int sph_read_cell_start(sph_sim *s, uint32_t *dst, uint64_t count) {
    if (!s || (count && !dst)) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    if (count > (uint64_t)s->uni.G + 1)
        return fail(SPH_ERR_INVALID_ARGUMENT, "count exceeds gridCount + 1");
    return copy_out(s, dst, s->b.start, 4 * (size_t)count);
}

int sph_read_sorted_order(sph_sim *s, uint32_t *dst, uint32_t count) {
    if (!s || (count && !dst)) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    if (count > s->n) return fail(SPH_ERR_INVALID_ARGUMENT, "count exceeds the particle count");
    return copy_out(s, dst, s->b.order2, 4 * (size_t)count);
}

int sph_get_counts(sph_sim *s, uint32_t *num_particles, uint64_t *grid_count) {
    if (!s) return fail(SPH_ERR_INVALID_ARGUMENT, "sim is NULL");
    int rc = sph::select(s);
    if (rc) return rc;
    if (num_particles) *num_particles = s->n;
    if (grid_count) *grid_count = s->uni.G;
    return SPH_OK;
}

int sph_time_steps(sph_sim *s, uint32_t substeps, float *ms) {
    if (!s || !ms) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
    for (uint32_t k = 0; k < substeps && e == hipSuccess && s->n; ++k)
        e = hipGraphLaunch(s->graph, s->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return sph::hip_fail(e, "sph_time_steps");
    return SPH_OK;
}

int sph_time_stages(sph_sim *s, uint32_t substeps, float *stage_ms) {
    if (!s || !stage_ms) return fail(SPH_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = sph::select(s);
    if (rc) return rc;
    for (int k = 0; k < SPH_STAGE_NUM; ++k) stage_ms[k] = 0.0f;
    if (s->n == 0) return SPH_OK;
    std::vector<hipEvent_t> ev(2 * SPH_STAGE_NUM, nullptr);
    hipError_t e = hipSuccess;
    for (auto &x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    for (uint32_t k = 0; k < substeps && e == hipSuccess; ++k) {
        for (int st = 0; st < SPH_STAGE_NUM && e == hipSuccess; ++st)
            e = sph::launch_stage(st, s->b, s->ntiles, s->stream, ev[2 * st], ev[2 * st + 1]);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        for (int st = 0; st < SPH_STAGE_NUM && e == hipSuccess; ++st) {
            float t = 0.0f;
            e = hipEventElapsedTime(&t, ev[2 * st], ev[2 * st + 1]);
            stage_ms[st] += t;
        }
    }
    for (auto x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return sph::hip_fail(e, "sph_time_stages");
    return SPH_OK;
}

}  // extern "C"
