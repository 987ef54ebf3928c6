// mpm_fused.h -- launch interface of the production (binned, LDS-windowed)
// substep kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "mpm_device.h"

namespace mpm {

// Bin keys are tile-major: tiles of (2^kTileBits)^3 cells in x-major tile
// order, cells x-major inside a tile.
constexpr int kTileBits = 2;

struct KeyGeom {
    int32_t dim[3];   // ceil(init_box): cells per axis used for keys
    uint32_t nt[3];   // tiles per axis
    uint64_t nkeys;   // nt0*nt1*nt2 * 2^(3*kTileBits)
};

KeyGeom make_key_geom(const float init_box[3]);

// Buffers of the substep kernels.
struct FusedArgs {
    float *state;          // SoA particle state in bin order
    uint32_t *id;          // storage slot -> caller particle index
    uint32_t stride;
    uint32_t n;
    int4 *grid;
    uint32_t G;
    float4 *posvel;        // indexed by caller particle index
    const Uniforms *uni;
    Params prm;
    int ppt;               // particles per thread (1, 2, 4): chunk = 256 * ppt
    int wmax;              // LDS window budget in cells (16 B each)
};

// Binning (counting sort by cell key): reads state/id of FusedArgs and writes
// the permuted copy to state_out/id_out.
struct BinArgs {
    KeyGeom kg;
    uint32_t *key;         // n
    uint32_t *rank;        // n
    uint32_t *count;       // nkeys, zero on entry, left zero on exit
    uint32_t *offset;      // nkeys
    uint32_t *block_sums;  // ceil(nkeys / 4096)
    float *state_out;
    uint32_t *id_out;
};

hipError_t launch_bin(const FusedArgs &a, const BinArgs &b, hipStream_t st);
hipError_t launch_p2g1(const FusedArgs &a, hipStream_t st);
hipError_t launch_p2g2(const FusedArgs &a, hipStream_t st);
hipError_t launch_g2p(const FusedArgs &a, hipStream_t st);
hipError_t launch_iota(uint32_t *id, uint32_t n, hipStream_t st);

}  // namespace mpm
