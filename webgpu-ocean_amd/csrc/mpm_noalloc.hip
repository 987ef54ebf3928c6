// mpm_noalloc.hip -- device scratch of the allocation-free step.
//
// The grid and uniforms of mpm_step_noalloc live in one file-scope device
// array, so that entry point never allocates.  Each GPU holds its own copy of
// the array at its own address, looked up once per device with
// hipGetSymbolAddress; host code uses only those addresses (the array's name
// in host code is a host-side stand-in, not a device address).
#include <mutex>
#include <vector>

#include "mpm_internal.h"

namespace mpm {
namespace {

// Cell cap of the reference's cell buffer: 64^3 cells of 16 B
// (mls-mpm/mls-mpm.ts:13-15, 122-130), plus 256 B for the uniforms.
constexpr uint32_t kMaxCells = 64u * 64u * 64u;
constexpr size_t kUniformBytes = 256;
static_assert(sizeof(Uniforms) <= kUniformBytes, "uniform slot");

std::mutex g_mu;
std::vector<void *> g_addr;  // per device ordinal, nullptr until looked up

}  // namespace

// Holds the uniforms (first 256 B) and up to kMaxCells grid cells; every call
// on a GPU shares that GPU's copy (4 MiB + 256 B per GPU, kept until the
// process ends).
static __device__ __align__(256) unsigned char scratch[kUniformBytes + 16 * (size_t)kMaxCells];

hipError_t noalloc_scratch(Uniforms **uni, int4 **grid, uint32_t *max_cells) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (dev < 0) return hipErrorInvalidDevice;
        if ((size_t)dev >= g_addr.size()) g_addr.resize((size_t)dev + 1, nullptr);
        if (!g_addr[dev]) {
            // the calling thread's current device is `dev`: its copy of the array
            e = hipGetSymbolAddress(&g_addr[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                g_addr[dev] = nullptr;
                return e;
            }
        }
        base = g_addr[dev];
    }
    *uni = static_cast<Uniforms *>(base);
    *grid = reinterpret_cast<int4 *>(static_cast<unsigned char *>(base) + kUniformBytes);
    *max_cells = kMaxCells;
    return hipSuccess;
}

}  // namespace mpm
