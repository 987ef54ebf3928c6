// Checks div_rn (webgpu-ocean_amd/csrc/div_rn.h) against IEEE single
// division, bit for bit: random operands over the ranges the raster uses
// (pixel-centre offsets a in [0, 2^14], rectangle extents b in [2^-10, 2^14]),
// random significands over whole binades, and structured cases (a = 0,
// a = b, powers of two, a = k + 0.5 - b / 3).  Prints the number of
// mismatches (0 expected) and exits 1 on any.  Built with
// -ffp-contract=off so that a / b is the IEEE quotient.
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "div_rn.h"

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static uint64_t next_u64() {  // splitmix64
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static float uniform(float lo, float hi) {
    return lo + (hi - lo) * (float)((next_u64() >> 40) * 0x1p-24);
}
static float with_exponent(int e) {  // random significand in [2^e, 2^(e+1))
    const uint32_t bits = (uint32_t)(e + 127) << 23 | (uint32_t)(next_u64() & 0x7fffff);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

#if defined(NO_CORRECTION)
// A planted weakening (built only by the test that shows the check catches
// it): the uncorrected product RN(a RN(1/b)), up to two ulps off.
static float quotient(float a, float b, float y) { return a * y; }
#else
static float quotient(float a, float b, float y) { return div_rn(a, b, y); }
#endif

static uint64_t g_bad = 0, g_checked = 0;
static void check(float a, float b) {
    const float want = a / b;
    const float got = quotient(a, b, 1.0f / b);
    uint32_t wb, gb;
    std::memcpy(&wb, &want, 4);
    std::memcpy(&gb, &got, 4);
    ++g_checked;
    if (wb != gb && g_bad++ < 10)
        std::printf("mismatch a=%a b=%a: %a vs %a\n", (double)a, (double)b, (double)got,
                    (double)want);
}

int main(int argc, char **argv) {
    const uint64_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 20000000ull;
    for (uint64_t i = 0; i < n; ++i) {
        const float b = uniform(0x1p-10f, 0x1p14f);
        check(uniform(0.0f, b), b);             // offsets inside the rectangle
        check(uniform(0.0f, 0x1p14f), b);
        check(with_exponent((int)(next_u64() % 38) - 14), with_exponent((int)(next_u64() % 24) - 10));
    }
    for (int k = 0; k < 4096; ++k) {
        const float b = 0.37f + 0.731f * (float)k;
        check(0.0f, b);
        check(b, b);
        check((float)k + 0.5f, 1.0f);
        check((float)k + 0.5f - b / 3.0f, b);
        check(std::ldexp(1.0f, k % 24), std::ldexp(1.0f, (k / 24) % 24 - 10));
    }
    std::printf("div_rn: %" PRIu64 " quotients, %" PRIu64 " mismatches\n", g_checked, g_bad);
    return g_bad ? 1 : 0;
}
