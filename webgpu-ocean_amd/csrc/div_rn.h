// div_rn.h -- correctly rounded single division from a precomputed
// reciprocal, used by the gfx950 renderer's raster and bilateral filter
// (render.hip) and by the host test that checks it against IEEE division
// (tests/emu/div_rn_check.cpp).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define DIV_RN_HD __host__ __device__
#else
#define DIV_RN_HD
#endif

// The correctly rounded quotient RN(a / b) from y = RN(1 / b), by two
// Markstein corrections (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller
// et al., Handbook of Floating-Point Arithmetic, 2nd ed., sec. 4.7):
// q0 = RN(a y) is within two ulps of a / b; r0 = RN(a - b q0) and
// q1 = RN(q0 + r0 y) make it faithful; for a faithful q1 the remainder
// r1 = a - b q1 is exact in one fma and q = RN(q1 + r1 y) is RN(a / b).
// Valid without underflow or overflow in the steps: the raster calls it with
// a in [0, 2^24] (pixel-centre offsets) and b > 0 a normal number, the
// bilateral filter with a = -(depth difference)^2 (0 or of magnitude
// >= ~1e-14) and b = 2 (depth threshold / 3)^2.  a = -0 gives +0 (IEEE: -0),
// which neither caller can tell apart (uv >= 0; exp(+-0) = 1).
DIV_RN_HD inline float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    const float r0 = fmaf(-q0, b, a);
    const float q1 = fmaf(r0, y, q0);
    const float r1 = fmaf(-q1, b, a);
    return fmaf(r1, y, q1);
}
