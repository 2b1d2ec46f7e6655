// Division- and square-root-free forms of a few IEEE float expressions of the
// raster, each equal to the expression it replaces for every input of its
// stated domain (so the integer G-buffer and DOM words do not change).
// The header is plain C++ as well as HIP: tests/exact_arith.cpp compiles it
// with the host compiler and checks every helper against the operation it
// replaces (whole domain where it is finite, edge cases and dense samples
// elsewhere). Host and device agree because each helper uses only correctly
// rounded IEEE single operations: add, multiply, fma, divide, sqrt.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define HM_EXACT_FN __host__ __device__ __forceinline__
#else
#define HM_EXACT_FN inline
#endif

namespace hm {

HM_EXACT_FN float exact_bits_f(uint32_t u) { return __builtin_bit_cast(float, u); }
HM_EXACT_FN uint32_t exact_f_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// a / b from y = 1 / b (the correctly rounded reciprocal) by one residual
// correction (Markstein's method, IBM J. Res. Dev. 34(1), 1990): q0 = a * y,
// r = a - q0 * b (exact, one FMA), q = q0 + r * y. Used where it has been
// checked against a / b over the whole domain:
//   - a = i, b = n, integers 0 <= i <= n <= 128 (camera sample parameters);
//   - a = q, b = 16777214, integers 0 <= q <= 16777214 (DOM depth codes).
HM_EXACT_FN float div_by_recip(float a, float b, float y) {
    const float q0 = a * y;
    const float r = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(r, y, q0);
}

// 1 / sqrt(d) with both operations rounded to nearest, through a closed form
// for d within 2^10 ulps of 1 (normalizing a vector that is already unit
// length to rounding): for d = 1 + j 2^-23 (0 <= j <= 1024), sqrt(d) rounds to
// 1 + floor(j/2) 2^-23 and its reciprocal to 1 - floor(j/2) 2^-23; for
// d = 1 - j 2^-24 (1 <= j <= 1024), sqrt(d) rounds to 1 - k 2^-24 with
// k = ceil(j/2), and its reciprocal to 1 + ceil(k/2) 2^-23. Elsewhere the
// two IEEE operations.
HM_EXACT_FN float inv_sqrt_rn(float d) {
    const int32_t j = (int32_t)(exact_f_bits(d) - 0x3F800000u);
    if (j >= 0 && j <= 1024) return exact_bits_f(0x3F800000u - 2u * ((uint32_t)j >> 1));
    if (j < 0 && j >= -1024) {
        const uint32_t k = (uint32_t)(1 - j) >> 1;
        return exact_bits_f(0x3F800000u + ((k + 1u) >> 1));
    }
    return 1.0f / __builtin_sqrtf(d);
}

// Depth codes span [0, 16777214]; their reciprocal scale, correctly rounded.
constexpr float kDepthCodes = 16777214.0f;
constexpr float kInvDepthCodes = 1.0f / 16777214.0f;

// code / 16777214 for a depth code (exactly the IEEE quotient).
HM_EXACT_FN float depth_code_fraction(uint32_t code) { return div_by_recip((float)code, kDepthCodes, kInvDepthCodes); }

// min(trunc(x / b), 15) for x >= 0 (NaN -> 0) and 2^-100 <= b <= 2^100,
// from y = 1 / b (correctly rounded) without dividing unless x * y lies
// within 2^-17 of an integer in [1, 15]: below 16, x * y and the rounded
// quotient differ by less than 2^-18 (two roundings of relative size 2^-24
// each on x / b, plus half an ulp of the quotient), so outside that band
// both truncate to the same integer.
HM_EXACT_FN uint32_t slice_of(float x, float b, float y) {
    const float t = x * y;
    if (!(t >= 1.0f - 0x1p-17f)) return 0u;  // also NaN
    if (t >= 15.0f + 0x1p-17f) return 15u;
    const float k = __builtin_floorf(t), frac = t - k;  // exact below 16
    if (frac >= 0x1p-17f && frac <= 1.0f - 0x1p-17f) return (uint32_t)k;
    const float q = x / b;
    return q > 0.0f ? (q < 15.0f ? (uint32_t)q : 15u) : 0u;
}

// True when the 6-bit code trunc(clamp(v, 0, 1) * 63 + 0.5) is the same
// for every v = c * (1 - t) + c * t, t in [0, 1] (both products and the sum
// rounded): such a v is within 3 ulps of c, which moves clamp(v) * 63 + 0.5
// by less than 2^-16 for |c| <= 2, so the code is c's own whenever
// clamp(c) * 63 + 0.5 lies at least 2^-12 from an integer.
HM_EXACT_FN bool quant63_fixed_under_mix(float c) {
    if (!(c >= -2.0f && c <= 2.0f)) return false;  // also NaN
    const float a = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    const float x = a * 63.0f + 0.5f;
    const float f = x - __builtin_floorf(x);
    return f >= 0x1p-12f && f <= 1.0f - 0x1p-12f;
}

// The largest float x with sqrt(x) <= h (sqrt correctly rounded), so that
// sqrt(d) <= h  <=>  d <= sqrt_le_bound(h) for every float d >= 0 (rounded
// sqrt is monotone). Domain: h = 0, h >= 2^-60 (h * h normal), +inf, or NaN
// (the bound is NaN, which no d satisfies, as sqrt(d) <= NaN fails).
HM_EXACT_FN float sqrt_le_bound(float h) {
    if (!(h > 0.0f)) return h;  // 0 -> 0 (only d = 0 passes), NaN -> NaN
    float x = h * h;
    if (!(x < __builtin_huge_valf()))  // h * h overflows: every finite d passes, inf only for h = inf
        return h < __builtin_huge_valf() ? exact_bits_f(0x7f7fffffu) : x;
    // x is within an ulp or two of the bound: walk down while too large, then up while the next one passes
    for (int k = 0; k < 8 && __builtin_sqrtf(x) > h; ++k) x = exact_bits_f(exact_f_bits(x) - 1u);
    for (int k = 0; k < 8; ++k) {
        const float nx = exact_bits_f(exact_f_bits(x) + 1u);
        if (!(__builtin_sqrtf(nx) <= h)) break;
        x = nx;
    }
    return x;
}

// (a / b < 0) and (a / b > 0) for b >= 1e-5 (or NaN), without dividing where
// the quotient cannot round to zero: for |a| >= 2^-60 and b <= 2^60 the
// quotient is at least 2^-120 in magnitude (normal), so its sign is a's.
HM_EXACT_FN bool div_below_zero(float a, float b) {
    const float m = a < 0.0f ? -a : a;
    if (m >= 0x1p-60f && b <= 0x1p60f) return a < 0.0f;
    return a / b < 0.0f;
}
HM_EXACT_FN bool div_above_zero(float a, float b) {
    const float m = a < 0.0f ? -a : a;
    if (m >= 0x1p-60f && b <= 0x1p60f) return a > 0.0f;
    return a / b > 0.0f;
}

}  // namespace hm
