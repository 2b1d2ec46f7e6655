// Host-side check of bevy_hair_amd/csrc/hm_exact.h: every helper against the
// IEEE operation it replaces. Prints one line per check and "ok" at the end;
// exits nonzero on the first mismatch. Built and run by test_exact_arith.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <vector>

#include "hm_exact.h"

using hm::exact_bits_f;
using hm::exact_f_bits;

static int fail(const char* what, double a, double b) {
    std::printf("FAIL %s: a=%.9g b=%.9g\n", what, a, b);
    return 1;
}

// a / b with the divide done as a volatile float operation (no folding).
static float ieee_div(float a, float b) {
    volatile float x = a, y = b;
    return x / y;
}

int main() {
    // i / n for the camera sample parameters.
    long pairs = 0;
    for (int n = 1; n <= 128; ++n) {
        const float fn = (float)n, y = ieee_div(1.0f, fn);
        for (int i = 0; i <= n; ++i, ++pairs)
            if (exact_f_bits(hm::div_by_recip((float)i, fn, y)) != exact_f_bits(ieee_div((float)i, fn)))
                return fail("div_by_recip i/n", i, n);
    }
    std::printf("i/n: %ld pairs equal\n", pairs);

    // 1 / sqrt(d) near 1 (closed form) and away from it.
    {
        long checked = 0;
        for (int j = -8192; j <= 8192; ++j, ++checked) {
            const float d = exact_bits_f(0x3F800000u + (uint32_t)j);
            const float want = ieee_div(1.0f, std::sqrt(d));
            if (exact_f_bits(hm::inv_sqrt_rn(d)) != exact_f_bits(want)) return fail("inv_sqrt_rn near 1", d, want);
        }
        for (uint64_t u = 0; u <= 0x7f800000ull; u += 1021u, ++checked) {
            const float d = exact_bits_f((uint32_t)u);
            const float want = ieee_div(1.0f, std::sqrt(d));
            if (exact_f_bits(hm::inv_sqrt_rn(d)) != exact_f_bits(want)) return fail("inv_sqrt_rn", d, want);
        }
        std::printf("inverse sqrt: %ld values\n", checked);
    }

    // 6-bit codes of a strand constant mixed with itself: where
    // quant63_fixed_under_mix(c) holds, every t = i / n (n <= 96) gives c's code.
    {
        auto code = [](float v) -> uint32_t {
            const float a = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);  // clampf on non-NaN input
            const float x = a * 63.0f + 0.5f;
            const uint32_t q = x > 0.0f ? (uint32_t)x : 0u;
            return q < 63u ? q : 63u;
        };
        long checked = 0, fixed = 0;
        std::vector<float> cs;
        for (int k = 0; k <= 64; ++k) {  // around every code boundary (k - 0.5) / 63, and dense elsewhere
            const float edge = ((float)k - 0.5f) / 63.0f;
            for (int d = -2000; d <= 2000; d += 3) cs.push_back(exact_bits_f(exact_f_bits(edge < 0 ? 0.001f : edge) + (uint32_t)d));
        }
        for (uint32_t u = exact_f_bits(1e-6f); u < exact_f_bits(1.0f); u += 65537u) cs.push_back(exact_bits_f(u));
        for (float c : {0.0f, 1.0f, -0.5f, 1.5f, 2.0f, -2.0f}) cs.push_back(c);
        for (float c : cs) {
            if (!hm::quant63_fixed_under_mix(c)) continue;
            ++fixed;
            const uint32_t want = code(c);
            for (int n = 1; n <= 96; ++n)
                for (int i = 0; i <= n; ++i, ++checked) {
                    const float t = ieee_div((float)i, (float)n);
                    const float v = c * (1.0f - t) + c * t;
                    if (code(v) != want) return fail("quant63_fixed_under_mix", c, t);
                }
        }
        std::printf("mixed codes: %ld fixed constants, %ld pairs\n", fixed, checked);
    }

    // DOM depth codes / 16777214.
    {
        const float b = 16777214.0f, y = ieee_div(1.0f, b);
        for (uint32_t q = 0; q <= 16777214u; ++q)
            if (exact_f_bits(hm::div_by_recip((float)q, b, y)) != exact_f_bits(ieee_div((float)q, b)))
                return fail("div_by_recip q/16777214", q, b);
        if (exact_f_bits(y) != exact_f_bits(hm::kInvDepthCodes)) return fail("kInvDepthCodes", y, hm::kInvDepthCodes);
        for (uint32_t q = 0; q <= 16777214u; q += 7u)
            if (exact_f_bits(hm::depth_code_fraction(q)) != exact_f_bits(ieee_div((float)q, b)))
                return fail("depth_code_fraction", q, b);
        std::printf("depth codes: 16777215 values equal\n");
    }

    // DOM slices: min(trunc(x / b), 15) over dense x for several layer
    // thicknesses, and x placed on and around every integer boundary.
    {
        long checked = 0;
        auto ref = [](float x, float b) -> uint32_t {
            const float q = ieee_div(x, b);
            const uint32_t s = q > 0.0f ? (q < 4294967040.0f ? (uint32_t)q : 4294967040u) : 0u;
            return s < 15u ? s : 15u;
        };
        const float bs[] = {6.25e-5f, 1e-3f / 16.0f, 0.0123f, 0.1f, 0.37f, 1.0f, 3.3333333f, 62.5f, 1000.0f, 0x1p-100f, 0x1p100f};
        for (float b : bs) {
            const float y = ieee_div(1.0f, b);
            const uint32_t b_lo = exact_f_bits(b * 0.001f), b_hi = exact_f_bits(b * 20.0f);
            for (uint32_t u = b_lo; u <= b_hi; u += 37u, ++checked) {
                const float x = exact_bits_f(u);
                if (hm::slice_of(x, b, y) != ref(x, b)) return fail("slice_of", x, b);
            }
            for (int k = 1; k <= 16; ++k) {
                const float c = (float)k * b;
                for (int d = -300; d <= 300; ++d, ++checked) {
                    const float x = exact_bits_f(exact_f_bits(c) + (uint32_t)d);
                    if (hm::slice_of(x, b, y) != ref(x, b)) return fail("slice_of boundary", x, b);
                }
            }
            for (float x : {0.0f, -0.0f, NAN, INFINITY, 1e30f})
                if (hm::slice_of(x, b, y) != ref(x, b)) return fail("slice_of edge", x, b);
        }
        std::printf("slices: %ld values\n", checked);
    }

    // sqrt_le_bound: sqrt(T) <= h < sqrt(next(T)) over a dense stride of
    // positive h (2^-60 .. max float), plus edge values.
    {
        long checked = 0;
        auto check = [&](float h) -> bool {
            const float t = hm::sqrt_le_bound(h);
            ++checked;
            if (std::isnan(h)) return std::isnan(t);
            if (h == 0.0f) return t == 0.0f;
            if (!(std::sqrt(t) <= h)) return false;
            if (std::isinf(t)) return std::isinf(h);
            const float nt = exact_bits_f(exact_f_bits(t) + 1u);
            return std::sqrt(nt) > h;
        };
        const uint32_t lo = exact_f_bits(0x1p-60f), hi = exact_f_bits(exact_bits_f(0x7f7fffffu));
        for (uint64_t b = lo; b <= hi; b += 331u)
            if (!check(exact_bits_f((uint32_t)b))) return fail("sqrt_le_bound", exact_bits_f((uint32_t)b), 0);
        const float edges[] = {0.0f, 0.025f, 0.5f * 0.05f, 0.5f, 1.0f, 1e19f, 1.8446743e19f, 1.9e19f, 3e38f,
                               exact_bits_f(0x7f7fffffu), INFINITY, NAN, 0x1p-60f};
        for (float h : edges)
            if (!check(h)) return fail("sqrt_le_bound edge", h, 0);
        // Exhaustive over d around the bound for a few h: sqrt(d) <= h <=> d <= T.
        for (float h : {0.025f, 0.03f, 0.1f, 0.37f, 1.0f, 2.5f}) {
            const float t = hm::sqrt_le_bound(h);
            for (int k = -64; k <= 64; ++k) {
                const float d = exact_bits_f(exact_f_bits(t) + (uint32_t)k);
                if ((std::sqrt(d) <= h) != (d <= t)) return fail("sqrt_le_bound neighbourhood", h, d);
            }
        }
        std::printf("sqrt bound: %ld values\n", checked);
    }

    // Quotient signs without dividing, for b >= 1e-5.
    {
        long checked = 0;
        const float bs[] = {1e-5f, 1.5e-5f, 0.01f, 0.3f, 1.0f, 7.25f, 1000.0f, 0x1p59f, 0x1p60f,
                            0x1.000002p60f, 1e30f, 3e38f, INFINITY, NAN};
        for (float b : bs) {
            for (uint64_t u = 0; u <= 0xffffffffull; u += 4093u) {
                const float a = exact_bits_f((uint32_t)u);
                ++checked;
                if (hm::div_below_zero(a, b) != (ieee_div(a, b) < 0.0f)) return fail("div_below_zero", a, b);
                if (hm::div_above_zero(a, b) != (ieee_div(a, b) > 0.0f)) return fail("div_above_zero", a, b);
            }
            for (float a : {0.0f, -0.0f, 0x1p-60f, -0x1p-60f, 0x1p-61f, -0x1p-149f, 0x1p-149f, INFINITY, -INFINITY, NAN}) {
                if (hm::div_below_zero(a, b) != (ieee_div(a, b) < 0.0f)) return fail("div_below_zero edge", a, b);
                if (hm::div_above_zero(a, b) != (ieee_div(a, b) > 0.0f)) return fail("div_above_zero edge", a, b);
            }
        }
        std::printf("quotient signs: %ld pairs\n", checked);
    }
    std::printf("ok\n");
    return 0;
}
