// Device-side math of the strand rasterizer (gfx950).
//
// Integer keys (G-buffer words, DOM depths and counts) must come out bit-exact
// against the CPU restatement, so every float operation here is written in the
// order the WGSL writes it, the library is compiled with -ffp-contract=off and
// correctly rounded f32 divide/sqrt, and min/max/conversions use explicit
// comparisons with WGSL semantics:
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z; mat*vec sums columns 0..3;
//   inverseSqrt(x) = 1/sqrt(x); normalize(v) = v * inverseSqrt(dot(v,v));
//   mix(a,b,t) = a*(1-t) + b*t; f32->u32/i32 truncate and saturate to
//   [0, 4294967040] / [-2^31, 2147483520]; i32 add/sub/abs wrap.
// Where a quotient or root has a cheaper form that equals it for every input
// of its domain (hm_exact.h), that form is used instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hair_mi.h"
#include "hm_exact.h"

namespace hm {

struct F2 { float x, y; };
struct F3 { float x, y, z; };
struct F4 { float x, y, z, w; };

__device__ __forceinline__ F2 operator+(F2 a, F2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ F2 operator-(F2 a, F2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ F2 operator*(F2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator/(F3 a, F3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F4 operator+(F4 a, F4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ F4 operator*(F4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
__device__ __forceinline__ F3 xyz(F4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ F3 splat3(float s) { return {s, s, s}; }

// min/max/clamp with the comparison the WGSL builtins reduce to on non-NaN input.
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
__device__ __forceinline__ F3 vmin(F3 a, F3 b) { return {mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
__device__ __forceinline__ F3 vmax(F3 a, F3 b) { return {mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
__device__ __forceinline__ F3 vclamp(F3 a, float lo, float hi) {
    return {clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi)};
}

__device__ __forceinline__ float dot2(F2 a, F2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len2(F2 a) { return __builtin_sqrtf(dot2(a, a)); }
__device__ __forceinline__ float len3(F3 a) { return __builtin_sqrtf(dot3(a, a)); }
__device__ __forceinline__ float rdiv(float a, float b) { return a / b; }
__device__ __forceinline__ float inv_sqrt(float x) { return 1.0f / __builtin_sqrtf(x); }

// Correctly rounded sqrt and divide on operands where the compiler's IEEE
// expansions reduce to their core: the same instruction sequences, hence the
// same bits, without the steps that only act outside these domains.
// sqrt: for 2^-96 <= x < inf the expansion neither scales x (it does below
// 2^-96) nor selects x itself (zero, infinity): what is left is the hardware
// root s and the choice among s and its two neighbours by the signs of the
// residuals x - s' * s.
__device__ __forceinline__ float sqrt_rn_mid(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    const float r = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : r;
}
// divide: for 2^-60 <= |a|, b <= 2^60 the expansion's v_div_scale leaves both
// operands as they are, its v_div_fmas is then a plain fma, and v_div_fixup
// has no zero, infinity or NaN to handle: what is left is the refined
// reciprocal y of b and two residual corrections of a * y.
__device__ __forceinline__ float div_rn_mid(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float y = __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
    const float q0 = a * y;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
    return __builtin_fmaf(__builtin_fmaf(-b, q1, a), y, q1);
}
// div_below_zero / div_above_zero (hm_exact.h, checked there on the host)
// with the division of their rare branch kept behind it: the compiler turns
// the branch into a select otherwise, and divides for every call.
__device__ __forceinline__ bool div_below_zero_rare(float a, float b) {
    if (fabsf(a) >= 0x1p-60f && b <= 0x1p60f) return a < 0.0f;
    asm volatile("" : "+v"(a));
    return a / b < 0.0f;
}
__device__ __forceinline__ bool div_above_zero_rare(float a, float b) {
    if (fabsf(a) >= 0x1p-60f && b <= 0x1p60f) return a > 0.0f;
    asm volatile("" : "+v"(a));
    return a / b > 0.0f;
}
__device__ __forceinline__ bool div_mid_ok(float a) {
    const float m = fabsf(a);
    return m >= 0x1p-60f && m <= 0x1p60f;
}
__device__ __forceinline__ F3 normalize3(F3 v) { return v * inv_sqrt(dot3(v, v)); }
// normalize3 for a vector expected to be unit length to rounding already
// (equal to normalize3 for every input; cheaper when the length is near 1).
__device__ __forceinline__ F3 normalize3_unit(F3 v) { return v * inv_sqrt_rn(dot3(v, v)); }
__device__ __forceinline__ float mixs(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ F2 mix2(F2 a, F2 b, float t) { return {mixs(a.x, b.x, t), mixs(a.y, b.y, t)}; }
__device__ __forceinline__ F3 mix3(F3 a, F3 b, float t) {
    return {mixs(a.x, b.x, t), mixs(a.y, b.y, t), mixs(a.z, b.z, t)};
}

// Saturating float -> integer conversions (NaN -> 0, out-of-range -> the
// nearest representable bound, else truncation). The input is clamped with
// selects first, so the conversion itself is always in range and the code is
// branch-free.
__device__ __forceinline__ uint32_t to_u32(float x) {
    float c = x > 0.0f ? x : 0.0f;  // NaN -> 0
    c = c < 4294967040.0f ? c : 4294967040.0f;
    return (uint32_t)c;
}
__device__ __forceinline__ int32_t to_i32(float x) {
    float c = x != x ? 0.0f : x;
    c = c > -2147483648.0f ? c : -2147483648.0f;
    c = c < 2147483520.0f ? c : 2147483520.0f;
    return (int32_t)c;
}
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int32_t wabs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }
__device__ __forceinline__ float u2f(uint32_t u) { return (float)u; }
__device__ __forceinline__ float i2f(int32_t i) { return (float)i; }

constexpr float kPi = 3.14159265f;

__device__ __forceinline__ F4 ld4(const float* p) {
    float4 v = *reinterpret_cast<const float4*>(p);
    return {v.x, v.y, v.z, v.w};
}

// Column-major 4x4 times (x, y, z, w); rows are independent, so callers that
// need only some rows compute exactly the values the full product would.
__device__ __forceinline__ float mat_row(const float* m, int r, float x, float y, float z, float w) {
    return m[r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r] * w;
}

__device__ __forceinline__ uint32_t pcg(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ F2 pcg_uv(uint32_t seed) {
    uint32_t a = pcg(seed);
    uint32_t b = pcg(a ^ 0x9e3779b9u);
    const float s = 1.0f / 4294967296.0f;
    return {u2f(a) * s, u2f(b) * s};
}

__device__ __forceinline__ uint64_t pack_gb(uint32_t depth, uint32_t tx, uint32_t ty, uint32_t u, uint32_t v,
                                            uint32_t w, uint32_t ao) {
    uint32_t high = (depth << 8u) | (tx & 0xffu);
    uint32_t low = ((ty & 0xffu) << 24u) | ((u & 0x3fu) << 18u) | ((v & 0x3fu) << 12u) | ((w & 0x3fu) << 6u) |
                   (ao & 0x3fu);
    return ((uint64_t)high << 32u) | (uint64_t)low;
}

__device__ __forceinline__ bool gb_empty(uint64_t word) { return ((uint32_t)(word >> 32u) >> 8u) == 0x00ffffffu; }
__device__ __forceinline__ uint32_t gb_depth(uint64_t word) { return (uint32_t)(word >> 32u) >> 8u; }
__device__ __forceinline__ uint32_t gb_tx(uint64_t word) { return (uint32_t)(word >> 32u) & 0xffu; }
__device__ __forceinline__ uint32_t gb_ty(uint64_t word) { return (uint32_t)word >> 24u; }

__device__ __forceinline__ uint32_t quantize_depth(float dist, float near_, float far_) {
    float q = clampf(rdiv(dist - near_, mx(far_ - near_, 1e-3f)), 0.0f, 1.0f);
    uint32_t v = to_u32(q * 16777214.0f);
    return v < 0x00fffffeu ? v : 0x00fffffeu;
}

__device__ __forceinline__ float dequantize_depth(uint32_t depth, float near_, float far_) {
    return mixs(near_, far_, depth_code_fraction(depth));
}

__device__ __forceinline__ void oct_encode(F3 n_in, uint32_t& ox, uint32_t& oy) {
    F3 n = normalize3_unit(n_in);  // callers pass normalized tangents
    float denom = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float m = mx(denom, 1e-6f);
    float px = rdiv(n.x, m), py = rdiv(n.y, m);
    if (n.z < 0.0f) {
        float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    float ex = clampf(px * 0.5f + 0.5f, 0.0f, 1.0f);
    float ey = clampf(py * 0.5f + 0.5f, 0.0f, 1.0f);
    ox = to_u32(ex * 255.0f + 0.5f);
    oy = to_u32(ey * 255.0f + 0.5f);
}

// oct_encode(safe_normalize(v)) through the hardware reciprocal square root
// and reciprocal (each within 1 ulp) instead of the IEEE divides and roots.
// Against the exact chain, the folded coordinates move by < 36 ulps of 1 and
// the codes' pre-truncation values ex * 255 + 0.5 by < 3.2e-4; when both lie
// farther than 2^-10 from an integer their truncations are the exact codes,
// and the function returns true. Otherwise (or for lengths near zero or
// huge, or components near zero whose sign decides the fold) it returns
// false and the caller takes the exact path.
__device__ __forceinline__ bool oct_encode_fast(F3 v, uint32_t& ox, uint32_t& oy) {
    const float l2 = dot3(v, v);
    if (!(l2 >= 1e-9f && l2 <= 1e30f)) return false;  // safe_normalize's 1e-10 cut-off, NaN
    const F3 n = v * __builtin_amdgcn_rsqf(l2);
    const float ax = fabsf(n.x), ay = fabsf(n.y);
    if (!(ax >= 0x1p-60f && ay >= 0x1p-60f)) return false;
    const float r = __builtin_amdgcn_rcpf(ax + ay + fabsf(n.z));
    float px = n.x * r, py = n.y * r;
    if (n.z < 0.0f) {
        const float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    const float dx = clampf(px * 0.5f + 0.5f, 0.0f, 1.0f) * 255.0f + 0.5f;
    const float dy = clampf(py * 0.5f + 0.5f, 0.0f, 1.0f) * 255.0f + 0.5f;
    const float fx = dx - floorf(dx), fy = dy - floorf(dy);
    if (!(fx >= 0x1p-10f && fx <= 1.0f - 0x1p-10f && fy >= 0x1p-10f && fy <= 1.0f - 0x1p-10f)) return false;
    ox = (uint32_t)dx;
    oy = (uint32_t)dy;
    return true;
}

__device__ __forceinline__ F3 oct_decode(uint32_t tx, uint32_t ty) {
    float ex = rdiv(u2f(tx), 255.0f), ey = rdiv(u2f(ty), 255.0f);
    float fx = ex * 2.0f - 1.0f, fy = ey * 2.0f - 1.0f;
    F3 n = {fx, fy, 1.0f - fabsf(fx) - fabsf(fy)};
    if (n.z < 0.0f) {
        float ox = n.x, oy = n.y;
        n.x = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
        n.y = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
    }
    return normalize3(n);
}

__device__ __forceinline__ F3 safe_normalize(F3 v) {
    float l2 = dot3(v, v);
    if (l2 < 1e-10f) return {0.0f, 1.0f, 0.0f};
    return v * inv_sqrt(l2);
}
// safe_normalize through the mid-range forms where 1e-10 <= |v|^2 < 2^120
// (the root then lies in [1e-5, 2^60]); the same results for every v.
__device__ __forceinline__ F3 safe_normalize_mid(F3 v) {
    const float l2 = dot3(v, v);
    if (!(l2 >= 1e-10f && l2 < 0x1p120f)) return safe_normalize(v);
    return v * div_rn_mid(1.0f, sqrt_rn_mid(l2));
}
// safe_normalize for a vector expected to be near unit length (same results).
__device__ __forceinline__ F3 safe_normalize_unit(F3 v) {
    float l2 = dot3(v, v);
    if (l2 < 1e-10f) return {0.0f, 1.0f, 0.0f};
    return v * inv_sqrt_rn(l2);
}

__device__ __forceinline__ F3 calc_tangent(uint32_t idx, uint32_t count, F3 prev, F3 curr, F3 next) {
    F3 d;
    float scale;
    if (idx == 0u) {
        d = next - curr;
        scale = len3(d);
    } else if (idx + 1u >= count) {
        d = curr - prev;
        scale = len3(d);
    } else {
        d = next - prev;
        scale = len3(d) * 0.5f;
    }
    return safe_normalize(d) * mx(scale, 1e-4f);
}

__device__ __forceinline__ F3 hermite(F3 p0, F3 p1, F3 m0, F3 m1, float t) {
    float t2 = t * t;
    float t3 = t2 * t;
    float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f;
    float h10 = t3 - 2.0f * t2 + t;
    float h01 = -2.0f * t3 + 3.0f * t2;
    float h11 = t3 - t2;
    return h00 * p0 + h10 * m0 + h01 * p1 + h11 * m1;
}

__device__ __forceinline__ F4 bilinear4(F4 c00, F4 c10, F4 c01, F4 c11, F2 uv) {
    float u = uv.x, v = uv.y;
    return c00 * ((1.0f - u) * (1.0f - v)) + c10 * (u * (1.0f - v)) + c01 * ((1.0f - u) * v) + c11 * (u * v);
}

__device__ __forceinline__ uint32_t quant63(float c) {
    uint32_t q = to_u32(clampf(c, 0.0f, 1.0f) * 63.0f + 0.5f);
    return q < 63u ? q : 63u;
}

// Per-pass view data the kernels read (copied from hm_params at launch).
struct View {
    float clip[16];
    float wfc[16];
    float wfm[16];
    float lclip[16];
    F4 camera_pos, camera_forward, screen, appearance, filter_params, albedo;
    uint32_t flags[4], pass_mode[4];
    F4 light_eye, light_forward, dom_info;
};

__device__ __forceinline__ F3 to_world(const View& p, F3 q) {
    return {mat_row(p.wfm, 0, q.x, q.y, q.z, 1.0f), mat_row(p.wfm, 1, q.x, q.y, q.z, 1.0f),
            mat_row(p.wfm, 2, q.x, q.y, q.z, 1.0f)};
}

// d is a normalized model-space direction: with a world_from_model of unit
// scale its image is unit length to rounding.
__device__ __forceinline__ F3 dir_world(const View& p, F3 d) {
    return safe_normalize_unit(F3{mat_row(p.wfm, 0, d.x, d.y, d.z, 0.0f), mat_row(p.wfm, 1, d.x, d.y, d.z, 0.0f),
                             mat_row(p.wfm, 2, d.x, d.y, d.z, 0.0f)});
}

struct Projected {
    F2 screen;
    float view_dist;
    bool in_front;
};

__device__ __forceinline__ Projected project_world(const View& p, F3 w) {
    float cx = mat_row(p.clip, 0, w.x, w.y, w.z, 1.0f);
    float cy = mat_row(p.clip, 1, w.x, w.y, w.z, 1.0f);
    float cw = mat_row(p.clip, 3, w.x, w.y, w.z, 1.0f);
    float ww = mx(cw, 1e-5f);
    float ndc_x = rdiv(cx, ww), ndc_y = rdiv(cy, ww);
    float sx = (ndc_x * 0.5f + 0.5f) * p.screen.x;
    float sy = (1.0f - (ndc_y * 0.5f + 0.5f)) * p.screen.y;
    float dist = dot3(w - xyz(p.camera_pos), xyz(p.camera_forward));
    return {{sx, sy}, dist, cw > 0.0f && dist > p.screen.z * 0.5f};
}

__device__ __forceinline__ F3 reconstruct_world(const View& p, int32_t px, int32_t py, float view_dist) {
    float ux = rdiv(i2f(px) + 0.5f, mx(p.screen.x, 1.0f));
    float uy = rdiv(i2f(py) + 0.5f, mx(p.screen.y, 1.0f));
    float ndc_x = ux * 2.0f - 1.0f;
    float ndc_y = (1.0f - uy) * 2.0f - 1.0f;
    float hx = mat_row(p.wfc, 0, ndc_x, ndc_y, 0.0f, 1.0f);
    float hy = mat_row(p.wfc, 1, ndc_x, ndc_y, 0.0f, 1.0f);
    float hz = mat_row(p.wfc, 2, ndc_x, ndc_y, 0.0f, 1.0f);
    float hw = mx(mat_row(p.wfc, 3, ndc_x, ndc_y, 0.0f, 1.0f), 1e-5f);
    F3 far_p = {rdiv(hx, hw), rdiv(hy, hw), rdiv(hz, hw)};
    F3 d = far_p - xyz(p.camera_pos);
    float denom = mx(dot3(d, xyz(p.camera_forward)), 1e-4f);
    return xyz(p.camera_pos) + d * rdiv(view_dist, denom);
}

__device__ __forceinline__ uint32_t control_points_for_lod(float l) {
    float c_raw = mx(floorf(__builtin_sqrtf(clampf(l, 0.0f, 1.0f)) * 127.0f), 16.0f);
    // c_raw is an integer in [16, 127]: pow(2, floor(log2(c_raw - 1))) is its
    // highest power of two not above c_raw - 1, taken exactly from the bits.
    uint32_t k = to_u32(c_raw - 1.0f);
    uint32_t snapped = (1u << (31 - __clz(k))) + 1u;
    return snapped < 127u ? snapped : 127u;
}

}  // namespace hm
