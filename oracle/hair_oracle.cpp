/*
 * THE ORACLE — test infrastructure, not product code.
 *
 * A scalar CPU restatement of bevy_hair's deferred strand-rasterization
 * kernels, written line by line from the reference WGSL so that the HIP
 * kernels can be checked against it:
 *   common.wgsl (helpers), clear.wgsl, clear_dom.wgsl, lod.wgsl, scan.wgsl,
 *   raster.wgsl (camera, DOM depth, DOM opacity), shade.wesl (fs, BSDF, DOM
 *   lookup; Bevy light helpers restated), filter.wgsl.
 * Paths are relative to the reference repository root.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark may
 * load this library. It is templated on the arithmetic type: the float
 * instantiation is the parity target (bit-exact integer outputs, 1e-3 relative
 * RGB), the double instantiation states the float arithmetic's error.
 *
 * Arithmetic conventions (the reference runs WGSL through naga/wgpu; these
 * are the build's definitions, shared with the HIP kernels):
 *   - no FMA contraction, operations evaluated left to right as written;
 *   - mat * vec sums columns 0..3 in order;
 *   - dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z;
 *   - inverseSqrt(x) = 1 / sqrt(x), both correctly rounded;
 *     normalize(v) = v * inverseSqrt(dot(v, v)); length(v) = sqrt(dot(v, v));
 *   - mix(a, b, t) = a*(1-t) + b*t (the WGSL definition);
 *   - f32 -> u32 / i32 conversions truncate and saturate to the representable
 *     range ([0, 4294967040], [-2^31, 2147483520]) as naga's clamps do;
 *   - i32 add/sub/abs wrap (WGSL two's-complement semantics).
 * Where the reference's own CPU mirrors pin a function (gbuffer.rs, lod.rs,
 * mesh.rs) the tests check this file against their known answers.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "../include/hair_mi.h"

namespace {

template <class T> struct v2 { T x, y; };
template <class T> struct v3 { T x, y, z; };
template <class T> struct v4 { T x, y, z, w; };

template <class T> inline v2<T> operator+(v2<T> a, v2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline v2<T> operator-(v2<T> a, v2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline v2<T> operator*(v2<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline v3<T> operator*(v3<T> a, v3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline v3<T> operator/(v3<T> a, v3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline v4<T> operator+(v4<T> a, v4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline v4<T> operator*(v4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline v3<T> xyz(v4<T> a) { return {a.x, a.y, a.z}; }
template <class T> inline v3<T> splat3(T s) { return {s, s, s}; }

template <class T> inline T fmin_(T a, T b) { return b < a ? b : a; }
template <class T> inline T fmax_(T a, T b) { return a < b ? b : a; }
template <class T> inline T clampf(T x, T lo, T hi) { return fmin_(fmax_(x, lo), hi); }
template <class T> inline v3<T> vmin(v3<T> a, v3<T> b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
template <class T> inline v3<T> vmax(v3<T> a, v3<T> b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
template <class T> inline v3<T> vclamp(v3<T> a, T lo, T hi) { return {clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi)}; }

template <class T> inline T dot2(v2<T> a, v2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T dot3(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T len2(v2<T> a) { return std::sqrt(dot2(a, a)); }
template <class T> inline T len3(v3<T> a) { return std::sqrt(dot3(a, a)); }
template <class T> inline T inverse_sqrt(T x) { return T(1) / std::sqrt(x); }
template <class T> inline v3<T> normalize3(v3<T> v) { return v * inverse_sqrt(dot3(v, v)); }
template <class T> inline T mixs(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline v2<T> mix2(v2<T> a, v2<T> b, T t) { return {mixs(a.x, b.x, t), mixs(a.y, b.y, t)}; }
template <class T> inline v3<T> mix3(v3<T> a, v3<T> b, T t) { return {mixs(a.x, b.x, t), mixs(a.y, b.y, t), mixs(a.z, b.z, t)}; }

// WGSL value conversions (naga clamps before OpConvertFTo[SU]).
template <class T> inline uint32_t to_u32(T x) {
    if (!(x > T(0))) return 0u;
    if (x >= T(4294967040.0)) return 4294967040u;
    return (uint32_t)x;
}
template <class T> inline int32_t to_i32(T x) {
    if (x != x) return 0;
    if (x <= T(-2147483648.0)) return INT32_MIN;
    if (x >= T(2147483520.0)) return 2147483520;
    return (int32_t)x;
}
inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline int32_t wabs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }

constexpr float kPi = 3.14159265f;  // common.wgsl:5

template <class T>
struct Params {
    T clip[16], wfc[16], wfm[16], lclip[16];
    v4<T> camera_pos, camera_forward, screen, appearance, filter_params, albedo;
    uint32_t flags[4], pass_mode[4];
    v4<T> light_eye, light_forward, dom_info;
};

template <class T> inline v4<T> ld4(const float* f) { return {T(f[0]), T(f[1]), T(f[2]), T(f[3])}; }

template <class T> Params<T> load_params(const hm_params& p) {
    Params<T> q;
    for (int i = 0; i < 16; ++i) {
        q.clip[i] = T(p.clip_from_world[i]);
        q.wfc[i] = T(p.world_from_clip[i]);
        q.wfm[i] = T(p.world_from_model[i]);
        q.lclip[i] = T(p.light_clip_from_world[i]);
    }
    q.camera_pos = ld4<T>(p.camera_pos);
    q.camera_forward = ld4<T>(p.camera_forward);
    q.screen = ld4<T>(p.screen);
    q.appearance = ld4<T>(p.appearance);
    q.filter_params = ld4<T>(p.filter_params);
    q.albedo = ld4<T>(p.albedo);
    for (int i = 0; i < 4; ++i) { q.flags[i] = p.flags[i]; q.pass_mode[i] = p.pass_mode[i]; }
    q.light_eye = ld4<T>(p.light_eye);
    q.light_forward = ld4<T>(p.light_forward);
    q.dom_info = ld4<T>(p.dom_info);
    return q;
}

template <class T> inline v4<T> mat_mul(const T* m, v4<T> v) {
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w,
            m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
            m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w,
            m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w};
}

// ---- common.wgsl ---------------------------------------------------------

inline uint32_t pcg(uint32_t input) {  // common.wgsl:56-60
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

template <class T> inline v2<T> pcg_uv(uint32_t seed) {  // common.wgsl:62-66
    uint32_t a = pcg(seed);
    uint32_t b = pcg(a ^ 0x9e3779b9u);
    const T s = T(1.0f / 4294967296.0f);
    return {T((float)a) * s, T((float)b) * s};
}

inline uint64_t pack_gb(uint32_t depth, uint32_t tx, uint32_t ty, uint32_t u, uint32_t v, uint32_t w, uint32_t ao) {
    // common.wgsl:68-76
    uint64_t high = (uint64_t)((depth << 8u) | (tx & 0xffu));
    uint64_t low = (uint64_t)(((ty & 0xffu) << 24u) | ((u & 0x3fu) << 18u) | ((v & 0x3fu) << 12u) |
                              ((w & 0x3fu) << 6u) | (ao & 0x3fu));
    return (high << 32u) | low;
}

struct Unpacked { uint32_t depth, tx, ty, u, v, w, ao; };

inline Unpacked unpack_gb(uint64_t word) {  // common.wgsl:94-104
    uint32_t low = (uint32_t)(word & 0xffffffffull);
    uint32_t high = (uint32_t)(word >> 32u);
    return {high >> 8u, high & 0xffu, low >> 24u, (low >> 18u) & 0x3fu, (low >> 12u) & 0x3fu, (low >> 6u) & 0x3fu,
            low & 0x3fu};
}

inline bool is_empty(uint64_t word) { return ((uint32_t)(word >> 32u) >> 8u) == 0x00ffffffu; }  // common.wgsl:106-108

template <class T> inline uint32_t quantize_depth(T dist, T near_, T far_) {  // common.wgsl:110-113
    T q = clampf((dist - near_) / fmax_(far_ - near_, T(1e-3f)), T(0), T(1));
    return std::min(to_u32(q * T(16777214.0f)), 0x00fffffeu);
}

template <class T> inline T dequantize_depth(uint32_t depth, T near_, T far_) {  // common.wgsl:115-117
    return mixs(near_, far_, T((float)depth) / T(16777214.0f));
}

template <class T> inline T sel_sign(bool positive) { return positive ? T(1) : T(-1); }

template <class T> inline void oct_encode(v3<T> n_in, uint32_t& ox, uint32_t& oy) {  // common.wgsl:119-130
    v3<T> n = normalize3(n_in);
    T denom = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
    T m = fmax_(denom, T(1e-6f));
    v2<T> p = {n.x / m, n.y / m};
    if (n.z < T(0)) {
        v2<T> old = p;
        p.x = (T(1) - std::fabs(old.y)) * sel_sign<T>(old.x >= T(0));
        p.y = (T(1) - std::fabs(old.x)) * sel_sign<T>(old.y >= T(0));
    }
    T ex = clampf(p.x * T(0.5f) + T(0.5f), T(0), T(1));
    T ey = clampf(p.y * T(0.5f) + T(0.5f), T(0), T(1));
    ox = to_u32(ex * T(255.0f) + T(0.5f));
    oy = to_u32(ey * T(255.0f) + T(0.5f));
}

template <class T> inline v3<T> oct_decode(uint32_t tx, uint32_t ty) {  // common.wgsl:132-142
    T ex = T((float)tx) / T(255.0f);
    T ey = T((float)ty) / T(255.0f);
    T fx = ex * T(2) - T(1);
    T fy = ey * T(2) - T(1);
    v3<T> n = {fx, fy, T(1) - std::fabs(fx) - std::fabs(fy)};
    if (n.z < T(0)) {
        T ox = n.x, oy = n.y;
        n.x = (T(1) - std::fabs(oy)) * sel_sign<T>(ox >= T(0));
        n.y = (T(1) - std::fabs(ox)) * sel_sign<T>(oy >= T(0));
    }
    return normalize3(n);
}

template <class T> inline v3<T> safe_normalize(v3<T> v) {  // common.wgsl:144-150
    T l2 = dot3(v, v);
    if (l2 < T(1e-10f)) return {T(0), T(1), T(0)};
    return v * inverse_sqrt(l2);
}

template <class T> inline v3<T> calc_tangent(uint32_t idx, uint32_t count, v3<T> prev, v3<T> curr, v3<T> next) {
    // common.wgsl:152-166
    v3<T> dir;
    T scale;
    if (idx == 0u) {
        dir = safe_normalize(next - curr);
        scale = len3(next - curr);
    } else if (idx + 1u >= count) {
        dir = safe_normalize(curr - prev);
        scale = len3(curr - prev);
    } else {
        dir = safe_normalize(next - prev);
        scale = len3(next - prev) * T(0.5f);
    }
    return dir * fmax_(scale, T(1e-4f));
}

template <class T> inline v3<T> hermite(v3<T> p0, v3<T> p1, v3<T> m0, v3<T> m1, T t) {  // common.wgsl:168-176
    T t2 = t * t;
    T t3 = t2 * t;
    T h00 = T(2) * t3 - T(3) * t2 + T(1);
    T h10 = t3 - T(2) * t2 + t;
    T h01 = T(-2) * t3 + T(3) * t2;
    T h11 = t3 - t2;
    return h00 * p0 + h10 * m0 + h01 * p1 + h11 * m1;
}

template <class T> inline v4<T> bilinear4(v4<T> c00, v4<T> c10, v4<T> c01, v4<T> c11, v2<T> uv) {
    // common.wgsl:178-182
    T u = uv.x, v = uv.y;
    return c00 * ((T(1) - u) * (T(1) - v)) + c10 * (u * (T(1) - v)) + c01 * ((T(1) - u) * v) + c11 * (u * v);
}

template <class T> inline void quant_uvw(v3<T> uvw, uint32_t q[3]) {  // common.wgsl:184-191
    v3<T> c = vclamp(uvw, T(0), T(1));
    q[0] = std::min(to_u32(c.x * T(63.0f) + T(0.5f)), 63u);
    q[1] = std::min(to_u32(c.y * T(63.0f) + T(0.5f)), 63u);
    q[2] = std::min(to_u32(c.z * T(63.0f) + T(0.5f)), 63u);
}

template <class T> inline v3<T> to_world(const Params<T>& p, v3<T> q) {  // common.wgsl:197-199
    return xyz(mat_mul(p.wfm, v4<T>{q.x, q.y, q.z, T(1)}));
}

template <class T> inline v3<T> dir_world(const Params<T>& p, v3<T> d) {  // common.wgsl:201-203
    return safe_normalize(xyz(mat_mul(p.wfm, v4<T>{d.x, d.y, d.z, T(0)})));
}

template <class T> struct Projected { v2<T> screen; T view_dist; T clip_depth; bool in_front; };

template <class T> inline Projected<T> project_world(const Params<T>& p, v3<T> world) {  // common.wgsl:212-219
    v4<T> clip = mat_mul(p.clip, v4<T>{world.x, world.y, world.z, T(1)});
    T w = fmax_(clip.w, T(1e-5f));
    T ndc_x = clip.x / w, ndc_y = clip.y / w;
    T sx = (ndc_x * T(0.5f) + T(0.5f)) * p.screen.x;
    T sy = (T(1) - (ndc_y * T(0.5f) + T(0.5f))) * p.screen.y;
    T dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
    return {{sx, sy}, dist, clip.z / w, clip.w > T(0) && dist > p.screen.z * T(0.5f)};
}

template <class T> inline v3<T> reconstruct_world(const Params<T>& p, int32_t px, int32_t py, T view_dist) {
    // common.wgsl:221-229
    T ux = (T((float)px) + T(0.5f)) / fmax_(p.screen.x, T(1));
    T uy = (T((float)py) + T(0.5f)) / fmax_(p.screen.y, T(1));
    T ndc_x = ux * T(2) - T(1);
    T ndc_y = (T(1) - uy) * T(2) - T(1);
    v4<T> far_h = mat_mul(p.wfc, v4<T>{ndc_x, ndc_y, T(0), T(1)});
    T hw = fmax_(far_h.w, T(1e-5f));
    v3<T> far_p = {far_h.x / hw, far_h.y / hw, far_h.z / hw};
    v3<T> dir = far_p - xyz(p.camera_pos);
    T denom = fmax_(dot3(dir, xyz(p.camera_forward)), T(1e-4f));
    return xyz(p.camera_pos) + dir * (view_dist / denom);
}

template <class T> inline uint32_t control_points_for_lod(T l) {  // common.wgsl:231-238
    T c_raw = fmax_(std::floor(std::sqrt(clampf(l, T(0), T(1))) * T(127.0f)), T(16.0f));
    if (c_raw < T(3)) return to_u32(fmax_(c_raw, T(2)));
    // pow(2, floor(log2(c_raw - 1))) + 1 on an integer-valued c_raw in [16, 127]:
    // the exact value is the highest power of two not above c_raw - 1.
    uint32_t k = to_u32(c_raw - T(1));
    uint32_t snapped = (1u << (31 - __builtin_clz(k))) + 1u;
    return std::min(snapped, 127u);
}

inline uint32_t style_index(uint32_t offset, uint32_t u, uint32_t v, uint32_t w) {  // common.wgsl:240-242
    return offset + (w * HM_STYLE_V + v) * HM_STYLE_U + u;
}

// ---- lod.wgsl -------------------------------------------------------------

template <class T>
void lod_bundle(const Params<T>& p, const hm_bundle* bundles, const hm_layer* layers, uint32_t bundle_id,
                hm_lod_record* out, const hm_lod_override* ovr) {
    // lod.wgsl:6-59
    const hm_bundle& bundle = bundles[bundle_id];
    v2<T> aabb_min = {T(1e20f), T(1e20f)};
    v2<T> aabb_max = {T(-1e20f), T(-1e20f)};
    bool any_front = false, any_point = false;
    for (uint32_t layer = 0; layer < HM_LAYER_COUNT; ++layer) {
        for (uint32_t corner = 0; corner < 4; ++corner) {
            const hm_layer& vert = layers[bundle.layer_offset + layer * 4u + corner];
            v3<T> world = to_world(p, v3<T>{T(vert.pos_ao[0]), T(vert.pos_ao[1]), T(vert.pos_ao[2])});
            v4<T> clip = mat_mul(p.clip, v4<T>{world.x, world.y, world.z, T(1)});
            T dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
            if (clip.w > T(0) && dist > T(0)) any_front = true;
            Projected<T> pr = project_world(p, world);
            aabb_min = {fmin_(aabb_min.x, pr.screen.x), fmin_(aabb_min.y, pr.screen.y)};
            aabb_max = {fmax_(aabb_max.x, pr.screen.x), fmax_(aabb_max.y, pr.screen.y)};
            any_point = true;
        }
    }
    bool culled = !any_point || !any_front || aabb_max.x < T(0) || aabb_max.y < T(0) || aabb_min.x > p.screen.x ||
                  aabb_min.y > p.screen.y;
    uint32_t n_lod = 0, cps = HM_LAYER_COUNT;
    if (!culled) {
        if (p.pass_mode[3] == 1u) {
            n_lod = bundle.strand_count;
            cps = control_points_for_lod(T(1));
        } else if (p.flags[0] == 0u) {
            n_lod = bundle.strand_count;
            cps = 127u;
        } else {
            T diag = len2(aabb_max - aabb_min);
            T l = clampf(diag / fmax_(p.screen.y, T(1)) * p.appearance.z, T(0), T(1));
            T delta = T((float)pcg(bundle_id ^ 0x9e3779b9u)) * T(1.0f / 4294967296.0f);
            T n = T((float)bundle.strand_count);
            T raw = std::ceil(l * (n + delta));
            n_lod = to_u32(clampf(raw, T(1), n));
            cps = control_points_for_lod(l);
        }
        if (ovr && p.pass_mode[3] == 0u) {  // harness override (camera pass only)
            if (ovr->control_points) cps = std::min(std::max(ovr->control_points, 2u), 127u);
            if (ovr->strand_fraction > 0.0f) {
                float n = (float)bundle.strand_count;
                float raw = std::ceil(ovr->strand_fraction * n);
                n_lod = to_u32(clampf(raw, 1.0f, n));
            }
        }
    }
    out[bundle_id] = {n_lod, cps, bundle.strand_count, 0u};
}

// ---- raster.wgsl ------------------------------------------------------------

struct Targets {
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    uint32_t* dom;
    const float* scene_depth;  // W*H or null
    float scene_clear;
    uint64_t counts[4];
};

// Plain (single-thread) or atomic (multi-thread CPU baseline) updates; min and
// add are order independent, so both give the same bytes.
template <bool ATOMIC> struct Sink;
template <> struct Sink<false> {
    static void min64(uint64_t* p, uint64_t v) { if (v < *p) *p = v; }
    static void min32(uint32_t* p, uint32_t v) { if (v < *p) *p = v; }
    static void add32(uint32_t* p, uint32_t v) { *p += v; }
    static uint32_t load32(const uint32_t* p) { return *p; }
};
template <> struct Sink<true> {
    static void min64(uint64_t* p, uint64_t v) {
        uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
        while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
    }
    static void min32(uint32_t* p, uint32_t v) {
        uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
        while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
    }
    static void add32(uint32_t* p, uint32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
    static uint32_t load32(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
};

template <class T, bool ATOMIC> struct Raster {
    const Params<T>& p;
    const hm_bundle* bundles;
    const hm_layer* layers;
    const float* style;
    Targets& tg;
    uint64_t counts[4] = {0, 0, 0, 0};

    struct SmLayer { v4<T> pos_ao, tangent; };
    SmLayer sm[16];
    v4<T> cp_pos[128], cp_tan[128], cp_uvw[128];

    SmLayer sample_layer(uint32_t offset, uint32_t layer, v2<T> uv) const {  // raster.wgsl:22-32
        uint32_t i = offset + layer * 4u;
        auto P = [&](uint32_t k) { return ld4<T>(layers[k].pos_ao); };
        auto D = [&](uint32_t k) { return ld4<T>(layers[k].tangent); };
        return {bilinear4(P(i), P(i + 1), P(i + 2), P(i + 3), uv), bilinear4(D(i), D(i + 1), D(i + 2), D(i + 3), uv)};
    }

    v3<T> texel(uint32_t idx) const { return {T(style[idx * 4]), T(style[idx * 4 + 1]), T(style[idx * 4 + 2])}; }

    v3<T> sample_style(uint32_t offset, v2<T> uv, T w) const {  // raster.wgsl:34-60
        T u = clampf(uv.x, T(0), T(1)) * T((float)(HM_STYLE_U - 1u));
        T v = clampf(uv.y, T(0), T(1)) * T((float)(HM_STYLE_V - 1u));
        T ww = clampf(w, T(0), T(1)) * T((float)(HM_STYLE_W - 1u));
        uint32_t u0 = to_u32(std::floor(u)), v0 = to_u32(std::floor(v)), w0 = to_u32(std::floor(ww));
        uint32_t u1 = std::min(u0 + 1u, HM_STYLE_U - 1u);
        uint32_t v1 = std::min(v0 + 1u, HM_STYLE_V - 1u);
        uint32_t w1 = std::min(w0 + 1u, HM_STYLE_W - 1u);
        T tu = u - T((float)u0), tv = v - T((float)v0), tw = ww - T((float)w0);
        v3<T> c000 = texel(style_index(offset, u0, v0, w0));
        v3<T> c100 = texel(style_index(offset, u1, v0, w0));
        v3<T> c010 = texel(style_index(offset, u0, v1, w0));
        v3<T> c110 = texel(style_index(offset, u1, v1, w0));
        v3<T> c001 = texel(style_index(offset, u0, v0, w1));
        v3<T> c101 = texel(style_index(offset, u1, v0, w1));
        v3<T> c011 = texel(style_index(offset, u0, v1, w1));
        v3<T> c111 = texel(style_index(offset, u1, v1, w1));
        v3<T> c00 = mix3(c000, c100, tu);
        v3<T> c01 = mix3(c010, c110, tu);
        v3<T> c10 = mix3(c001, c101, tu);
        v3<T> c11 = mix3(c011, c111, tu);
        return mix3(mix3(c00, c01, tv), mix3(c10, c11, tv), tw);
    }

    v4<T> eval_cp(uint32_t idx, uint32_t ncp, v2<T> uv, uint32_t style_offset) {  // raster.wgsl:62-75
        T w = T((float)idx) / T((float)std::max(ncp - 1u, 1u));
        T layers_f = w * T((float)(HM_LAYER_COUNT - 1u));
        uint32_t i0 = std::min(to_u32(std::floor(layers_f)), HM_LAYER_COUNT - 1u);
        uint32_t i1 = std::min(i0 + 1u, HM_LAYER_COUNT - 1u);
        T t = layers_f - T((float)i0);
        const SmLayer& a = sm[i0];
        const SmLayer& b = sm[i1];
        v3<T> base = hermite(xyz(a.pos_ao), xyz(b.pos_ao), xyz(a.tangent), xyz(b.tangent), t);
        T ao = mixs(a.pos_ao.w, b.pos_ao.w, t);
        v3<T> residual = sample_style(style_offset, uv, w);
        cp_uvw[idx] = {uv.x, uv.y, w, ao};
        v3<T> q = base + residual;
        return {q.x, q.y, q.z, ao};
    }

    T scene_at(int32_t x, int32_t y) const {
        if (!tg.scene_depth) return T(tg.scene_clear);
        return T(tg.scene_depth[(size_t)y * (size_t)to_i32(p.screen.x) + (size_t)x]);
    }

    void write_pixel(int32_t px, int32_t py, T t, v3<T> p0, v3<T> p1, v3<T> t0, v3<T> t1, v3<T> uvw0, v3<T> uvw1,
                     T ao0, T ao1, bool center_hit, T beta) {  // raster.wgsl:77-116
        int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        if (px < 0 || py < 0 || px >= width || py >= height) return;
        v3<T> world = mix3(p0, p1, t);
        v3<T> tangent = safe_normalize(mix3(t0, t1, t));
        v3<T> uvw = mix3(uvw0, uvw1, t);
        T ao = mixs(ao0, ao1, t);
        T dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
        if (dist <= p.screen.z) return;
        v4<T> clip = mat_mul(p.clip, v4<T>{world.x, world.y, world.z, T(1)});
        T clip_depth = clip.z / fmax_(clip.w, T(1e-5f));
        T scene = scene_at(px, py);
        if (p.pass_mode[2] == 1u) {
            if (clip_depth < scene) return;
        } else if (clip_depth > scene) {
            return;
        }
        uint32_t depth = quantize_depth(dist, p.screen.z, p.screen.w);
        uint32_t ox, oy;
        oct_encode(tangent, ox, oy);
        uint32_t ao_q = to_u32(clampf(ao, T(0), T(1)) * T(63.0f) + T(0.5f));
        if (p.flags[2] == 0u) ao_q = 63u;
        uint32_t q[3];
        quant_uvw(uvw, q);
        uint64_t payload = pack_gb(depth, ox, oy, q[0], q[1], q[2], ao_q);
        size_t pix = (size_t)(uint32_t)py * (uint32_t)width + (uint32_t)px;
        counts[0]++;
        if (tg.cons) Sink<ATOMIC>::min64(&tg.cons[pix], payload);
        if (center_hit) {
            counts[1]++;
            uint32_t beta_q = to_u32(clampf(beta, T(0), T(1)) * T(255.0f) + T(0.5f));
            if (tg.center) Sink<ATOMIC>::min64(&tg.center[pix], payload);
            if (tg.beta) Sink<ATOMIC>::min32(&tg.beta[pix], (depth << 8u) | beta_q);
        }
    }

    void write_dom_pixel(int32_t px, int32_t py, v3<T> world) {  // raster.wgsl:118-148
        int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        if (px < 0 || py < 0 || px >= width || py >= height) return;
        T near_ = p.screen.z, far_ = p.screen.w;
        T dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
        if (dist <= near_ || dist >= far_) return;
        uint32_t dom_w = to_u32(p.dom_info.x);
        size_t pix = (size_t)(uint32_t)py * dom_w + (uint32_t)px;
        uint32_t q = quantize_depth(dist, near_, far_);
        if (p.pass_mode[0] == 1u) {
            counts[2]++;
            if (tg.dom) Sink<ATOMIC>::min32(&tg.dom[pix], q);
            return;
        }
        uint32_t front_q = tg.dom ? Sink<ATOMIC>::load32(&tg.dom[pix]) : 0u;
        if (front_q == HM_DOM_EMPTY) return;
        T z_front = dequantize_depth(front_q, near_, far_);
        T thickness = fmax_(far_ - near_, T(1e-3f)) / T((float)HM_DOM_LAYERS);
        T t = fmax_(dist - z_front, T(0)) / thickness;
        uint32_t slice = std::min(to_u32(t), HM_DOM_LAYERS - 1u);
        size_t pixels = (size_t)dom_w * to_u32(p.dom_info.y);
        counts[2]++;
        if (tg.dom) Sink<ATOMIC>::add32(&tg.dom[pixels + pix * HM_DOM_LAYERS + slice], 1u);
    }

    void draw_dom_segment(v3<T> p0, v3<T> p1, bool include_start) {  // raster.wgsl:150-190
        Projected<T> a = project_world(p, p0);
        Projected<T> b = project_world(p, p1);
        if (!a.in_front && !b.in_front) return;
        counts[3]++;
        int32_t x0 = to_i32(std::floor(a.screen.x));
        int32_t y0 = to_i32(std::floor(a.screen.y));
        int32_t x1 = to_i32(std::floor(b.screen.x));
        int32_t y1 = to_i32(std::floor(b.screen.y));
        int32_t dx = wabs(wsub(x1, x0));
        int32_t dy = wabs(wsub(y1, y0));
        int32_t sx = x0 < x1 ? 1 : -1;
        int32_t sy = y0 < y1 ? 1 : -1;
        int32_t err = wsub(dx, dy);
        v2<T> seg = b.screen - a.screen;
        T seg_len2 = fmax_(dot2(seg, seg), T(1e-6f));
        bool skip = !include_start && (x0 != x1 || y0 != y1);
        for (int step = 0; step < 1024; ++step) {
            if (!skip) {
                v2<T> center = {T((float)x0) + T(0.5f), T((float)y0) + T(0.5f)};
                T t = clampf(dot2(center - a.screen, seg) / seg_len2, T(0), T(1));
                write_dom_pixel(x0, y0, mix3(p0, p1, t));
            }
            skip = false;
            if (x0 == x1 && y0 == y1) break;
            int32_t e2 = wadd(err, err);
            if (e2 > wsub(0, dy)) {
                err = wsub(err, dy);
                x0 = wadd(x0, sx);
            }
            if (e2 < dx) {
                err = wadd(err, dx);
                y0 = wadd(y0, sy);
            }
        }
    }

    void draw_segment(v3<T> p0, v3<T> p1, v3<T> t0, v3<T> t1, v3<T> uvw0, v3<T> uvw1, T ao0, T ao1, T beta) {
        // raster.wgsl:192-223
        Projected<T> a = project_world(p, p0);
        Projected<T> b = project_world(p, p1);
        if (!a.in_front && !b.in_front) return;
        counts[3]++;
        v2<T> delta = b.screen - a.screen;
        int32_t steps = to_i32(std::ceil(fmax_(std::fabs(delta.x), std::fabs(delta.y))));
        int32_t n = std::min(std::max(steps, 1), 96);
        T diameter = fmax_(p.appearance.w, T(0.05f));
        v2<T> prev = a.screen;
        for (int32_t i = 0; i <= n; ++i) {
            T t = T((float)i) / T((float)n);
            v2<T> screen = mix2(a.screen, b.screen, t);
            int32_t pxi = to_i32(std::floor(screen.x)), pyi = to_i32(std::floor(screen.y));
            v2<T> center = {T((float)pxi) + T(0.5f), T((float)pyi) + T(0.5f)};
            v2<T> seg = b.screen - a.screen;
            T denom = fmax_(dot2(seg, seg), T(1e-6f));
            T closest = clampf(dot2(center - a.screen, seg) / denom, T(0), T(1));
            T dist = len2(center - (a.screen + seg * closest));
            bool center_hit = dist <= diameter * T(0.5f);
            write_pixel(pxi, pyi, t, p0, p1, t0, t1, uvw0, uvw1, ao0, ao1, center_hit, beta);
            if (i > 0) {
                int32_t ppx = to_i32(std::floor(prev.x)), ppy = to_i32(std::floor(prev.y));
                if (ppx != pxi && ppy != pyi) {
                    write_pixel(ppx, pyi, t, p0, p1, t0, t1, uvw0, uvw1, ao0, ao1, false, beta);
                    write_pixel(pxi, ppy, t, p0, p1, t0, t1, uvw0, uvw1, ao0, ao1, false, beta);
                }
            }
            prev = screen;
        }
    }

    void strand(const hm_strand_ref& ref) {  // raster.wgsl:225-337, one workgroup
        const hm_bundle& bundle = bundles[ref.bundle];
        v2<T> uv = pcg_uv<T>(ref.strand ^ (ref.bundle * 0x85ebca6bu));
        v4<T> loaded[16];
        for (uint32_t lane = 0; lane < HM_LAYER_COUNT; ++lane) {
            SmLayer s = sample_layer(bundle.layer_offset, lane, uv);
            loaded[lane] = s.pos_ao;
            sm[lane] = s;
        }
        // The LDS path (subgroup size != 32); the shuffle path reads the same values.
        SmLayer next_sm[16];
        for (uint32_t lane = 0; lane < HM_LAYER_COUNT; ++lane) {
            uint32_t prev_i = lane == 0u ? 0u : lane - 1u;
            uint32_t next_i = std::min(lane + 1u, HM_LAYER_COUNT - 1u);
            v3<T> prev_p = xyz(sm[prev_i].pos_ao);
            v3<T> next_p = xyz(sm[next_i].pos_ao);
            if (lane == 0u) prev_p = xyz(loaded[lane]);
            if (lane + 1u >= HM_LAYER_COUNT) next_p = xyz(loaded[lane]);
            v3<T> tangent = calc_tangent(lane, HM_LAYER_COUNT, prev_p, xyz(loaded[lane]), next_p);
            next_sm[lane] = {loaded[lane], {tangent.x, tangent.y, tangent.z, T(0)}};
        }
        for (uint32_t lane = 0; lane < HM_LAYER_COUNT; ++lane) sm[lane] = next_sm[lane];

        uint32_t ncp = std::min(std::max(ref.control_points, 2u), 127u);
        for (uint32_t idx = 0; idx < ncp; ++idx) cp_pos[idx] = eval_cp(idx, ncp, uv, bundle.style_offset);
        for (uint32_t idx = 0; idx < ncp; ++idx) {
            v3<T> cur = xyz(cp_pos[idx]);
            v3<T> prev_p = idx > 0u ? xyz(cp_pos[idx - 1u]) : cur;
            v3<T> next_p = idx + 1u < ncp ? xyz(cp_pos[idx + 1u]) : cur;
            v3<T> tn = safe_normalize(calc_tangent(idx, ncp, prev_p, cur, next_p));
            cp_tan[idx] = {tn.x, tn.y, tn.z, T(0)};
        }
        T beta = T((float)ref.n_lod) / T((float)std::max(bundle.strand_count, 1u));
        for (uint32_t idx = 0; idx + 1u < ncp; ++idx) {
            v3<T> p0 = to_world(p, xyz(cp_pos[idx]));
            v3<T> p1 = to_world(p, xyz(cp_pos[idx + 1u]));
            if (p.pass_mode[0] != 0u) {
                draw_dom_segment(p0, p1, idx == 0u);
            } else {
                v3<T> t0 = dir_world(p, xyz(cp_tan[idx]));
                v3<T> t1 = dir_world(p, xyz(cp_tan[idx + 1u]));
                draw_segment(p0, p1, t0, t1, xyz(cp_uvw[idx]), xyz(cp_uvw[idx + 1u]), cp_uvw[idx].w,
                             cp_uvw[idx + 1u].w, beta);
            }
        }
    }
};

// ---- shade.wesl -----------------------------------------------------------

template <class T> struct Lobes { v3<T> r, tt, trt; };
template <class T> struct Shape { v3<T> alpha, beta; };
template <class T> struct DualScatter {
    Lobes<T> fwd;
    v3<T> a_f, alpha_f, beta_f2, a_back, delta_b, sigma_b2;
};

template <class T> inline v3<T> vexp(v3<T> a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }
template <class T> inline v3<T> vsqrt(v3<T> a) { return {std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z)}; }

template <class T> inline v3<T> gaussian3(v3<T> b, v3<T> x) {  // shade.wesl:51-54
    v3<T> bb = vmax(b, splat3(T(1e-3f)));
    v3<T> num = vexp(T(-0.5f) * x * x / (bb * bb));
    return num / (T(std::sqrt(T(2) * T(kPi))) * bb);
}

template <class T> inline T fresnel(T cos_t) {  // shade.wesl:56-60
    const T ior = T(1.55f);
    T r0 = (T(1) - ior) / (T(1) + ior);
    T f0 = r0 * r0;
    return f0 + (T(1) - f0) * std::pow(T(1) - clampf(cos_t, T(0), T(1)), T(5));
}

template <class T> inline v3<T> absorption_from_color(v3<T> color, T roughness) {  // shade.wesl:70-77
    T b = clampf(roughness, T(0), T(1));
    T b2 = b * b;
    T b3 = b2 * b;
    T d = T(5.969f) - T(0.215f) * b + T(2.532f) * b2 - T(10.73f) * b3 + T(5.574f) * b2 * b2 + T(0.245f) * b3 * b2;
    v3<T> c = vclamp(color, T(1e-4f), T(1));
    v3<T> l = {std::log(c.x) / d, std::log(c.y) / d, std::log(c.z) / d};
    return l * l;
}

template <class T> inline Lobes<T> azimuthal(T cos_theta_d, T cos_phi, v3<T> sigma_a) {  // shade.wesl:94-116
    T cos_half_phi = std::sqrt(clampf(T(0.5f) + T(0.5f) * cos_phi, T(0), T(1)));
    T n_r = T(0.25f) * cos_half_phi * fresnel(cos_theta_d * cos_half_phi);
    T eta = T(1.19f) / cos_theta_d + T(0.36f) * cos_theta_d;
    T a = T(1) / eta;
    T h = cos_half_phi * (T(1) + a * (T(0.6f) - T(0.8f) * cos_phi));
    T f_tt = fresnel(cos_theta_d * std::sqrt(clampf(T(1) - h * h, T(0), T(1))));
    T s_tt = std::sqrt(clampf(T(1) - h * h * a * a, T(0), T(1)));
    v3<T> t_tt = vexp(sigma_a * T(-1) * T(2) * s_tt / cos_theta_d);
    T k_tt = std::exp(T(-3.65f) * cos_phi - T(3.98f)) * (T(1) - f_tt) * (T(1) - f_tt);
    v3<T> n_tt = t_tt * k_tt;
    T f_trt = fresnel(cos_theta_d * T(0.5f));
    v3<T> t_trt = vexp(sigma_a * T(-1) * T(3.2f) / cos_theta_d);
    T k_trt = std::exp(T(17.0f) * cos_phi - T(16.78f)) * (T(1) - f_trt) * (T(1) - f_trt) * f_trt;
    v3<T> n_trt = t_trt * k_trt;
    return {splat3(n_r), n_tt, n_trt};
}

template <class T> inline Shape<T> lobe_shape(T roughness, T tilt) {  // shade.wesl:124-128
    T b = clampf(roughness * roughness, T(0.02f), T(1));
    return {{tilt, T(-0.5f) * tilt, T(-1.5f) * tilt}, {b, T(0.5f) * b, T(2) * b}};
}

template <class T> inline v3<T> lobe_mean(const Lobes<T>& l, v3<T> total, v3<T> q) {  // shade.wesl:148-150
    v3<T> num = l.r * q.x + l.tt * q.y + l.trt * q.z;
    return num / vmax(total, splat3(T(1e-5f)));
}

template <class T> struct ScatterSums { Lobes<T> fwd, back; };

// The midpoint-rule azimuthal integrals of dual_scatter (shade.wesl:153-165).
template <class T> ScatterSums<T> scatter_sums(T cos_theta_d, v3<T> sigma_a) {
    Lobes<T> fwd = {splat3(T(0)), splat3(T(0)), splat3(T(0))};
    Lobes<T> back = fwd;
    const uint32_t steps = 8u;
    for (uint32_t i = 0; i < steps; ++i) {
        T u = (T((float)i) + T(0.5f)) / T((float)steps);
        Lobes<T> b = azimuthal(cos_theta_d, std::cos(u * T(0.5f) * T(kPi)), sigma_a);
        Lobes<T> f = azimuthal(cos_theta_d, std::cos((T(1) + u) * T(0.5f) * T(kPi)), sigma_a);
        back = {back.r + b.r, back.tt + b.tt, back.trt + b.trt};
        fwd = {fwd.r + f.r, fwd.tt + f.tt, fwd.trt + f.trt};
    }
    T w = T(kPi) / T((float)steps);
    fwd = {fwd.r * w, fwd.tt * w, fwd.trt * w};
    back = {back.r * w, back.tt * w, back.trt * w};
    return {fwd, back};
}

// The dual-scattering terms from the lobe sums (shade.wesl:167-189, Zinke
// et al. 2008 Eqs. 13-17 as the reference writes them).
template <class T> DualScatter<T> dual_terms(const Lobes<T>& fwd, const Lobes<T>& back, const Shape<T>& shape) {
    v3<T> a_f = vmin(fwd.r + fwd.tt + fwd.trt, splat3(T(0.95f)));
    v3<T> a_b = back.r + back.tt + back.trt;
    v3<T> beta2 = shape.beta * shape.beta;
    v3<T> alpha_f = lobe_mean(fwd, a_f, shape.alpha);
    v3<T> beta_f2 = lobe_mean(fwd, a_f, beta2);
    v3<T> alpha_b = lobe_mean(back, a_b, shape.alpha);
    v3<T> beta_b2 = lobe_mean(back, a_b, beta2);
    v3<T> af2 = a_f * a_f;
    v3<T> rest = splat3(T(1)) - af2;
    v3<T> ab2 = a_b * a_b;
    v3<T> ab3 = ab2 * a_b;
    v3<T> a_back = a_b * af2 / rest + ab3 * af2 / (rest * rest * rest);
    v3<T> delta_b = alpha_b * (splat3(T(1)) - T(2) * ab2 / (rest * rest)) +
                    alpha_b * (T(2) * rest * rest + T(4) * af2 * ab2) / (rest * rest * rest);
    v3<T> beta_f = vsqrt(beta_f2);
    v3<T> beta_b = vsqrt(beta_b2);
    v3<T> sigma_b = (splat3(T(1)) + T(0.7f) * af2) *
                    (a_b * vsqrt(T(2) * beta_f2 + beta_b2) + ab3 * vsqrt(T(2) * beta_f2 + T(3) * beta_b2)) /
                    vmax(a_b + ab3 * (T(2) * beta_f + T(3) * beta_b), splat3(T(1e-5f)));
    return {fwd, a_f, alpha_f, beta_f2, a_back, delta_b, sigma_b * sigma_b};
}

template <class T> DualScatter<T> dual_scatter(T cos_theta_d, v3<T> sigma_a, const Shape<T>& shape) {
    // shade.wesl:153-190
    ScatterSums<T> s = scatter_sums(cos_theta_d, sigma_a);
    return dual_terms(s.fwd, s.back, shape);
}

// hair_bsdf's terms: single scattering, the local back lobe, the global
// forward-scattering term (zero unless depth > 1e-3) and exp(-depth).
template <class T> struct BsdfParts { v3<T> single, back, global; T direct; };

template <class T>
BsdfParts<T> hair_bsdf_parts(v3<T> tangent, v3<T> view, v3<T> light, v3<T> sigma_a, const Shape<T>& shape,
                             const DualScatter<T>& ds, T depth) {  // shade.wesl:193-242
    v3<T> t = safe_normalize(tangent);
    v3<T> v = safe_normalize(view);
    v3<T> l = safe_normalize(light);
    T sin_i = clampf(dot3(t, l), T(-1), T(1));
    T sin_o = clampf(dot3(t, v), T(-1), T(1));
    T theta_i = std::asin(sin_i);
    T theta_o = std::asin(sin_o);
    T theta_h = theta_i + theta_o;
    T cos_theta_d = fmax_(std::cos(T(0.5f) * std::fabs(theta_i - theta_o)), T(1e-3f));
    v3<T> t_i = l - t * sin_i;
    v3<T> t_o = v - t * sin_o;
    T proj = std::sqrt(fmax_(dot3(t_i, t_i) * dot3(t_o, t_o), T(1e-8f)));
    T cos_phi = clampf(dot3(t_i, t_o) / proj, T(-1), T(1));
    Lobes<T> n = azimuthal(cos_theta_d, cos_phi, sigma_a);
    v3<T> m = gaussian3(shape.beta, splat3(theta_h) - shape.alpha);
    v3<T> single = m.x * n.r + m.y * n.tt + m.z * n.trt;
    v3<T> back = T(2) * ds.a_back * gaussian3(vsqrt(ds.sigma_b2), splat3(theta_h) - ds.delta_b) / T(kPi);
    T direct = std::exp(-depth);
    BsdfParts<T> out = {single, back, splat3(T(0)), direct};
    if (depth > T(1e-3f)) {
        v3<T> sigma_f2 = depth * ds.beta_f2;
        v3<T> transmit = {std::pow(ds.a_f.x, depth), std::pow(ds.a_f.y, depth), std::pow(ds.a_f.z, depth)};
        v3<T> beta2 = shape.beta * shape.beta;
        v3<T> x = splat3(theta_h);
        v3<T> spread = gaussian3(vsqrt(splat3(beta2.x) + sigma_f2), x - splat3(shape.alpha.x)) * ds.fwd.r +
                       gaussian3(vsqrt(splat3(beta2.y) + sigma_f2), x - splat3(shape.alpha.y)) * ds.fwd.tt +
                       gaussian3(vsqrt(splat3(beta2.z) + sigma_f2), x - splat3(shape.alpha.z)) * ds.fwd.trt;
        v3<T> back_spread = T(2) * ds.a_back * gaussian3(vsqrt(ds.sigma_b2 + sigma_f2), x - ds.delta_b) / T(kPi);
        out.global = (T(1) - direct) * transmit * T(0.7f) * (spread / T(kPi) + T(kPi) * T(0.7f) * back_spread);
    }
    return out;
}

template <class T>
v3<T> hair_bsdf(v3<T> tangent, v3<T> view, v3<T> light, v3<T> sigma_a, const Shape<T>& shape,
                const DualScatter<T>& ds, T shadow, T depth) {  // shade.wesl:193-244
    BsdfParts<T> b = hair_bsdf_parts(tangent, view, light, sigma_a, shape, ds, depth);
    v3<T> color = b.direct * (b.single + T(0.7f) * b.back);
    if (depth > T(1e-3f)) color = color + b.global;
    return color * shadow;
}

template <class T> struct DomView {
    const Params<T>& p;
    const uint32_t* dom;

    T optical_depth(uint32_t pix, T t) const {  // shade.wesl:246-260
        size_t pixels = (size_t)to_u32(p.dom_info.x) * to_u32(p.dom_info.y);
        size_t base = pixels + (size_t)pix * HM_DOM_LAYERS;
        T occ = T(0);
        for (uint32_t slice = 0; slice < HM_DOM_LAYERS; ++slice) {
            T count = T((float)dom[base + slice]);
            T i = T((float)slice);
            if (i + T(1) <= t) occ = occ + count;
            else if (i < t) occ = occ + count * (t - i);
        }
        return occ;
    }

    T texel_visibility(int32_t x, int32_t y, T dist, T near_, T far_, uint32_t dom_w, uint32_t dom_h) const {
        // shade.wesl:262-277
        if (x < 0 || y < 0 || x >= (int32_t)dom_w || y >= (int32_t)dom_h) return T(1);
        uint32_t pix = (uint32_t)y * dom_w + (uint32_t)x;
        uint32_t front_q = dom[pix];
        if (front_q == HM_DOM_EMPTY) return T(1);
        T z_front = dequantize_depth(front_q, near_, far_);
        T thickness = fmax_(far_ - near_, T(1e-3f)) / T((float)HM_DOM_LAYERS);
        T rel = fmax_(dist - z_front - thickness * T(0.05f), T(0));
        return std::exp(-optical_depth(pix, rel / thickness) * p.filter_params.w);
    }

    T visibility(v3<T> world, T beta) const {  // shade.wesl:279-310
        if (p.flags[3] == 0u) return T(1);
        v4<T> clip = mat_mul(p.lclip, v4<T>{world.x, world.y, world.z, T(1)});
        if (clip.w <= T(0)) return T(1);
        T ndc_x = clip.x / clip.w, ndc_y = clip.y / clip.w;
        T ux = ndc_x * T(0.5f) + T(0.5f);
        T uy = T(1) - (ndc_y * T(0.5f) + T(0.5f));
        if (ux < T(0) || uy < T(0) || ux > T(1) || uy > T(1)) return T(1);
        uint32_t dom_w = to_u32(p.dom_info.x), dom_h = to_u32(p.dom_info.y);
        T near_ = p.dom_info.z, far_ = p.dom_info.w;
        T dist = dot3(world - xyz(p.light_eye), xyz(p.light_forward)) -
                 (-std::log(fmax_(beta, T(1e-4f))) * p.filter_params.z);
        T fx = ux * T((float)dom_w) - T(0.5f);
        T fy = uy * T((float)dom_h) - T(0.5f);
        int32_t x0 = to_i32(std::floor(fx)), y0 = to_i32(std::floor(fy));
        T tx = fx - std::floor(fx), ty = fy - std::floor(fy);
        T v00 = texel_visibility(x0, y0, dist, near_, far_, dom_w, dom_h);
        T v10 = texel_visibility(x0 + 1, y0, dist, near_, far_, dom_w, dom_h);
        T v01 = texel_visibility(x0, y0 + 1, dist, near_, far_, dom_w, dom_h);
        T v11 = texel_visibility(x0 + 1, y0 + 1, dist, near_, far_, dom_w, dom_h);
        return mixs(mixs(v00, v10, tx), mixs(v01, v11, tx), ty);
    }
};

// Bevy's smooth range window (bevy_pbr 0.20.1 pbr_lighting.wgsl
// getRangeFalloff): saturate(1 - (d^2 / r^2)^2)^2, zero from the range on.
template <class T> inline T range_falloff(T dist2, T inv_range2) {
    T factor = dist2 * inv_range2;
    T smooth = clampf(T(1) - factor * factor, T(0), T(1));
    return smooth * smooth;
}

// getDistanceAttenuation: the range window over the squared distance.
template <class T> inline T distance_attenuation(T dist2, T inv_range2) {
    T att = range_falloff(dist2, inv_range2);
    return att * T(1) / fmax_(dist2, T(1e-4f));
}

template <class T> inline T spot_mask(const hm_point_light& L, v3<T> world) {  // shade.wesl:312-327
    v3<T> spot_dir = {T(L.light_custom_data[0]), T(0), T(L.light_custom_data[1])};
    spot_dir.y = std::sqrt(fmax_(T(0), T(1) - spot_dir.x * spot_dir.x - spot_dir.z * spot_dir.z));
    if (L.flags & HM_LIGHT_FLAG_SPOT_Y_NEGATIVE) spot_dir.y = -spot_dir.y;
    v3<T> light_to_frag = v3<T>{T(L.position_radius[0]), T(L.position_radius[1]), T(L.position_radius[2])} - world;
    T cd = dot3(spot_dir * T(-1), safe_normalize(light_to_frag));
    T att = clampf(cd * T(L.light_custom_data[2]) + T(L.light_custom_data[3]), T(0), T(1));
    return att * att;
}

struct Lights {
    const hm_point_light* points;
    uint32_t np;
    const hm_directional_light* dirs;
    uint32_t nd;
    const hm_rect_light* rects;
    uint32_t nr;
    const hm_env_probe* env;  // null: no environment map
};

template <class T> inline v3<T> cross3(v3<T> a, v3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---- the environment probe (shade.wesl:329-427) ------------------------------

template <class T> inline T luminance(v3<T> c) {  // shade.wesl:62-64
    return c.x * T(0.2126f) + c.y * T(0.7152f) + c.z * T(0.0722f);
}

template <class T> inline v3<T> scatter_tint(v3<T> albedo, T shadow) {  // shade.wesl:81-85
    v3<T> hue = albedo * (T(1) / fmax_(luminance(albedo), T(1e-4f)));
    T e = T(0.5f) - T(0.5f) * clampf(shadow, T(0), T(1));
    return {std::pow(hue.x, e), std::pow(hue.y, e), std::pow(hue.z, e)};
}

template <class T> inline v3<T> quat_rotate(const float q[4], v3<T> d) {  // shade.wesl:329-332
    v3<T> u = {T(q[0]), T(q[1]), T(q[2])};
    v3<T> t = T(2) * cross3(u, d);
    return d + T(q[3]) * t + cross3(u, t);
}

template <class T> inline v3<T> cubemap_dir(const hm_env_probe& e, v3<T> d) {  // shade.wesl:334-339
    v3<T> s = quat_rotate(e.view_rotation, d);
    s.z = -s.z;  // cube maps are left-handed
    return s;
}

// The cube face a direction selects and its face coordinates in [0, 1]
// (Vulkan's cube map face selection table; ties go to x, then y; a zero
// direction reads the middle of face +x).
template <class T> struct CubeUV { uint32_t face; T s, t; };

template <class T> CubeUV<T> cube_face(v3<T> d) {
    T ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    uint32_t face;
    T ma, sc, tc;
    if (ax >= ay && ax >= az) {
        ma = ax;
        face = d.x >= T(0) ? 0u : 1u;
        sc = d.x >= T(0) ? -d.z : d.z;
        tc = -d.y;
    } else if (ay >= az) {
        ma = ay;
        face = d.y >= T(0) ? 2u : 3u;
        sc = d.x;
        tc = d.y >= T(0) ? d.z : -d.z;
    } else {
        ma = az;
        face = d.z >= T(0) ? 4u : 5u;
        sc = d.z >= T(0) ? d.x : -d.x;
        tc = -d.y;
    }
    if (!(ma > T(0))) return {0u, T(0.5f), T(0.5f)};
    return {face, T(0.5f) * (sc / ma + T(1)), T(0.5f) * (tc / ma + T(1))};
}

// Bilinear filtering inside one face with clamp-to-edge (the cube map
// sampler's linear filter without filtering across faces), RGB of RGBA texels.
template <class T> v3<T> cube_bilinear(const float* level, uint32_t size, CubeUV<T> c) {
    const float* face = level + (size_t)c.face * size * size * 4u;
    T u = c.s * T((float)size) - T(0.5f), v = c.t * T((float)size) - T(0.5f);
    T fu = std::floor(u), fv = std::floor(v);
    T wx = u - fu, wy = v - fv;
    int32_t x0 = (int32_t)fu, y0 = (int32_t)fv;
    int32_t hi = (int32_t)size - 1;
    int32_t xs[2] = {std::min(std::max(x0, 0), hi), std::min(std::max(x0 + 1, 0), hi)};
    int32_t ys[2] = {std::min(std::max(y0, 0), hi), std::min(std::max(y0 + 1, 0), hi)};
    v3<T> tap[4];
    for (int k = 0; k < 4; ++k) {
        const float* q = face + ((size_t)ys[k >> 1] * size + (size_t)xs[k & 1]) * 4u;
        tap[k] = {T(q[0]), T(q[1]), T(q[2])};
    }
    return mix3(mix3(tap[0], tap[1], wx), mix3(tap[2], tap[3], wx), wy);
}

// textureSampleLevel of a cube map holding `mips` levels one after another
// (level m: six faces of max(size >> m, 1)^2 texels), linear between levels;
// the level is clamped to [0, mips - 1].
template <class T> v3<T> cube_sample(const float* data, uint32_t size, uint32_t mips, v3<T> dir, T level) {
    CubeUV<T> c = cube_face(dir);
    level = clampf(level, T(0), T((float)(mips - 1u)));
    T fl = std::floor(level);
    uint32_t l0 = (uint32_t)fl, l1 = std::min(l0 + 1u, mips - 1u);
    T f = level - fl;
    size_t base = 0;
    const float* lv[2] = {nullptr, nullptr};
    uint32_t sz[2] = {0, 0};
    for (uint32_t m = 0; m <= l1; ++m) {
        uint32_t sm = std::max(size >> m, 1u);
        if (m == l0) { lv[0] = data + base; sz[0] = sm; }
        if (m == l1) { lv[1] = data + base; sz[1] = sm; }
        base += (size_t)6u * sm * sm * 4u;
    }
    v3<T> a = cube_bilinear(lv[0], sz[0], c);
    if (l1 == l0) return a;
    return mix3(a, cube_bilinear(lv[1], sz[1], c), f);
}

// shade.wesl:396-427 (the single-probe sample_environment, shade.wesl:365-382).
template <class T>
v3<T> probe_radiance(const hm_env_probe* env, v3<T> tangent, v3<T> fiber_n, v3<T> view, v3<T> albedo, T roughness,
                     T tilt) {
    if (!env) return splat3(T(0));
    T theta_o = std::asin(clampf(dot3(tangent, view), T(-1), T(1)));
    T theta_i = tilt - theta_o;
    v3<T> spec_dir = cubemap_dir(*env, tangent * std::sin(theta_i) + fiber_n * std::cos(theta_i));
    v3<T> diff_dir = cubemap_dir(*env, fiber_n);
    v3<T> spec = cube_sample(env->specular, env->specular_size, env->specular_mips, spec_dir,
                             roughness * T((float)(env->specular_mips - 1u)));
    v3<T> diff = cube_sample(env->diffuse, env->diffuse_size, 1u, diff_dir, T(0));
    T cos_o = fmax_(std::cos(theta_o), T(1e-3f));
    T f_r = fresnel(cos_o);
    T f_trt = fresnel(cos_o * T(0.5f));
    v3<T> sigma_a = absorption_from_color(albedo, roughness);
    v3<T> e = sigma_a * (T(-3.2f) / cos_o);
    v3<T> trt = v3<T>{std::exp(e.x), std::exp(e.y), std::exp(e.z)} *
                (T(0.76f) * (T(1) - f_trt) * (T(1) - f_trt) * f_trt);
    v3<T> scatter = albedo * scatter_tint(albedo, T(0.5f));
    return (spec * (splat3(f_r) + trt) + diff * scatter) * T(env->intensity);
}

template <class T>
v3<T> direct_light(const Params<T>& p, const DomView<T>& dv, const Lights& lights, v3<T> world, v3<T> tangent,
                   v3<T> view, v3<T> albedo, T roughness, T tilt, T beta) {  // shade.wesl:429-521
    (void)p;
    Shape<T> shape = lobe_shape(roughness, tilt);
    T tv = dot3(tangent, view);
    T cos_theta_o = std::sqrt(fmax_(T(1) - tv * tv, T(1e-6f)));
    v3<T> sigma_a = absorption_from_color(albedo, roughness);
    DualScatter<T> ds = dual_scatter(cos_theta_o, sigma_a, shape);
    v3<T> color = splat3(T(0));
    // Point lights, then spot lights (Bevy's cluster ranges list points first).
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t i = 0; i < lights.np; ++i) {
            const hm_point_light& L = lights.points[i];
            if ((pass == 0) != (L.kind == HM_LIGHT_POINT)) continue;
            v3<T> lp = {T(L.position_radius[0]), T(L.position_radius[1]), T(L.position_radius[2])};
            v3<T> light_to_frag = lp - world;
            T dist2 = dot3(light_to_frag, light_to_frag);
            v3<T> incident = safe_normalize(light_to_frag);
            T att = distance_attenuation(dist2, T(L.color_inverse_square_range[3]));
            v3<T> lobe = hair_bsdf(tangent, view, incident, sigma_a, shape, ds, T(1), T(0));
            v3<T> lc = {T(L.color_inverse_square_range[0]), T(L.color_inverse_square_range[1]),
                        T(L.color_inverse_square_range[2])};
            v3<T> c = lobe * lc * att;
            if (pass == 1) c = c * spot_mask<T>(L, world);
            color = color + c;
        }
    }
    bool dom_left = true;
    for (uint32_t i = 0; i < lights.nd; ++i) {
        const hm_directional_light& L = lights.dirs[i];
        v3<T> incident = {T(L.direction_to_light[0]), T(L.direction_to_light[1]), T(L.direction_to_light[2])};
        bool casts = (L.flags & HM_LIGHT_FLAG_SHADOWS) != 0u;
        T depth = T(0);
        if (dom_left && casts) {
            depth = -std::log(fmax_(dv.visibility(world, beta), T(1e-6f)));
            dom_left = false;
        }
        v3<T> lobe = hair_bsdf(tangent, view, incident, sigma_a, shape, ds, T(1), depth);
        color = color + lobe * v3<T>{T(L.color[0]), T(L.color[1]), T(L.color[2])};
    }
    for (uint32_t i = 0; i < lights.nr; ++i) {  // shade.wesl:503-518
        const hm_rect_light& L = lights.rects[i];
        v3<T> to_center = v3<T>{T(L.position[0]), T(L.position[1]), T(L.position[2])} - world;
        T dist2 = fmax_(dot3(to_center, to_center), T(1e-4f));
        v3<T> incident = to_center * inverse_sqrt(dist2);
        v3<T> normal = safe_normalize(cross3(v3<T>{T(L.up[0]), T(L.up[1]), T(L.up[2])},
                                             v3<T>{T(L.right[0]), T(L.right[1]), T(L.right[2])}));
        T facing = fmax_(dot3(normal, incident), T(0));
        T solid = T(L.width) * T(L.height) * facing / dist2;
        T range2 = fmax_(T(L.range) * T(L.range), T(1e-4f));
        T falloff = range_falloff(dist2, T(1) / range2);
        v3<T> lobe = hair_bsdf(tangent, view, incident, sigma_a, shape, ds, T(1), T(0));
        color = color + lobe * v3<T>{T(L.color[0]), T(L.color[1]), T(L.color[2])} * solid * falloff;
    }
    return color;
}

// The per-pixel inputs shade.wesl:fs derives from a covered pixel's words
// (shade.wesl:539-571).
template <class T> struct PixelIn { v3<T> world, tangent, view, albedo; T roughness, beta, view_dist, ao; bool ortho; };

template <class T>
PixelIn<T> pixel_inputs(const Params<T>& p, uint64_t word, uint32_t beta_word, uint32_t x, uint32_t y) {
    PixelIn<T> in;
    Unpacked u = unpack_gb(word);
    in.view_dist = dequantize_depth(u.depth, p.screen.z, p.screen.w);
    in.world = reconstruct_world(p, (int32_t)x, (int32_t)y, in.view_dist);
    in.tangent = oct_decode<T>(u.tx, u.ty);
    v3<T> uvw = {T((float)u.u) / T(63.0f), T((float)u.v) / T(63.0f), T((float)u.w) / T(63.0f)};
    in.beta = T((float)(beta_word & 0xffu)) / T(255.0f);
    in.ao = p.flags[2] == 0u ? T(1) : T((float)u.ao) / T(63.0f);
    in.roughness = fmax_(p.appearance.x * (T(0.65f) + T(0.7f) * uvw.z), T(0.04f));
    in.albedo = xyz(p.albedo) * (T(0.65f) + T(0.35f) * uvw.z);
    // Orthographic when the projection's last row is (0, 0, 0, 1).
    in.ortho = p.clip[3] == T(0) && p.clip[7] == T(0) && p.clip[11] == T(0) && p.clip[15] == T(1);
    if (in.ortho) in.view = safe_normalize(v3<T>{p.clip[2], p.clip[6], p.clip[10]});
    else in.view = safe_normalize(xyz(p.camera_pos) - in.world);
    return in;
}

template <class T>
void shade_pixel(const Params<T>& p, const DomView<T>& dv, const Lights& lights, const uint64_t* center,
                 const uint32_t* beta_buf, T exposure, uint32_t x, uint32_t y, float* shaded, float* shaded_depth) {
    // shade.wesl:523-593
    uint32_t width = to_u32(p.screen.x);
    size_t pix = (size_t)y * width + x;
    uint64_t word = center[pix];
    if (is_empty(word)) {
        shaded[pix * 4 + 0] = shaded[pix * 4 + 1] = shaded[pix * 4 + 2] = shaded[pix * 4 + 3] = 0.0f;
        shaded_depth[pix] = 0.0f;
        return;
    }
    PixelIn<T> in = pixel_inputs(p, word, beta_buf[pix], x, y);
    T view_dist = in.view_dist;
    v3<T> color = direct_light(p, dv, lights, in.world, in.tangent, in.view, in.albedo, in.roughness,
                               p.appearance.y, in.beta);
    // probe_radiance x ao (shade.wesl:585-586); zero without an environment
    // map (shade.wesl:384-394).
    if (lights.env) {
        v3<T> fiber_n = safe_normalize(in.view - in.tangent * dot3(in.view, in.tangent));
        color = color + probe_radiance(lights.env, in.tangent, fiber_n, in.view, in.albedo, in.roughness,
                                       p.appearance.y) * in.ao;
    }
    color = color * exposure;
    shaded[pix * 4 + 0] = (float)color.x;
    shaded[pix * 4 + 1] = (float)color.y;
    shaded[pix * 4 + 2] = (float)color.z;
    shaded[pix * 4 + 3] = 1.0f;
    shaded_depth[pix] = (float)view_dist;
}

// ---- hair_shadow.wesl: HairShadowMaterial's receiver lookup -----------------

// hair_shadow.wesl:30-50 with the map size at run time (the shader fixes 512).
template <class T>
T receiver_texel_visibility(const hm_shadow_light& L, const uint32_t* dom, uint32_t S, int32_t x, int32_t y, T dist) {
    if (x < 0 || y < 0 || x >= (int32_t)S || y >= (int32_t)S) return T(1);
    uint32_t pix = (uint32_t)y * S + (uint32_t)x;
    uint32_t front_q = dom[pix];
    if (front_q == HM_DOM_EMPTY) return T(1);
    T near_ = T(L.depth[0]), far_ = T(L.depth[1]);
    T z_front = mixs(near_, far_, T((float)front_q) / T(16777214.0f));
    T thickness = fmax_(far_ - near_, T(1e-3f)) / T((float)HM_DOM_LAYERS);
    T t = fmax_(dist - z_front, T(0)) / thickness;
    size_t base = (size_t)S * S + (size_t)pix * HM_DOM_LAYERS;
    T hits = T(0);
    for (uint32_t slice = 0; slice < HM_DOM_LAYERS; ++slice)
        hits = hits + T((float)dom[base + slice]) * clampf(t - T((float)slice), T(0), T(1));
    return std::exp(-hits * T(L.depth[2]));
}

template <class T> T hair_visibility(const hm_shadow_light& L, const uint32_t* dom, uint32_t S, v3<T> world) {
    // hair_shadow.wesl:52-75
    if (T(L.depth[3]) == T(0)) return T(1);
    T m[16];
    for (int i = 0; i < 16; ++i) m[i] = T(L.light_clip_from_world[i]);
    v4<T> clip = mat_mul(m, v4<T>{world.x, world.y, world.z, T(1)});
    if (clip.w <= T(0)) return T(1);
    T ndc_x = clip.x / clip.w, ndc_y = clip.y / clip.w;
    T ux = ndc_x * T(0.5f) + T(0.5f);
    T uy = T(1) - (ndc_y * T(0.5f) + T(0.5f));
    if (ux < T(0) || uy < T(0) || ux > T(1) || uy > T(1)) return T(1);
    v3<T> eye = {T(L.light_eye[0]), T(L.light_eye[1]), T(L.light_eye[2])};
    v3<T> fwd = {T(L.light_forward[0]), T(L.light_forward[1]), T(L.light_forward[2])};
    T dist = dot3(world - eye, fwd);
    T fx = ux * T((float)S) - T(0.5f);
    T fy = uy * T((float)S) - T(0.5f);
    int32_t x0 = to_i32(std::floor(fx)), y0 = to_i32(std::floor(fy));
    T wx = fx - std::floor(fx), wy = fy - std::floor(fy);
    T v00 = receiver_texel_visibility(L, dom, S, x0, y0, dist);
    T v10 = receiver_texel_visibility(L, dom, S, x0 + 1, y0, dist);
    T v01 = receiver_texel_visibility(L, dom, S, x0, y0 + 1, dist);
    T v11 = receiver_texel_visibility(L, dom, S, x0 + 1, y0 + 1, dist);
    return mixs(mixs(v00, v10, wx), mixs(v01, v11, wx), wy);
}

// ---- composite.wgsl with its pipeline state (pipeline.rs:299-339) -----------

// One fragment: colour with alpha >= 0.001 at the reconstructed point's clip
// depth (clamped to [0, 1]); depth compare >= with reverse Z (pass.rs:916-921),
// else <=; depth written; colour alpha-blended (rgb src * a + dst * (1 - a),
// alpha src + dst * (1 - a)).
template <class T>
void composite_pixel(const Params<T>& p, const float* color, const float* depth_buf, float* tc, float* td, int32_t x,
                     int32_t y) {
    int32_t width = to_i32(p.screen.x);
    size_t pix = (size_t)y * (uint32_t)width + (uint32_t)x;
    const float* c = color + pix * 4;
    if (c[3] < 0.001f) return;
    T view_dist = T(depth_buf[pix]);
    v3<T> world = reconstruct_world(p, x, y, view_dist);
    v4<T> clip = mat_mul(p.clip, v4<T>{world.x, world.y, world.z, T(1)});
    T depth = clampf(clip.z / fmax_(clip.w, T(1e-5f)), T(0), T(1));
    float d = (float)depth;
    bool pass = p.pass_mode[2] == 1u ? d >= td[pix] : d <= td[pix];
    if (!pass) return;
    td[pix] = d;
    float a = c[3];
    float* o = tc + pix * 4;
    for (int k = 0; k < 3; ++k) o[k] = c[k] * a + o[k] * (1.0f - a);
    o[3] = a + o[3] * (1.0f - a);
}

// ---- filter.wgsl -----------------------------------------------------------

template <class T> inline v2<T> screen_tangent(const Params<T>& p, v3<T> world, v3<T> tangent) {  // filter.wgsl:9-18
    Projected<T> a = project_world(p, world);
    Projected<T> b = project_world(p, world + tangent);
    v2<T> d = b.screen - a.screen;
    T l = len2(d);
    if (l < T(1e-3f)) return {T(1), T(0)};
    return {d.x / l, d.y / l};
}

template <class T>
void filter_pixel(const Params<T>& p, const uint64_t* center, const uint64_t* cons, const float* shaded,
                  const float* shaded_depth, float* filtered, float* filtered_depth, int32_t x, int32_t y) {
    // filter.wgsl:20-135
    (void)shaded_depth;
    int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
    size_t pix = (size_t)y * (uint32_t)width + (uint32_t)x;
    uint64_t center_word = center[pix];
    uint64_t cons_word = cons[pix];
    bool center_hit = !is_empty(center_word);
    bool cons_hit = !is_empty(cons_word);
    if (!center_hit && !cons_hit) {
        for (int c = 0; c < 4; ++c) filtered[pix * 4 + c] = 0.0f;
        filtered_depth[pix] = 0.0f;
        return;
    }
    Unpacked up = unpack_gb(center_hit ? center_word : cons_word);
    v3<T> tangent = oct_decode<T>(up.tx, up.ty);
    uint32_t depth_q = up.depth;
    T view_dist = dequantize_depth(depth_q, p.screen.z, p.screen.w);
    v3<T> world = reconstruct_world(p, x, y, view_dist);
    v2<T> t_screen = screen_tangent(p, world, tangent);
    T cc[4] = {T(shaded[pix * 4]), T(shaded[pix * 4 + 1]), T(shaded[pix * 4 + 2]), T(shaded[pix * 4 + 3])};
    const int radius = 5;
    T s_par = fmax_(p.filter_params.x, T(0.05f));
    T s_perp = fmax_(p.filter_params.y, T(0.02f));
    T sigma_par = T((float)radius) * s_par;
    T sigma_perp = T((float)radius) * s_perp;
    const T sigma_c = T(0.9f), depth_reject = T(1.45e-3f), closer_eps = T(5e-4f), strip = T(0.5f);
    v3<T> accum = splat3(T(0));
    T weight = T(0), h_center = T(0), h_cons = T(0);
    bool promote = !center_hit;
    for (int dy = -radius; dy <= radius; ++dy) {
        for (int dx = -radius; dx <= radius; ++dx) {
            int32_t qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= width || qy >= height) continue;
            size_t qpix = (size_t)qy * (uint32_t)width + (uint32_t)qx;
            v2<T> d = {T((float)dx), T((float)dy)};
            T d_par = dot2(d, t_screen);
            T d_perp = len2(d - t_screen * d_par);
            uint64_t q_center = center[qpix];
            uint64_t q_cons = cons[qpix];
            if (d_perp <= strip) {
                if (!is_empty(q_center)) h_center = h_center + T(1);
                else if (!is_empty(q_cons)) h_cons = h_cons + T(1);
            }
            if (is_empty(q_center)) continue;
            Unpacked qu = unpack_gb(q_center);
            T q_norm = T((float)qu.depth) / T(16777214.0f);
            T p_norm = T((float)depth_q) / T(16777214.0f);
            if (std::fabs(q_norm - p_norm) > depth_reject) continue;
            T spatial = std::exp(-d_par * d_par / (sigma_par * sigma_par) - d_perp * d_perp / (sigma_perp * sigma_perp));
            T color_w = T(1);
            v3<T> q_color = {T(shaded[qpix * 4]), T(shaded[qpix * 4 + 1]), T(shaded[qpix * 4 + 2])};
            if (!promote && cc[3] > T(0)) {
                T chroma = len3(v3<T>{cc[0], cc[1], cc[2]} - q_color);
                color_w = std::exp(-chroma * chroma / (sigma_c * sigma_c));
            }
            T w = spatial * color_w;
            accum = accum + q_color * w;
            weight = weight + w;
        }
    }
    v3<T> color = {cc[0], cc[1], cc[2]};
    if (weight > T(1e-5f)) color = accum / weight;
    T alpha;
    T coverage = h_center / fmax_(h_center + h_cons, T(1));
    T cons_depth = cons_hit ? T((float)unpack_gb(cons_word).depth) / T(16777214.0f) : T(1);
    T center_depth = T((float)depth_q) / T(16777214.0f);
    bool closer = cons_hit && center_hit && (center_depth - cons_depth) >= closer_eps;
    if (closer) alpha = T(1);
    else alpha = clampf(coverage, T(0), T(1));
    if (promote && weight <= T(1e-5f)) {
        alpha = T(0);
        color = splat3(T(0));
    }
    filtered[pix * 4 + 0] = (float)color.x;
    filtered[pix * 4 + 1] = (float)color.y;
    filtered[pix * 4 + 2] = (float)color.z;
    filtered[pix * 4 + 3] = (float)alpha;
    filtered_depth[pix] = (float)view_dist;
}

template <class T, bool ATOMIC>
void raster_range(const hm_params* prm, const hm_bundle* bundles, const hm_layer* layers, const float* style,
                  const hm_strand_ref* refs, uint32_t begin, uint32_t end, Targets* tg, uint64_t counts[4]) {
    Params<T> p = load_params<T>(*prm);
    Raster<T, ATOMIC>* r = new Raster<T, ATOMIC>{p, bundles, layers, style, *tg};
    for (uint32_t i = begin; i < end; ++i) r->strand(refs[i]);
    for (int k = 0; k < 4; ++k) counts[k] += r->counts[k];
    delete r;
}

}  // namespace

// ---- C entry points -----------------------------------------------------------

extern "C" {

uint32_t or_pcg(uint32_t x) { return pcg(x); }

void or_pcg_uv(uint32_t seed, float out[2]) {
    v2<float> uv = pcg_uv<float>(seed);
    out[0] = uv.x;
    out[1] = uv.y;
}

uint64_t or_pack_gb(uint32_t depth, uint32_t tx, uint32_t ty, uint32_t u, uint32_t v, uint32_t w, uint32_t ao) {
    return pack_gb(depth, tx, ty, u, v, w, ao);
}

void or_unpack_gb(uint64_t word, uint32_t out[7]) {
    Unpacked u = unpack_gb(word);
    uint32_t vals[7] = {u.depth, u.tx, u.ty, u.u, u.v, u.w, u.ao};
    std::memcpy(out, vals, sizeof(vals));
}

int or_is_empty(uint64_t word) { return is_empty(word) ? 1 : 0; }

uint32_t or_quantize_depth(float dist, float near_, float far_) { return quantize_depth<float>(dist, near_, far_); }
float or_dequantize_depth(uint32_t d, float near_, float far_) { return dequantize_depth<float>(d, near_, far_); }

void or_oct_encode(const float n[3], uint32_t out[2]) { oct_encode<float>({n[0], n[1], n[2]}, out[0], out[1]); }
void or_oct_decode(uint32_t tx, uint32_t ty, float out[3]) {
    v3<float> n = oct_decode<float>(tx, ty);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
}

uint32_t or_control_points_for_lod(float l) { return control_points_for_lod<float>(l); }

/* Hermite-cage strand point as eval_cp computes it (raster.wgsl:62-75 with the
 * layer stage raster.wgsl:236-267), for one strand uv and control point. */
void or_eval_cp(const hm_bundle* bundles, const hm_layer* layers, const float* style, uint32_t bundle_id, float u,
                float v, uint32_t idx, uint32_t ncp, float out[4]) {
    hm_params dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    Params<float> p = load_params<float>(dummy);
    Targets tg{};
    Raster<float, false>* r = new Raster<float, false>{p, bundles, layers, style, tg};
    const hm_bundle& b = bundles[bundle_id];
    v2<float> uv = {u, v};
    v4<float> loaded[16];
    for (uint32_t lane = 0; lane < 16; ++lane) {
        auto s = r->sample_layer(b.layer_offset, lane, uv);
        loaded[lane] = s.pos_ao;
        r->sm[lane] = s;
    }
    for (uint32_t lane = 0; lane < 16; ++lane) {
        v3<float> prev_p = xyz(loaded[lane == 0 ? 0 : lane - 1]);
        v3<float> next_p = xyz(loaded[std::min(lane + 1u, 15u)]);
        v3<float> tn = calc_tangent(lane, 16u, prev_p, xyz(loaded[lane]), next_p);
        r->sm[lane].tangent = {tn.x, tn.y, tn.z, 0.0f};
    }
    v4<float> cp = r->eval_cp(idx, ncp, uv, b.style_offset);
    out[0] = cp.x;
    out[1] = cp.y;
    out[2] = cp.z;
    out[3] = cp.w;
    delete r;
}

float or_dom_optical_depth(const uint32_t* layers16, float t) {
    hm_params prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.dom_info[0] = 0.0f;
    prm.dom_info[1] = 0.0f;
    Params<float> p = load_params<float>(prm);
    DomView<float> dv{p, layers16};
    return dv.optical_depth(0, t);
}

void or_clear(uint64_t* center, uint64_t* cons, uint32_t* beta, size_t pixels) {  // clear.wgsl:7-16
    for (size_t i = 0; i < pixels; ++i) {
        center[i] = HM_GB_EMPTY;
        cons[i] = HM_GB_EMPTY;
        beta[i] = 0xffffffffu;
    }
}

void or_clear_dom(uint32_t* dom, uint32_t dom_size) {  // clear_dom.wgsl:8-21 with a runtime size
    size_t texels = (size_t)dom_size * dom_size;
    size_t total = texels * (1 + HM_DOM_LAYERS);
    for (size_t i = 0; i < total; ++i) dom[i] = i < texels ? 0xffffffffu : 0u;
}

void or_lod(const hm_params* prm, const hm_bundle* bundles, const hm_layer* layers, hm_lod_record* out,
            const hm_lod_override* ovr, int f64) {
    uint32_t nb = prm->pass_mode[1];
    if (f64) {
        Params<double> p = load_params<double>(*prm);
        for (uint32_t b = 0; b < nb; ++b) lod_bundle(p, bundles, layers, b, out, ovr);
    } else {
        Params<float> p = load_params<float>(*prm);
        for (uint32_t b = 0; b < nb; ++b) lod_bundle(p, bundles, layers, b, out, ovr);
    }
}

uint32_t or_scan(const hm_params* prm, const hm_lod_record* lod, hm_strand_ref* refs, uint32_t capacity,
                 uint32_t indirect[3]) {  // scan.wgsl:6-24
    uint32_t total = 0;
    uint32_t bundle_count = prm->pass_mode[1];
    for (uint32_t bundle = 0; bundle < bundle_count; ++bundle) {
        hm_lod_record rec = lod[bundle];
        for (uint32_t strand = 0; strand < rec.n_lod; ++strand) {
            if (total >= capacity) break;
            refs[total] = {bundle, strand, rec.n_lod, rec.control_points};
            total += 1;
        }
    }
    indirect[0] = total;
    indirect[1] = 1;
    indirect[2] = 1;
    return total;
}

/* raster.wgsl:225-337 over refs[0..count). Targets may be NULL (counting only). */
void or_raster(const hm_params* prm, const hm_bundle* bundles, const hm_layer* layers, const float* style,
               const hm_strand_ref* refs, uint32_t count, uint64_t* center, uint64_t* cons, uint32_t* beta,
               uint32_t* dom, const float* scene_depth, float scene_clear, uint64_t counts[4], int f64,
               int threads) {
    Targets tg{center, cons, beta, dom, scene_depth, scene_clear, {0, 0, 0, 0}};
    uint64_t c[4] = {0, 0, 0, 0};
    if (threads <= 1) {
        if (f64) raster_range<double, false>(prm, bundles, layers, style, refs, 0, count, &tg, c);
        else raster_range<float, false>(prm, bundles, layers, style, refs, 0, count, &tg, c);
    } else {
        std::vector<std::thread> pool;
        std::vector<uint64_t> part((size_t)threads * 4, 0);
        for (int t = 0; t < threads; ++t) {
            uint32_t b = (uint32_t)((uint64_t)count * t / threads);
            uint32_t e = (uint32_t)((uint64_t)count * (t + 1) / threads);
            pool.emplace_back([=, &tg, &part]() {
                if (f64) raster_range<double, true>(prm, bundles, layers, style, refs, b, e, &tg, &part[t * 4]);
                else raster_range<float, true>(prm, bundles, layers, style, refs, b, e, &tg, &part[t * 4]);
            });
        }
        for (auto& th : pool) th.join();
        for (int t = 0; t < threads; ++t)
            for (int k = 0; k < 4; ++k) c[k] += part[(size_t)t * 4 + k];
    }
    if (counts)
        for (int k = 0; k < 4; ++k) counts[k] = c[k];
}

/* shade.wesl:523-593 for every pixel, rows split over `threads`. */
void or_shade(const hm_params* prm, const uint64_t* center, const uint32_t* beta, const uint32_t* dom,
              const hm_point_light* points, uint32_t np, const hm_directional_light* dirs, uint32_t nd,
              const hm_rect_light* rects, uint32_t nr, const hm_env_probe* env, float exposure, float* shaded,
              float* shaded_depth, int f64, int threads, uint32_t row_begin, uint32_t row_end) {
    Lights lights{points, np, dirs, nd, rects, nr, env};
    uint32_t width = (uint32_t)prm->screen[0];
    uint32_t height = std::min(row_end, (uint32_t)prm->screen[1]);
    auto run = [&](uint32_t y0, uint32_t y1) {
        if (f64) {
            Params<double> p = load_params<double>(*prm);
            DomView<double> dv{p, dom};
            for (uint32_t y = y0; y < y1; ++y)
                for (uint32_t x = 0; x < width; ++x)
                    shade_pixel<double>(p, dv, lights, center, beta, exposure, x, y, shaded, shaded_depth);
        } else {
            Params<float> p = load_params<float>(*prm);
            DomView<float> dv{p, dom};
            for (uint32_t y = y0; y < y1; ++y)
                for (uint32_t x = 0; x < width; ++x)
                    shade_pixel<float>(p, dv, lights, center, beta, exposure, x, y, shaded, shaded_depth);
        }
    };
    if (row_begin >= height) return;
    uint32_t rows = height - row_begin;
    if (threads <= 1) {
        run(row_begin, height);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back(run, row_begin + (uint32_t)((uint64_t)rows * t / threads),
                          row_begin + (uint32_t)((uint64_t)rows * (t + 1) / threads));
    for (auto& th : pool) th.join();
}

/* filter.wgsl:20-135 for every pixel. */
void or_filter(const hm_params* prm, const uint64_t* center, const uint64_t* cons, const float* shaded,
               const float* shaded_depth, float* filtered, float* filtered_depth, int f64, int threads,
               uint32_t row_begin, uint32_t row_end) {
    int32_t width = (int32_t)prm->screen[0];
    int32_t height = (int32_t)std::min(row_end, (uint32_t)prm->screen[1]);
    int32_t first = (int32_t)row_begin;
    auto run = [&](int32_t y0, int32_t y1) {
        if (f64) {
            Params<double> p = load_params<double>(*prm);
            for (int32_t y = y0; y < y1; ++y)
                for (int32_t x = 0; x < width; ++x)
                    filter_pixel<double>(p, center, cons, shaded, shaded_depth, filtered, filtered_depth, x, y);
        } else {
            Params<float> p = load_params<float>(*prm);
            for (int32_t y = y0; y < y1; ++y)
                for (int32_t x = 0; x < width; ++x)
                    filter_pixel<float>(p, center, cons, shaded, shaded_depth, filtered, filtered_depth, x, y);
        }
    };
    if (first >= height) return;
    int64_t rows = height - first;
    if (threads <= 1) {
        run(first, height);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back(run, first + (int32_t)(rows * t / threads), first + (int32_t)(rows * (t + 1) / threads));
    for (auto& th : pool) th.join();
}

/* BSDF pieces of shade.wesl in double precision, for textbook checks:
 * out = N_R, N_TT.rgb, N_TRT.rgb of azimuthal() (shade.wesl:94-116). */
void or_azimuthal(double cos_theta_d, double cos_phi, const double sigma_a[3], double out[7]) {
    Lobes<double> l = azimuthal<double>(cos_theta_d, cos_phi, {sigma_a[0], sigma_a[1], sigma_a[2]});
    double v[7] = {l.r.x, l.tt.x, l.tt.y, l.tt.z, l.trt.x, l.trt.y, l.trt.z};
    std::memcpy(out, v, sizeof(v));
}

/* dual_scatter (shade.wesl:153-190): out = fwd.r, fwd.tt, fwd.trt, a_f, beta_f2,
 * a_back, delta_b, sigma_b2 as 8 rgb triples. */
void or_dual_scatter(double cos_theta_d, const double sigma_a[3], double roughness, double tilt, double out[24]) {
    Shape<double> sh = lobe_shape<double>(roughness, tilt);
    DualScatter<double> d = dual_scatter<double>(cos_theta_d, {sigma_a[0], sigma_a[1], sigma_a[2]}, sh);
    v3<double> parts[8] = {d.fwd.r, d.fwd.tt, d.fwd.trt, d.a_f, d.beta_f2, d.a_back, d.delta_b, d.sigma_b2};
    for (int i = 0; i < 8; ++i) {
        out[i * 3] = parts[i].x;
        out[i * 3 + 1] = parts[i].y;
        out[i * 3 + 2] = parts[i].z;
    }
}

/* hair_bsdf (shade.wesl:193-244) for one light, shadow 1. */
void or_hair_bsdf(const double t[3], const double v[3], const double l[3], const double albedo[3], double roughness,
                  double tilt, double depth, double out[3]) {
    Shape<double> sh = lobe_shape<double>(roughness, tilt);
    v3<double> tt = {t[0], t[1], t[2]}, vv = {v[0], v[1], v[2]}, ll = {l[0], l[1], l[2]};
    v3<double> sigma = absorption_from_color<double>({albedo[0], albedo[1], albedo[2]}, roughness);
    double tv = dot3(tt, vv);
    double cos_theta_o = std::sqrt(fmax_(1.0 - tv * tv, 1e-6));
    DualScatter<double> ds = dual_scatter<double>(cos_theta_o, sigma, sh);
    v3<double> c = hair_bsdf<double>(tt, vv, ll, sigma, sh, ds, 1.0, depth);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

/* scatter_sums (shade.wesl:153-165): out = fwd r, tt, trt then back r, tt, trt, rgb each. */
void or_scatter_sums(double cos_theta_d, const double sigma_a[3], double out[18]) {
    ScatterSums<double> s = scatter_sums<double>(cos_theta_d, {sigma_a[0], sigma_a[1], sigma_a[2]});
    v3<double> parts[6] = {s.fwd.r, s.fwd.tt, s.fwd.trt, s.back.r, s.back.tt, s.back.trt};
    for (int i = 0; i < 6; ++i) {
        out[i * 3] = parts[i].x;
        out[i * 3 + 1] = parts[i].y;
        out[i * 3 + 2] = parts[i].z;
    }
}
/* dual_terms (shade.wesl:167-189) from given lobe sums fwd / back (r, tt, trt,
 * rgb each): out = a_f, alpha_f, beta_f2, a_back, delta_b, sigma_b2. */
void or_dual_terms(const double fwd[9], const double back[9], double roughness, double tilt, double out[18]) {
    Shape<double> sh = lobe_shape<double>(roughness, tilt);
    auto lobes = [](const double* a) {
        return Lobes<double>{{a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8]}};
    };
    DualScatter<double> d = dual_terms<double>(lobes(fwd), lobes(back), sh);
    v3<double> parts[6] = {d.a_f, d.alpha_f, d.beta_f2, d.a_back, d.delta_b, d.sigma_b2};
    for (int i = 0; i < 6; ++i) {
        out[i * 3] = parts[i].x;
        out[i * 3 + 1] = parts[i].y;
        out[i * 3 + 2] = parts[i].z;
    }
}
/* hair_bsdf_parts for one light: out = single, back, global (rgb each), exp(-depth). */
void or_bsdf_parts(const double t[3], const double v[3], const double l[3], const double albedo[3], double roughness,
                   double tilt, double depth, double out[10]) {
    Shape<double> sh = lobe_shape<double>(roughness, tilt);
    v3<double> tt = {t[0], t[1], t[2]}, vv = {v[0], v[1], v[2]}, ll = {l[0], l[1], l[2]};
    v3<double> sigma = absorption_from_color<double>({albedo[0], albedo[1], albedo[2]}, roughness);
    double tv = dot3(tt, vv);
    double cos_theta_o = std::sqrt(fmax_(1.0 - tv * tv, 1e-6));
    DualScatter<double> ds = dual_scatter<double>(cos_theta_o, sigma, sh);
    BsdfParts<double> b = hair_bsdf_parts<double>(tt, vv, ll, sigma, sh, ds, depth);
    v3<double> parts[3] = {b.single, b.back, b.global};
    for (int i = 0; i < 3; ++i) {
        out[i * 3] = parts[i].x;
        out[i * 3 + 1] = parts[i].y;
        out[i * 3 + 2] = parts[i].z;
    }
    out[9] = b.direct;
}
/* dom_visibility (shade.wesl:279-310), float path, for a world point and beta. */
float or_dom_visibility(const hm_params* prm, const uint32_t* dom, const float world[3], float beta) {
    Params<float> p = load_params<float>(*prm);
    DomView<float> dv{p, dom};
    return dv.visibility({world[0], world[1], world[2]}, beta);
}
float or_spot_mask(const hm_point_light* light, const float world[3]) {  // shade.wesl:312-327
    return spot_mask<float>(*light, {world[0], world[1], world[2]});
}
float or_distance_attenuation(float dist2, float inv_range2) { return distance_attenuation<float>(dist2, inv_range2); }
double or_range_falloff(double dist2, double inv_range2) { return range_falloff<double>(dist2, inv_range2); }

/* scatter_tint (shade.wesl:81-85). */
void or_scatter_tint(const double albedo[3], double shadow, double out[3]) {
    v3<double> r = scatter_tint<double>({albedo[0], albedo[1], albedo[2]}, shadow);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* cube_face: out = face, s, t. */
void or_cube_face(const double dir[3], double out[3]) {
    CubeUV<double> c = cube_face<double>({dir[0], dir[1], dir[2]});
    out[0] = c.face; out[1] = c.s; out[2] = c.t;
}

/* cube_sample of a cube map with `mips` levels at `level`, no view rotation. */
void or_cube_sample(const float* data, uint32_t size, uint32_t mips, const double dir[3], double level,
                    double out[3]) {
    v3<double> r = cube_sample<double>(data, size, mips, {dir[0], dir[1], dir[2]}, level);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* cubemap_dir (shade.wesl:334-339). */
void or_cubemap_dir(const hm_env_probe* env, const double dir[3], double out[3]) {
    v3<double> r = cubemap_dir<double>(*env, {dir[0], dir[1], dir[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* probe_radiance (shade.wesl:396-427) with fiber_n formed as shade.wesl:fs does. */
void or_probe_radiance(const hm_env_probe* env, const double tangent[3], const double view[3],
                       const double albedo[3], double roughness, double tilt, double out[3]) {
    v3<double> t = {tangent[0], tangent[1], tangent[2]}, v = {view[0], view[1], view[2]};
    v3<double> fiber_n = safe_normalize(v - t * dot3(v, t));
    v3<double> r = probe_radiance<double>(env, t, fiber_n, v, {albedo[0], albedo[1], albedo[2]}, roughness, tilt);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void or_reconstruct_world(const hm_params* prm, int32_t px, int32_t py, float view_dist, float out[3]) {
    Params<float> p = load_params<float>(*prm);
    v3<float> w = reconstruct_world(p, px, py, view_dist);
    out[0] = w.x;
    out[1] = w.y;
    out[2] = w.z;
}
/* project_world (common.wgsl:212-219): out = screen x, y, view distance, clip depth, in_front. */
void or_project_world(const hm_params* prm, const float world[3], float out[5]) {
    Params<float> p = load_params<float>(*prm);
    Projected<float> q = project_world(p, v3<float>{world[0], world[1], world[2]});
    out[0] = q.screen.x;
    out[1] = q.screen.y;
    out[2] = q.view_dist;
    out[3] = q.clip_depth;
    out[4] = q.in_front ? 1.0f : 0.0f;
}
void or_screen_tangent(const hm_params* prm, const float world[3], const float tangent[3], float out[2]) {
    Params<float> p = load_params<float>(*prm);  // filter.wgsl:9-18
    v2<float> d = screen_tangent(p, v3<float>{world[0], world[1], world[2]}, v3<float>{tangent[0], tangent[1], tangent[2]});
    out[0] = d.x;
    out[1] = d.y;
}
/* pixel_inputs: out = world, tangent, view, albedo (3 each), roughness, beta, view distance, ortho. */
void or_pixel_inputs(const hm_params* prm, uint64_t word, uint32_t beta_word, uint32_t x, uint32_t y, float out[16]) {
    Params<float> p = load_params<float>(*prm);
    PixelIn<float> in = pixel_inputs(p, word, beta_word, x, y);
    v3<float> parts[4] = {in.world, in.tangent, in.view, in.albedo};
    for (int i = 0; i < 4; ++i) {
        out[i * 3] = parts[i].x;
        out[i * 3 + 1] = parts[i].y;
        out[i * 3 + 2] = parts[i].z;
    }
    out[12] = in.roughness;
    out[13] = in.beta;
    out[14] = in.view_dist;
    out[15] = in.ortho ? 1.0f : 0.0f;
}
/* hair_visibility (hair_shadow.wesl:52-75) of a world point, map of S x S texels. */
float or_hair_visibility(const hm_shadow_light* light, const uint32_t* dom, uint32_t S, const float world[3]) {
    return hair_visibility<float>(*light, dom, S, v3<float>{world[0], world[1], world[2]});
}
/* composite.wgsl over the whole screen into target_color (float4) and target_depth. */
void or_composite(const hm_params* prm, const float* color, const float* depth_buf, float* target_color,
                  float* target_depth) {
    Params<float> p = load_params<float>(*prm);
    int32_t width = (int32_t)prm->screen[0], height = (int32_t)prm->screen[1];
    for (int32_t y = 0; y < height; ++y)
        for (int32_t x = 0; x < width; ++x) composite_pixel<float>(p, color, depth_buf, target_color, target_depth, x, y);
}
double or_fresnel(double c) { return fresnel<double>(c); }

void or_absorption_from_color(const double color[3], double roughness, double out[3]) {
    v3<double> s = absorption_from_color<double>({color[0], color[1], color[2]}, roughness);
    out[0] = s.x;
    out[1] = s.y;
    out[2] = s.z;
}

}  // extern "C"
