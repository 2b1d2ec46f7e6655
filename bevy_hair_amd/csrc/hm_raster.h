// Strand generation and segment drawing of raster.wgsl, shared by the direct
// (global-atomic) raster and the binned (LDS-tile) raster. Paths are relative
// to the reference repository.
//
// gen_strand:   raster.wgsl:231-317 — one wave per strand. Lanes 0..15 sample
//               the 16 cage layers (raster.wgsl:22-32), lanes 16..31 the 16
//               style levels at the strand's uv (the u/v half of
//               raster.wgsl:34-60, shared by every control point of the
//               strand), control point idx lives in lane idx % 64 (slot idx/64).
//               strand_layers / strand_cp are its two halves, also used by
//               the binned generation with two strands per wave.
// Drawer:       draw_segment/write_pixel (raster.wgsl:77-116, 192-223) and
//               draw_dom_segment/write_dom_pixel (raster.wgsl:118-190). The
//               write target is a Sink: global atomics, counters, or an LDS tile.
#pragma once

#include "hm_device.h"
#include "hm_kernels.h"

namespace hm {

__device__ __forceinline__ F3 shfl3(F3 v, int src) { return {__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)}; }
__device__ __forceinline__ F4 shfl4(F4 v, int src) {
    return {__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src)};
}

// Layer and style samples of one strand, held by a group of lanes starting
// at lane gb (r = lane - gb): layer k (position + ao, and its tangent) and
// style level k both sit in lane gb + k for k < 16 (raster.wgsl:22-60,
// 244-267). Lanes r < 16 sample the layers, lanes 16 <= r < 32 the style
// levels; other lanes of the group hold unused values.
struct StrandLayers {
    F4 lay;         // layer position (xyz) and ao (w)
    F3 ltan;        // layer tangent
    F3 sty;         // style offset of level r
    F2 uv;
    float beta;     // n_lod / max(N_b, 1)
    uint32_t ncp;   // clamp(control_points, 2, 127)
    float den, inv_den;  // max(ncp - 1, 1) and its reciprocal
};

__device__ __forceinline__ StrandLayers strand_layers(const hm_strand_ref& ref, const hm_bundle* bundles,
                                                      const hm_layer* layers, const float* style, int r, int gb) {
    StrandLayers o;
    const hm_bundle bundle = bundles[ref.bundle];
    const F2 uv = pcg_uv(ref.strand ^ (ref.bundle * 0x85ebca6bu));
    F4 lay = {0, 0, 0, 0};
    F3 sty = {0, 0, 0};
    if (r < 16) {
        const hm_layer* L = layers + bundle.layer_offset + (uint32_t)r * 4u;
        lay = bilinear4(ld4(L[0].pos_ao), ld4(L[1].pos_ao), ld4(L[2].pos_ao), ld4(L[3].pos_ao), uv);
    } else if (r < 32) {
        uint32_t k = (uint32_t)r - 16u;
        float u = clampf(uv.x, 0.0f, 1.0f) * 7.0f;
        float v = clampf(uv.y, 0.0f, 1.0f) * 7.0f;
        uint32_t u0 = to_u32(floorf(u)), v0 = to_u32(floorf(v));
        uint32_t u1 = u0 + 1u < 7u ? u0 + 1u : 7u;
        uint32_t v1 = v0 + 1u < 7u ? v0 + 1u : 7u;
        float tu = u - u2f(u0), tv = v - u2f(v0);
        uint32_t off = bundle.style_offset + k * 64u;
        F3 c000 = xyz(ld4(style + 4u * (off + v0 * 8u + u0)));
        F3 c100 = xyz(ld4(style + 4u * (off + v0 * 8u + u1)));
        F3 c010 = xyz(ld4(style + 4u * (off + v1 * 8u + u0)));
        F3 c110 = xyz(ld4(style + 4u * (off + v1 * 8u + u1)));
        sty = mix3(mix3(c000, c100, tu), mix3(c010, c110, tu), tv);
    }
    o.sty = shfl3(sty, gb + ((r + 16) & 31));  // style level k now sits in lane gb + k

    // Layer tangents (raster.wgsl:244-267).
    F3 lp = xyz(lay);
    F3 prev_l = shfl3(lp, gb + (r > 0 ? r - 1 : 0));
    F3 next_l = shfl3(lp, gb + (r < 15 ? r + 1 : 15));
    o.ltan = calc_tangent((uint32_t)r, HM_LAYER_COUNT, prev_l, lp, next_l);
    o.lay = lay;
    o.uv = uv;
    o.ncp = ref.control_points < 2u ? 2u : (ref.control_points > 127u ? 127u : ref.control_points);
    o.den = u2f(o.ncp - 1u > 1u ? o.ncp - 1u : 1u);
    o.inv_den = rdiv(1.0f, o.den);
    o.beta = rdiv(u2f(ref.n_lod), u2f(bundle.strand_count > 1u ? bundle.strand_count : 1u));
    return o;
}

// Control point idx of the strand whose samples the group at lane gb holds
// (raster.wgsl:270-284, eval_cp raster.wgsl:62-75): model-space position,
// ao and the parameter w = idx / (ncp - 1).
__device__ __forceinline__ F3 strand_cp(const StrandLayers& L, uint32_t idx, int gb, float& w, float& ao) {
    w = div_by_recip(u2f(idx), L.den, L.inv_den);  // exact for idx <= den; other lanes are unused
    float layers_f = w * 15.0f;
    uint32_t i0 = to_u32(floorf(layers_f));
    i0 = i0 < 15u ? i0 : 15u;
    uint32_t i1 = i0 + 1u < 15u ? i0 + 1u : 15u;
    float t = layers_f - u2f(i0);
    F4 a0 = shfl4(L.lay, gb + (int)i0), a1 = shfl4(L.lay, gb + (int)i1);
    F3 m0 = shfl3(L.ltan, gb + (int)i0), m1 = shfl3(L.ltan, gb + (int)i1);
    F3 base = hermite(xyz(a0), xyz(a1), m0, m1, t);
    ao = mixs(a0.w, a1.w, t);
    // w is in [0, 1], so the style level index and blend equal i0/i1/t.
    F3 s0 = shfl3(L.sty, gb + (int)i0), s1 = shfl3(L.sty, gb + (int)i1);
    return base + mix3(s0, s1, t);
}

// Per-lane result of gen_strand: control points idx = lane + 64 * j.
struct StrandCps {
    F3 cw[2];       // world position (to_world)
    F3 ct[2];       // control-point tangent, model space, normalized
    float w[2], ao[2];
    F2 uv;
    float beta;     // n_lod / max(N_b, 1)
    uint32_t ncp;   // clamp(control_points, 2, 127)
};

// One strand per wave (raster.wgsl:231-317).
__device__ __forceinline__ StrandCps gen_strand(const View& p, const hm_strand_ref& ref, const hm_bundle* bundles,
                                                const hm_layer* layers, const float* style, int lane) {
    StrandCps o;
    const StrandLayers L = strand_layers(ref, bundles, layers, style, lane, 0);
    const uint32_t ncp = L.ncp;
    F3 cp[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) cp[j] = strand_cp(L, (uint32_t)lane + 64u * (uint32_t)j, 0, o.w[j], o.ao[j]);
    // Control-point tangents (raster.wgsl:286-317).
    F3 up0 = shfl3(cp[0], (lane + 63) & 63), up1 = shfl3(cp[1], (lane + 63) & 63);
    F3 dn0 = shfl3(cp[0], (lane + 1) & 63), dn1 = shfl3(cp[1], (lane + 1) & 63);
    {
        uint32_t idx = (uint32_t)lane;
        F3 prev = lane > 0 ? up0 : cp[0];
        F3 next = idx + 1u < ncp ? (lane < 63 ? dn0 : dn1) : cp[0];
        o.ct[0] = safe_normalize(calc_tangent(idx, ncp, prev, cp[0], next));
        idx += 64u;
        prev = lane > 0 ? up1 : up0;
        next = idx + 1u < ncp ? dn1 : cp[1];
        o.ct[1] = safe_normalize(calc_tangent(idx, ncp, prev, cp[1], next));
    }
    o.beta = L.beta;
#pragma unroll
    for (int j = 0; j < 2; ++j) o.cw[j] = to_world(p, cp[j]);
    o.uv = L.uv;
    o.ncp = ncp;
    return o;
}

// ---- write targets ------------------------------------------------------------

struct Counters {
    uint64_t cam_writes, cam_centers, dom_writes, segments;
};

// Global atomics, optionally screened by a plain load: targets only decrease
// during a pass, so a loaded value bounds the current one from above and a
// payload that is not below it cannot change the word.
struct GlobalSink {
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    uint32_t* dom;
    size_t dom_pixels;
    bool check;
    __device__ __forceinline__ bool accepts(int32_t, int32_t) const { return true; }
    __device__ __forceinline__ bool walked_out(int32_t, int32_t, int32_t, int32_t) const { return false; }
    __device__ __forceinline__ void min64(uint64_t* a, uint64_t v) const {
        if (!check || v < *a) atomicMin((unsigned long long*)a, (unsigned long long)v);
    }
    __device__ __forceinline__ void min32(uint32_t* a, uint32_t v) const {
        if (!check || v < *a) atomicMin(a, v);
    }
    __device__ __forceinline__ void cam(size_t pix, int32_t, int32_t, uint64_t payload, bool center_hit,
                                        uint32_t beta_word, Counters&) {
        min64(cons + pix, payload);
        if (center_hit) {
            min64(center + pix, payload);
            min32(beta + pix, beta_word);
        }
    }
    __device__ __forceinline__ void dom_min(size_t pix, int32_t, int32_t, uint32_t q, float, Counters&) {
        min32(dom + pix, q);
    }
    __device__ __forceinline__ uint32_t dom_front(size_t pix, int32_t, int32_t) const { return dom[pix]; }
    __device__ __forceinline__ void dom_add(size_t pix, int32_t, int32_t, uint32_t slice, Counters&) {
        atomicAdd(dom + dom_pixels + pix * HM_DOM_LAYERS + slice, 1u);
    }
};

// Counts the writes that would reach the targets; the DOM front is read.
struct CountSink {
    const uint32_t* dom;
    __device__ __forceinline__ bool accepts(int32_t, int32_t) const { return true; }
    __device__ __forceinline__ bool walked_out(int32_t, int32_t, int32_t, int32_t) const { return false; }
    __device__ __forceinline__ void cam(size_t, int32_t, int32_t, uint64_t, bool center_hit, uint32_t, Counters& c) {
        c.cam_writes++;
        if (center_hit) c.cam_centers++;
    }
    __device__ __forceinline__ void dom_min(size_t, int32_t, int32_t, uint32_t, float, Counters& c) { c.dom_writes++; }
    __device__ __forceinline__ uint32_t dom_front(size_t pix, int32_t, int32_t) const { return dom[pix]; }
    __device__ __forceinline__ void dom_add(size_t, int32_t, int32_t, uint32_t, Counters& c) { c.dom_writes++; }
};

// The opacity pass's slice of a sample at view distance `dist` behind the
// texel's front (raster.wgsl:138-146): slab thickness (far - near) / 16 and
// min(u32(max(dist - z_front, 0) / thickness), 15), exactly.
struct DomSlab {
    float near_, far_, thickness, inv_thickness;
};
__device__ __forceinline__ DomSlab dom_slab(float near_, float far_) {
    const float thickness = rdiv(mx(far_ - near_, 1e-3f), 16.0f);
    return {near_, far_, thickness, rdiv(1.0f, thickness)};
}
__device__ __forceinline__ uint32_t dom_slice(const DomSlab& s, uint32_t front_q, float dist) {
    const float z_front = dequantize_depth(front_q, s.near_, s.far_);
    return slice_of(mx(dist - z_front, 0.0f), s.thickness, s.inv_thickness);
}

// ---- segment drawing ------------------------------------------------------------

template <int MODE, class Sink>
struct Drawer {
    const View& p;
    const float* scene_depth;
    float scene_clear;
    Sink& sink;
    Counters c;
    // The centre-hit test dist <= diameter / 2 of raster.wgsl as
    // dist^2 <= hit_d2 (dist = sqrt(dist^2) rounded); once per drawer.
    float hit_d2 = sqrt_le_bound(mx(p.appearance.w, 0.05f) * 0.5f);

    __device__ __forceinline__ float scene_at(int32_t x, int32_t y) const {
        if (!scene_depth) return scene_clear;
        return scene_depth[(size_t)(uint32_t)y * (uint32_t)to_i32(p.screen.x) + (uint32_t)x];
    }

    // With no scene depth buffer and a zero clear value the scene test of
    // raster.wgsl:93-98 depends only on the sign of the clip depth, which is
    // decided once per sample (no division); otherwise put() tests each pixel.
    __device__ __forceinline__ bool scene_sign_only() const { return !scene_depth && scene_clear == 0.0f; }

    // One camera write of a payload (raster.wgsl:93-115 once the payload is
    // built: it depends only on t, not on the pixel).
    __device__ __forceinline__ void put(int32_t x, int32_t y, int32_t width, float clip_depth, uint64_t payload,
                                        bool center_hit, uint32_t beta_word) {
        if (!scene_sign_only()) {
            float scene = scene_at(x, y);
            if (p.pass_mode[2] == 1u) {
                if (clip_depth < scene) return;
            } else if (clip_depth > scene) {
                return;
            }
        }
        sink.cam((size_t)(uint32_t)y * (uint32_t)width + (uint32_t)x, x, y, payload, center_hit, beta_word, c);
    }

    __device__ void segment(F3 p0, F3 p1, F3 t0, F3 t1, F3 uvw0, F3 uvw1, float ao0, float ao1, float beta) {
        Projected A = project_world(p, p0);
        Projected B = project_world(p, p1);
        if (!A.in_front && !B.in_front) return;
        c.segments++;
        segment_walk(p0, p1, A.screen, B.screen, t0, t1, uvw0, uvw1, ao0, ao1, beta);
    }

    // The DDA of raster.wgsl:77-116 between projected end points
    // a = project(p0).screen and b = project(p1).screen.
    __device__ void segment_walk(F3 p0, F3 p1, F2 a_screen, F2 b_screen, F3 t0, F3 t1, F3 uvw0, F3 uvw1, float ao0,
                                 float ao1, float beta) {
        const struct {
            F2 screen;
        } A{a_screen}, B{b_screen};
        F2 delta = B.screen - A.screen;
        int32_t steps = to_i32(ceilf(mx(fabsf(delta.x), fabsf(delta.y))));
        int32_t n = steps < 1 ? 1 : (steps > 96 ? 96 : steps);
        const int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        F2 seg = B.screen - A.screen;
        float denom = mx(dot2(seg, seg), 1e-6f);
        uint32_t beta_q = to_u32(clampf(beta, 0.0f, 1.0f) * 255.0f + 0.5f);
        // A strand's u and v are the same at both ends: their codes are the
        // ends' own unless a mix could cross a code boundary.
        const bool uv_fixed = uvw0.x == uvw1.x && uvw0.y == uvw1.y && quant63_fixed_under_mix(uvw0.x) &&
                              quant63_fixed_under_mix(uvw0.y);
        const uint32_t qu_end = quant63(uvw0.x), qv_end = quant63(uvw0.y);
        const float fn = i2f(n), inv_n = rdiv(1.0f, fn);  // t = i / n below, exactly (i <= n <= 96)
        const bool sign_only = scene_sign_only();
        int32_t ppx = 0, ppy = 0;
        for (int32_t i = 0; i <= n; ++i) {
            float t = div_by_recip(i2f(i), fn, inv_n);
            F2 screen = mix2(A.screen, B.screen, t);
            int32_t px = to_i32(floorf(screen.x)), py = to_i32(floorf(screen.y));
            bool diag = i > 0 && ppx != px && ppy != py;
            bool in0 = px >= 0 && py >= 0 && px < width && py < height && sink.accepts(px, py);
            bool in1 = diag && ppx >= 0 && py >= 0 && ppx < width && py < height && sink.accepts(ppx, py);
            bool in2 = diag && px >= 0 && ppy >= 0 && px < width && ppy < height && sink.accepts(px, ppy);
            if (in0 || in1 || in2) {
                F3 world = mix3(p0, p1, t);
                float vd = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
                float cz = mat_row(p.clip, 2, world.x, world.y, world.z, 1.0f);
                float cw = mx(mat_row(p.clip, 3, world.x, world.y, world.z, 1.0f), 1e-5f);
                float clip_depth = 0.0f;  // cz / cw, formed only when put() compares it per pixel
                bool visible = vd > p.screen.z;
                if (visible && sign_only)
                    visible = !(p.pass_mode[2] == 1u ? div_below_zero_rare(cz, cw) : div_above_zero_rare(cz, cw));
                else if (visible)
                    clip_depth = rdiv(cz, cw);
                if (visible) {
                    const F3 tangent_raw = mix3(t0, t1, t);
                    uint32_t qu = qu_end, qv = qv_end;
                    if (!uv_fixed) {
                        qu = quant63(mixs(uvw0.x, uvw1.x, t));
                        qv = quant63(mixs(uvw0.y, uvw1.y, t));
                    }
                    const float uvw_z = mixs(uvw0.z, uvw1.z, t);
                    float ao = mixs(ao0, ao1, t);
                    uint32_t depth = quantize_depth(vd, p.screen.z, p.screen.w);
                    uint32_t ox, oy;
                    if (!oct_encode_fast(tangent_raw, ox, oy)) oct_encode(safe_normalize(tangent_raw), ox, oy);
                    uint32_t ao_q = to_u32(clampf(ao, 0.0f, 1.0f) * 63.0f + 0.5f);
                    if (p.flags[2] == 0u) ao_q = 63u;
                    uint64_t payload = pack_gb(depth, ox, oy, qu, qv, quant63(uvw_z), ao_q);
                    uint32_t beta_word = (depth << 8u) | beta_q;
                    if (in0) {
                        F2 center = {i2f(px) + 0.5f, i2f(py) + 0.5f};
                        float closest = clampf(rdiv(dot2(center - A.screen, seg), denom), 0.0f, 1.0f);
                        F2 off = center - (A.screen + seg * closest);
                        bool hit = dot2(off, off) <= hit_d2;
                        put(px, py, width, clip_depth, payload, hit, beta_word);
                    }
                    if (in1) put(ppx, py, width, clip_depth, payload, false, beta_word);
                    if (in2) put(px, ppy, width, clip_depth, payload, false, beta_word);
                }
            }
            ppx = px;
            ppy = py;
        }
    }

    // raster.wgsl:118-148. The sample position (raster.wgsl:176-179) is only
    // formed for pixels that pass the bounds test.
    __device__ __forceinline__ void dom_pixel(int32_t x, int32_t y, F2 a, F2 seg, float seg_len2, F3 p0, F3 p1) {
        int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        if (x < 0 || y < 0 || x >= width || y >= height) return;
        if (!sink.accepts(x, y)) return;
        F2 center = {i2f(x) + 0.5f, i2f(y) + 0.5f};
        float ts = clampf(rdiv(dot2(center - a, seg), seg_len2), 0.0f, 1.0f);
        F3 world = mix3(p0, p1, ts);
        float near_ = p.screen.z, far_ = p.screen.w;
        float dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
        if (dist <= near_ || dist >= far_) return;
        uint32_t dom_w = to_u32(p.dom_info.x);
        size_t pix = (size_t)(uint32_t)y * dom_w + (uint32_t)x;
        uint32_t q = quantize_depth(dist, near_, far_);
        if (MODE == 1) {
            sink.dom_min(pix, x, y, q, dist, c);
            return;
        }
        uint32_t front_q = sink.dom_front(pix, x, y);  // final: the depth pass has finished
        if (front_q == HM_DOM_EMPTY) return;
        const DomSlab slab = dom_slab(near_, far_);
        sink.dom_add(pix, x, y, dom_slice(slab, front_q, dist), c);
    }

    __device__ void dom_segment(F3 p0, F3 p1, bool include_start) {  // raster.wgsl:150-190
        Projected A = project_world(p, p0);
        Projected B = project_world(p, p1);
        if (!A.in_front && !B.in_front) return;
        c.segments++;
        dom_walk(p0, p1, A.screen, B.screen, include_start);
    }

    // The Bresenham walk of raster.wgsl:157-189 between projected end points
    // a = project(p0).screen and b = project(p1).screen.
    __device__ void dom_walk(F3 p0, F3 p1, F2 a, F2 b, bool include_start) {
        int32_t x0 = to_i32(floorf(a.x)), y0 = to_i32(floorf(a.y));
        int32_t x1 = to_i32(floorf(b.x)), y1 = to_i32(floorf(b.y));
        int32_t dx = wabs(wsub(x1, x0)), dy = wabs(wsub(y1, y0));
        int32_t sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
        int32_t err = wsub(dx, dy);
        F2 seg = b - a;
        float seg_len2 = mx(dot2(seg, seg), 1e-6f);
        bool skip = !include_start && (x0 != x1 || y0 != y1);
        int32_t ndy = wsub(0, dy);
        for (int step = 0; step < 1024; ++step) {
            if (!skip) dom_pixel(x0, y0, a, seg, seg_len2, p0, p1);
            skip = false;
            if (x0 == x1 && y0 == y1) break;
            if (sink.walked_out(x0, y0, sx, sy)) break;
            int32_t e2 = wadd(err, err);
            if (e2 > ndy) {
                err = wsub(err, dy);
                x0 = wadd(x0, sx);
            }
            if (e2 < dx) {
                err = wadd(err, dx);
                y0 = wadd(y0, sy);
            }
        }
    }
};

}  // namespace hm
