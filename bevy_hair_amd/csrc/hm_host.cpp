// Host side of the C ABI (include/hair_mi.h): context and buffer ownership,
// HairParams construction and the per-groom dispatch sequence of hair_pass.
//
// Reference (paths relative to the reference repository):
//   GpuGroom::upload   src/render/pass.rs:98-167
//   ensure_targets     src/render/pass.rs:789-815
//   hair_pass          src/render/pass.rs:375-643 (per-groom sequence)
//   make_params        src/render/pass.rs:676-709
//   key_direction      src/render/pass.rs:711-719
//   world_extent       src/render/pass.rs:721-745
//   light_fit          src/render/pass.rs:747-776
//   depth_is_reverse   src/render/pass.rs:778-787
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hm_kernels.h"

#include "hm_ctx.h"

namespace {

hm_status fail(hm_ctx* ctx, hm_status st, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return st;
}

hm_status hip_fail(hm_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? HM_ERR_OOM : HM_ERR_HIP, m);
}

#define HM_HIP(ctx, call)                                              \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);       \
    } while (0)

template <class T> hipError_t dev_alloc(T** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}

template <class T> void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// ---- column-major float mat4 helpers (glam conventions) -------------------------

struct M4 { float m[16]; };
struct V3 { float x, y, z; };

V3 v3(float x, float y, float z) { return {x, y, z}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float length(V3 a) { return std::sqrt(dot(a, a)); }
V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
V3 normalize_or(V3 v, V3 fallback) {  // glam Vec3::normalize_or
    float rcp = 1.0f / length(v);
    if (std::isfinite(rcp) && rcp > 0.0f) return scale(v, rcp);
    return fallback;
}

M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r.m[c * 4 + row] = a.m[0 * 4 + row] * b.m[c * 4 + 0] + a.m[1 * 4 + row] * b.m[c * 4 + 1] +
                               a.m[2 * 4 + row] * b.m[c * 4 + 2] + a.m[3 * 4 + row] * b.m[c * 4 + 3];
    return r;
}

void mul4(const M4& a, const float v[4], float out[4]) {
    for (int row = 0; row < 4; ++row)
        out[row] = a.m[0 * 4 + row] * v[0] + a.m[1 * 4 + row] * v[1] + a.m[2 * 4 + row] * v[2] + a.m[3 * 4 + row] * v[3];
}

V3 transform_point3(const M4& a, V3 p) {
    float v[4] = {p.x, p.y, p.z, 1.0f}, o[4];
    mul4(a, v, o);
    return {o[0], o[1], o[2]};
}

V3 transform_vector3(const M4& a, V3 p) {
    float v[4] = {p.x, p.y, p.z, 0.0f}, o[4];
    mul4(a, v, o);
    return {o[0], o[1], o[2]};
}

// General 4x4 inverse by cofactor expansion.
M4 inverse(const M4& a) {
    const float* m = a.m;
    float inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    M4 r;
    float rdet = 1.0f / det;
    for (int i = 0; i < 16; ++i) r.m[i] = inv[i] * rdet;
    return r;
}

M4 load_m4(const float* f) {
    M4 r;
    std::memcpy(r.m, f, sizeof(r.m));
    return r;
}

void world_corners(const float bmin[3], const float bmax[3], const M4& wfm, V3 out[8]) {  // pass.rs:721-735
    int index = 0;
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz)
                out[index++] = transform_point3(wfm, v3(ix ? bmax[0] : bmin[0], iy ? bmax[1] : bmin[1],
                                                        iz ? bmax[2] : bmin[2]));
}

void world_extent(const float bmin[3], const float bmax[3], const M4& wfm, V3& center, float& radius) {
    // pass.rs:737-745: mean of the 8 corners; radius folds from 1.0.
    V3 c[8];
    world_corners(bmin, bmax, wfm, c);
    V3 sum = v3(0, 0, 0);
    for (int i = 0; i < 8; ++i) sum = add(sum, c[i]);
    center = scale(sum, 1.0f / 8.0f);
    radius = 1.0f;
    for (int i = 0; i < 8; ++i) radius = std::fmax(radius, length(sub(c[i], center)));
}

// Transform::looking_to(forward, up) then to_matrix().inverse(): the view
// matrix whose -Z axis is `forward` (bevy_transform; glam look_to_rh).
M4 look_to_view(V3 eye, V3 forward, V3 up) {
    V3 back = scale(normalize_or(forward, v3(0, 0, -1)), -1.0f);
    V3 right = normalize_or(cross(up, back), v3(1, 0, 0));
    V3 upv = cross(back, right);
    M4 v;
    // rows of the rotation are right, up, back; translation -R * eye
    v.m[0] = right.x; v.m[4] = right.y; v.m[8] = right.z; v.m[12] = -dot(right, eye);
    v.m[1] = upv.x;   v.m[5] = upv.y;   v.m[9] = upv.z;   v.m[13] = -dot(upv, eye);
    v.m[2] = back.x;  v.m[6] = back.y;  v.m[10] = back.z; v.m[14] = -dot(back, eye);
    v.m[3] = 0;       v.m[7] = 0;       v.m[11] = 0;      v.m[15] = 1;
    return v;
}

// Right-handed orthographic projection with depth in [0, 1] (glam
// Mat4::orthographic_rh). Only its x, y and w rows reach the hair path.
M4 orthographic(float l, float r, float b, float t, float n, float f) {
    float rw = 1.0f / (r - l), rh = 1.0f / (t - b), rd = 1.0f / (n - f);
    M4 m;
    std::memset(m.m, 0, sizeof(m.m));
    m.m[0] = rw + rw;
    m.m[5] = rh + rh;
    m.m[10] = rd;
    m.m[12] = -(l + r) * rw;
    m.m[13] = -(t + b) * rh;
    m.m[14] = rd * n;
    m.m[15] = 1.0f;
    return m;
}

void light_fit(const float bmin[3], const float bmax[3], const M4& wfm, V3 direction_to_light, V3& eye, V3& fwd,
               M4& clip, float& near_out, float& far_out) {  // pass.rs:747-776
    V3 center;
    float radius;
    world_extent(bmin, bmax, wfm, center, radius);
    V3 to_light = normalize_or(direction_to_light, v3(0, 1, 0));
    fwd = scale(to_light, -1.0f);
    eye = add(center, scale(to_light, radius));
    float near_ = 3.402823466e38f, far_ = -3.402823466e38f;
    V3 c[8];
    world_corners(bmin, bmax, wfm, c);
    for (int i = 0; i < 8; ++i) {
        float depth = dot(sub(c[i], eye), fwd);
        near_ = std::fmin(near_, depth);
        far_ = std::fmax(far_, depth);
    }
    near_ = std::fmax(near_ - radius * 0.05f, 0.05f);
    far_ = std::fmax(far_ + radius * 0.05f, near_ + 1.0f);
    V3 up = std::fabs(dot(fwd, v3(0, 1, 0))) > 0.95f ? v3(0, 0, 1) : v3(0, 1, 0);
    M4 view = look_to_view(eye, fwd, up);
    M4 proj = orthographic(-radius, radius, -radius, radius, near_, far_);
    clip = mul(proj, view);
    near_out = near_;
    far_out = far_;
}

float clip_z(const M4& clip, V3 p) {  // pass.rs:784-787
    float v[4] = {p.x, p.y, p.z, 1.0f}, c[4];
    mul4(clip, v, c);
    return c[2] / std::fmax(c[3], 1e-5f);
}

struct ViewArgs {
    M4 clip_from_world, world_from_clip, world_from_model, light_clip;
    V3 camera_pos, camera_forward;
    float screen[4];
    V3 light_forward;
    float dom_near, dom_far, dom_scale;
    uint32_t pass_mode[4];
    V3 light_eye;
    uint32_t dom_size;
};

void make_params(const ViewArgs& a, const hm_groom_config& g, hm_params& out) {  // pass.rs:676-709
    std::memset(&out, 0, sizeof(out));
    std::memcpy(out.clip_from_world, a.clip_from_world.m, 64);
    std::memcpy(out.world_from_clip, a.world_from_clip.m, 64);
    std::memcpy(out.world_from_model, a.world_from_model.m, 64);
    std::memcpy(out.light_clip_from_world, a.light_clip.m, 64);
    float cp[4] = {a.camera_pos.x, a.camera_pos.y, a.camera_pos.z, 1.0f};
    float cf[4] = {a.camera_forward.x, a.camera_forward.y, a.camera_forward.z, 0.0f};
    std::memcpy(out.camera_pos, cp, 16);
    std::memcpy(out.camera_forward, cf, 16);
    std::memcpy(out.screen, a.screen, 16);
    float app[4] = {g.roughness, g.tilt, g.lambda, g.center_diameter};
    std::memcpy(out.appearance, app, 16);
    float fp[4] = {1.0f, 0.2f, std::fmax(a.dom_scale, 1e-3f), HM_DOM_SIGMA};
    std::memcpy(out.filter_params, fp, 16);
    float alb[4] = {g.albedo_linear[0], g.albedo_linear[1], g.albedo_linear[2], 1.0f};
    std::memcpy(out.albedo, alb, 16);
    uint32_t flags[4] = {g.lod ? 1u : 0u, g.filter ? 1u : 0u, g.ambient_occlusion ? 1u : 0u, g.deep_opacity ? 1u : 0u};
    std::memcpy(out.flags, flags, 16);
    std::memcpy(out.pass_mode, a.pass_mode, 16);
    float le[4] = {a.light_eye.x, a.light_eye.y, a.light_eye.z, 1.0f};
    float lf[4] = {a.light_forward.x, a.light_forward.y, a.light_forward.z, 0.0f};
    std::memcpy(out.light_eye, le, 16);
    std::memcpy(out.light_forward, lf, 16);
    float di[4] = {(float)a.dom_size, (float)a.dom_size, a.dom_near, a.dom_far};
    std::memcpy(out.dom_info, di, 16);
}

void free_targets(hm_ctx* c) {
    dev_free(c->gb_slab);
    c->center = nullptr;
    c->cons = nullptr;
    c->beta = nullptr;
    dev_free(c->shaded);
    dev_free(c->shaded_depth);
    dev_free(c->filtered);
    dev_free(c->filtered_depth);
    dev_free(c->dom);
    dev_free(c->pix_list);
    dev_free(c->pix_list2);
    dev_free(c->pix_count);
    dev_free(c->scene_depth);
    c->width = c->height = c->dom_size = 0;
    c->cap_gb_words = c->cap_px = c->cap_dom_words = 0;
    c->list_len = 0;
}

void free_bin(hm_ctx* c) {
    hm_bin_state& b = c->bin;
    dev_free(b.cps);
    dev_free(b.cpy);
    dev_free(b.dcp);
    dev_free(b.cpt);
    dev_free(b.sinfo);
    for (int pass = 0; pass < 2; ++pass) {
        dev_free(b.meta[pass]);
        dev_free(b.seg_box[pass]);
        dev_free(b.long_list[pass]);
        dev_free(b.records[pass]);
        dev_free(b.lists[pass]);
        dev_free(b.wg_list[pass]);
    }
    dev_free(b.chunk_tile);
    b = hm_bin_state();
}

// Grows the binning buffers to hold `slots` strand slots, `meta_words`
// status + per-tile words of pass `pass`, `records` tile records, `lists`
// (tile, rank) entries, `wgs` binning workgroups and `chunks` tile-workgroup
// chunks. Only the shared control-point buffers and pass `pass`'s own
// buffers are touched; before any of them is replaced, both streams are
// drained (an earlier frame, or the other pass of this one, may still read
// them).
hipError_t ensure_bin(hm_ctx* c, int pass, size_t slots, size_t meta_words, size_t records, size_t lists,
                      size_t wgs, size_t chunks) {
    hm_bin_state& b = c->bin;
    hipError_t e;
    if (b.meta_dirty[pass] && b.meta[pass]) {
        // (a fill kernel rather than a memset: this may run under graph capture)
        if ((e = hm::launch_fill32(c->stream, b.meta[pass], b.meta_words[pass], 0u, 1)) != hipSuccess) return e;
        b.meta_dirty[pass] = false;
    }
    const bool chunks_ok = pass != HM_PASS_CAMERA || chunks <= b.chunk_capacity;
    if (slots <= b.slots && slots <= b.seg_slots[pass] && meta_words <= b.meta_words[pass] &&
        records <= b.record_capacity[pass] && lists <= b.list_capacity[pass] && wgs <= b.wg_capacity[pass] && chunks_ok)
        return hipSuccess;
    if ((e = hipStreamSynchronize(c->main)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(c->side)) != hipSuccess) return e;
    c->alloc_gen++;
    if (slots > b.slots) {
        dev_free(b.cps);
        dev_free(b.cpy);
        dev_free(b.dcp);
        dev_free(b.cpt);
        dev_free(b.sinfo);
        b.slots = 0;
        size_t segs = slots * hm::BIN_STRIDE;
        if ((e = dev_alloc(&b.cps, segs * sizeof(float4))) != hipSuccess) return e;
        if ((e = dev_alloc(&b.cpy, segs * sizeof(float))) != hipSuccess) return e;
        if ((e = dev_alloc(&b.dcp, segs * 5 * sizeof(float))) != hipSuccess) return e;
        if ((e = dev_alloc(&b.cpt, segs * sizeof(float4))) != hipSuccess) return e;
        if ((e = dev_alloc(&b.sinfo, slots * sizeof(float4))) != hipSuccess) return e;
        b.slots = slots;
    }
    if (slots > b.seg_slots[pass]) {
        dev_free(b.seg_box[pass]);
        dev_free(b.long_list[pass]);
        b.seg_slots[pass] = 0;
        size_t segs = slots * hm::BIN_STRIDE;
        if ((e = dev_alloc(&b.seg_box[pass], segs * 4)) != hipSuccess) return e;
        if ((e = dev_alloc(&b.long_list[pass], segs * 4)) != hipSuccess) return e;
        b.seg_slots[pass] = slots;
    }
    if (meta_words > b.meta_words[pass]) {
        dev_free(b.meta[pass]);
        b.meta_words[pass] = 0;
        if ((e = dev_alloc(&b.meta[pass], meta_words * 4)) != hipSuccess) return e;
        // k_bin_gen expects the long-segment cursor at zero (k_bin_scan
        // re-zeroes it after every pass).
        if ((e = hm::launch_fill32(c->stream, b.meta[pass], meta_words, 0u, 1)) != hipSuccess) return e;
        b.meta_words[pass] = meta_words;
        b.meta_dirty[pass] = false;
    }
    if (records > b.record_capacity[pass]) {
        dev_free(b.records[pass]);
        b.record_capacity[pass] = 0;
        if ((e = dev_alloc(&b.records[pass], records * 4)) != hipSuccess) return e;
        b.record_capacity[pass] = records;
    }
    if (lists > b.list_capacity[pass]) {
        dev_free(b.lists[pass]);
        b.list_capacity[pass] = 0;
        if ((e = dev_alloc(&b.lists[pass], lists * sizeof(uint2))) != hipSuccess) return e;
        b.list_capacity[pass] = lists;
    }
    if (wgs > b.wg_capacity[pass]) {
        dev_free(b.wg_list[pass]);
        b.wg_capacity[pass] = 0;
        if ((e = dev_alloc(&b.wg_list[pass], wgs * sizeof(uint2))) != hipSuccess) return e;
        b.wg_capacity[pass] = wgs;
    }
    if (!chunks_ok) {
        dev_free(b.chunk_tile);
        b.chunk_capacity = 0;
        if ((e = dev_alloc(&b.chunk_tile, chunks * 4)) != hipSuccess) return e;
        b.chunk_capacity = chunks;
    }
    return hipSuccess;
}

// Binning grid of a pass: workgroups per CU times CUs. The camera pass's
// tile histogram is small (one LDS word per 32x32 tile), so it takes its own,
// larger count; the DOM pass's (one word per 16x16 texel tile) limits how many
// fit on a CU.
static int bin_grid(const hm_ctx* c, int pass) {
    const uint32_t cam = c->tuning.cam_bin_blocks_per_cu;
    const uint32_t per_cu = pass == HM_PASS_CAMERA && cam ? cam : c->tuning.bin_blocks_per_cu;
    return c->cu_count * (int)per_cu;
}

int raster_blocks(const hm_ctx* c, uint32_t capacity, uint32_t mode) {
    uint32_t want = (capacity + 3u) / 4u;
    uint32_t bpc = c->tuning.raster_blocks_per_cu[mode < 3u ? mode : 0u];
    uint32_t cap = (uint32_t)c->cu_count * (bpc ? bpc : 16u);
    return (int)(want < cap ? (want ? want : 1u) : cap);
}

// Records an event pair around one kernel into the current frame slot.
struct Timer {
    hm_ctx* c;
    size_t base = 0;
    bool on = false;
    Timer(hm_ctx* ctx, int kernel) : c(ctx) {
        uint32_t slot = c->frame;
        on = c->slots && slot < c->slots;
        if (!on) return;
        base = ((size_t)slot * HM_K_COUNT + (size_t)kernel) * 2;
        (void)hipEventRecord(c->ev[base], c->stream);
    }
    ~Timer() {
        if (!on) return;
        (void)hipEventRecord(c->ev[base + 1], c->stream);
        c->ran[base / 2] = 1;
    }
};

// The direct kernels run for `pass`: its status words say so (records, binned
// and fallback all zero), so that an earlier binned run's status is never read
// as this call's. The long-segment cursor and the sub-list counters after it
// are left alone (the cursor is zero between passes).
hipError_t mark_direct(hm_ctx* c, int pass) {
    if (!c->bin.meta[pass]) return hipSuccess;
    return hm::launch_fill32(c->stream, c->bin.meta[pass], hm::BIN_LONG_APPEND, 0u, 1);
}

hm_status check_targets(hm_ctx* ctx) {
    if (!ctx) return HM_ERR_INVALID;
    if (!ctx->gb_slab) return fail(ctx, HM_ERR_NOT_READY, "targets not allocated (hm_targets)");
    return HM_OK;
}

hm_status check_pass_params(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass) {
    if (!ctx || !g || !p || (pass != HM_PASS_CAMERA && pass != HM_PASS_DOM))
        return fail(ctx, HM_ERR_INVALID, "null handle or bad pass");
    if (p->pass_mode[1] > g->nb) return fail(ctx, HM_ERR_INVALID, "pass_mode.y exceeds the groom's bundle count");
    return HM_OK;
}

}  // namespace

hm_status hm_fail(hm_ctx* ctx, hm_status st, const std::string& msg) { return fail(ctx, st, msg); }
hm_status hm_hip_fail(hm_ctx* ctx, hipError_t e, const char* what) { return hip_fail(ctx, e, what); }

// ---- ABI ------------------------------------------------------------------------

extern "C" {

const char* hm_status_string(hm_status s) {
    switch (s) {
        case HM_OK: return "ok";
        case HM_ERR_INVALID: return "invalid argument";
        case HM_ERR_OOM: return "out of device memory";
        case HM_ERR_HIP: return "HIP runtime error";
        case HM_ERR_NO_DEVICE: return "no usable gfx950 device";
        case HM_ERR_NOT_READY: return "not ready";
        case HM_ERR_CAPACITY: return "capacity exceeded";
    }
    return "unknown status";
}

const char* hm_last_error(const hm_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

hm_status hm_create(int device, hm_ctx** out) {
    if (!out) return HM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return HM_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HM_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HM_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return HM_ERR_NO_DEVICE;
    if (hm_noalloc_setup(device) != hipSuccess) return HM_ERR_HIP;
    hm_ctx* c = new hm_ctx();
    c->device = device;
    c->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&c->main, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return HM_ERR_HIP;
    }
    c->stream = c->main;
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
        hm_destroy(c);
        return HM_ERR_HIP;
    }
    // The fork and join events only order the context's two streams on one
    // device and the host never waits on them, so they leave out the
    // system-scope fence an event record makes by default.
    for (int j = 0; j < 3; ++j) {
        if (hipEventCreateWithFlags(&c->fork_ev[j], hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
            hm_destroy(c);
            return HM_ERR_HIP;
        }
    }
    (void)hipEventCreate(&c->frame_ev[0]);
    (void)hipEventCreate(&c->frame_ev[1]);
    if (dev_alloc(&c->counters, 64) != hipSuccess) {
        hm_destroy(c);
        return HM_ERR_OOM;
    }
    *out = c;
    return HM_OK;
}

void hm_destroy(hm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->stream = c->main;
    if (c->main) (void)hipStreamSynchronize(c->main);
    if (c->side) (void)hipStreamSynchronize(c->side);
    free_targets(c);
    free_bin(c);
    hm_mlp_release(c);
    if (c->borrowed) {  // the parent's buffers
        c->points = nullptr;
        c->dirs = nullptr;
        c->rects = nullptr;
        c->env_maps = nullptr;
    }
    dev_free(c->points);
    dev_free(c->dirs);
    dev_free(c->rects);
    dev_free(c->env_maps);
    dev_free(c->counters);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (int j = 0; j < 2; ++j)
        if (c->frame_ev[j]) (void)hipEventDestroy(c->frame_ev[j]);
    for (hipEvent_t e : c->fork_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->main) (void)hipStreamDestroy(c->main);
    delete c;
}

}  // extern "C"

hm_status hm_groom_shadow(hm_ctx* ctx, const hm_groom* g, hm_groom** out) {
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    hm_groom* s = new hm_groom();
    s->borrowed = true;
    s->nb = g->nb;
    s->nl = g->nl;
    s->ns = g->ns;
    s->strands = g->strands;
    s->capacity = g->capacity;
    std::memcpy(s->bmin, g->bmin, 12);
    std::memcpy(s->bmax, g->bmax, 12);
    s->bundles = g->bundles;
    s->layers = g->layers;
    s->style = g->style;
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) e = dev_alloc(p, bytes);
    };
    for (int k = 0; k < 2; ++k) {
        A(&s->lod[k], (size_t)s->nb * sizeof(hm_lod_record));
        A(&s->refs[k], (size_t)s->capacity * sizeof(hm_strand_ref));
        A(&s->prefix[k], ((size_t)s->nb + 1) * 4);
        A(&s->indirect[k], 16);
    }
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMemset(s->indirect[k], 0, 16);
    if (e != hipSuccess) {
        hm_groom_free(nullptr, s);
        return hip_fail(ctx, e, "batch lane groom");
    }
    *out = s;
    return HM_OK;
}

hm_status hm_lane_create(hm_ctx* parent, hm_ctx** out) {
    *out = nullptr;
    hm_ctx* c = nullptr;
    hm_status st = hm_create(parent->device, &c);
    if (st != HM_OK) return fail(parent, st, "batch lane: context creation failed");
    c->borrowed = true;
    c->tuning = parent->tuning;
    c->points = parent->points;
    c->point_count = parent->point_count;
    c->point_cap = parent->point_cap;
    c->spot_first = parent->spot_first;
    c->dirs = parent->dirs;
    c->dir_count = parent->dir_count;
    c->dir_cap = parent->dir_cap;
    c->host_points = parent->host_points;
    c->light_gen = parent->light_gen;
    c->rects = parent->rects;
    c->rect_count = parent->rect_count;
    c->rect_cap = parent->rect_cap;
    c->env_maps = parent->env_maps;
    c->env_bytes = parent->env_bytes;
    c->env = parent->env;
    c->shade_gen = parent->shade_gen;
    c->scene_clear = parent->scene_clear;
    if ((st = hm_mlp_borrow(c, parent)) != HM_OK) {
        parent->err = "batch lane: " + c->err;
        hm_destroy(c);
        return st;
    }
    *out = c;
    return HM_OK;
}

extern "C" {

hm_status hm_groom_upload(hm_ctx* ctx, const hm_groom_desc* d, hm_groom** out) {
    if (!ctx || !d || !out) return fail(ctx, HM_ERR_INVALID, "null argument");
    *out = nullptr;
    if ((d->bundle_count && !d->bundles) || (d->layer_count && !d->layers) || (d->style_count && !d->style))
        return fail(ctx, HM_ERR_INVALID, "null array with nonzero count");
    // Every bundle must address 16x4 corners and 8x8x16 texels inside the arrays.
    uint64_t strands = 0;
    for (uint32_t b = 0; b < d->bundle_count; ++b) {
        const hm_bundle& B = d->bundles[b];
        if ((uint64_t)B.layer_offset + HM_LAYER_COUNT * 4u > d->layer_count ||
            (uint64_t)B.style_offset + HM_STYLE_TEXELS > d->style_count)
            return fail(ctx, HM_ERR_INVALID, "bundle " + std::to_string(b) + " addresses outside the cage/style arrays");
        strands += B.strand_count;
    }
    if (strands > 0x7fffffffull) return fail(ctx, HM_ERR_CAPACITY, "more than 2^31 strands");
    (void)hipSetDevice(ctx->device);
    hm_groom* g = new hm_groom();
    // An empty mesh keeps one zero bundle, like the 16-byte placeholder
    // buffers of GpuGroom::upload (pass.rs:118-119, pass.rs:170-175).
    g->nb = d->bundle_count ? d->bundle_count : 1u;
    g->nl = d->bundle_count ? d->layer_count : HM_LAYER_COUNT * 4u;
    g->ns = d->bundle_count ? d->style_count : HM_STYLE_TEXELS;
    g->strands = (uint32_t)strands;
    g->capacity = g->strands ? g->strands : 1u;
    std::memcpy(g->bmin, d->bounds_min, 12);
    std::memcpy(g->bmax, d->bounds_max, 12);
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) e = dev_alloc(p, bytes);
    };
    A(&g->bundles, (size_t)g->nb * sizeof(hm_bundle));
    A(&g->layers, (size_t)g->nl * sizeof(hm_layer));
    A(&g->style, (size_t)g->ns * 16);
    for (int k = 0; k < 2; ++k) {
        A(&g->lod[k], (size_t)g->nb * sizeof(hm_lod_record));
        A(&g->refs[k], (size_t)g->capacity * sizeof(hm_strand_ref));
        A(&g->prefix[k], ((size_t)g->nb + 1) * 4);
        A(&g->indirect[k], 16);
    }
    if (e == hipSuccess) {
        if (d->bundle_count) {
            e = hipMemcpy(g->bundles, d->bundles, (size_t)g->nb * sizeof(hm_bundle), hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(g->layers, d->layers, (size_t)g->nl * sizeof(hm_layer), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(g->style, d->style, (size_t)g->ns * 16, hipMemcpyHostToDevice);
        } else {
            e = hipMemset(g->bundles, 0, sizeof(hm_bundle));
            if (e == hipSuccess) e = hipMemset(g->layers, 0, (size_t)g->nl * sizeof(hm_layer));
            if (e == hipSuccess) e = hipMemset(g->style, 0, (size_t)g->ns * 16);
        }
    }
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMemset(g->indirect[k], 0, 16);
    if (e != hipSuccess) {
        hm_groom_free(ctx, g);
        return hip_fail(ctx, e, "hm_groom_upload");
    }
    *out = g;
    return HM_OK;
}

void hm_groom_free(hm_ctx* ctx, hm_groom* g) {
    if (!g) return;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (g->borrowed) {  // the shadowed groom's buffers
        g->bundles = nullptr;
        g->layers = nullptr;
        g->style = nullptr;
    }
    dev_free(g->bundles);
    dev_free(g->layers);
    dev_free(g->style);
    for (int k = 0; k < 2; ++k) {
        dev_free(g->lod[k]);
        dev_free(g->refs[k]);
        dev_free(g->prefix[k]);
        dev_free(g->indirect[k]);
    }
    delete g;
}

uint32_t hm_groom_bundle_count(const hm_groom* g) { return g ? g->nb : 0u; }
uint32_t hm_groom_strand_count(const hm_groom* g) { return g ? g->strands : 0u; }

hm_status hm_targets(hm_ctx* ctx, uint32_t width, uint32_t height, uint32_t dom_size) {
    if (!ctx) return HM_ERR_INVALID;
    if (width == 0 || height == 0) return fail(ctx, HM_ERR_NOT_READY, "zero-size target");  // pass.rs:320-322
    if (dom_size == 0 || (dom_size & 1u) || dom_size > 16384u)
        return fail(ctx, HM_ERR_INVALID, "dom_size must be even and in [2, 16384]");
    if ((uint64_t)width * height > (1ull << 31)) return fail(ctx, HM_ERR_CAPACITY, "target larger than 2^31 pixels");
    if (ctx->gb_slab && ctx->width == width && ctx->height == height && ctx->dom_size == dom_size) return HM_OK;
    (void)hipSetDevice(ctx->device);
    const size_t px = (size_t)width * height;
    const size_t dom_words = (size_t)dom_size * dom_size * (1 + HM_DOM_LAYERS);
    // center (8 B) | conservative (8 B) | beta (4 B, padded to 16 B)
    const size_t beta_words = (px + 3) & ~(size_t)3;
    const size_t gb_words = px * 2 + px * 2 + beta_words;
    // Covered-pixel sub-lists: room for every pixel of every (partial) tile
    // up to 32 x 32 plus one tile's slack per sub-list.
    const size_t list_len = ((size_t)width + 32u) * ((size_t)height + 32u) + (size_t)hm::LIST_SUBS * 1024u;
    const bool fits = ctx->gb_slab && gb_words <= ctx->cap_gb_words && px <= ctx->cap_px &&
                      dom_words <= ctx->cap_dom_words && list_len <= ctx->list_len;
    if (!fits) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->alloc_gen++;
        float keep_clear = ctx->scene_clear;
        const size_t c_gb = std::max(gb_words, ctx->cap_gb_words), c_px = std::max(px, ctx->cap_px);
        const size_t c_dom = std::max(dom_words, ctx->cap_dom_words), c_list = std::max(list_len, ctx->list_len);
        free_targets(ctx);
        hipError_t e = dev_alloc(&ctx->gb_slab, c_gb * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->shaded, c_px * 16);
        if (e == hipSuccess) e = dev_alloc(&ctx->shaded_depth, c_px * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->filtered, c_px * 16);
        if (e == hipSuccess) e = dev_alloc(&ctx->filtered_depth, c_px * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->dom, c_dom * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->pix_list, c_list * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->pix_list2, c_list * 4);
        if (e == hipSuccess) e = dev_alloc(&ctx->pix_count, 2 * hm::LIST_SUBS * 4);
        if (e == hipSuccess) e = hipMemset(ctx->dom, 0, c_dom * 4);
        if (e != hipSuccess) {
            free_targets(ctx);
            return hip_fail(ctx, e, "hm_targets");
        }
        ctx->cap_gb_words = c_gb;
        ctx->cap_px = c_px;
        ctx->cap_dom_words = c_dom;
        ctx->list_len = c_list;
        ctx->scene_clear = keep_clear;
    } else if (ctx->scene_depth && (width != ctx->width || height != ctx->height)) {
        // the scene depth belongs to the previous size
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx->scene_depth);
        ctx->alloc_gen++;
    }
    if (width != ctx->width || height != ctx->height) ctx->scene_depth_ext = nullptr;  // bound for the old size
    uint32_t* slab = static_cast<uint32_t*>(ctx->gb_slab);
    ctx->center = reinterpret_cast<uint64_t*>(slab);
    ctx->cons = reinterpret_cast<uint64_t*>(slab + px * 2);
    ctx->beta = slab + px * 4;
    ctx->gb_words = gb_words;
    ctx->width = width;
    ctx->height = height;
    ctx->dom_size = dom_size;
    return HM_OK;
}

hm_status hm_set_scene_depth(hm_ctx* ctx, const float* host_depth, float clear_value) {
    if (!ctx) return HM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    ctx->scene_depth_ext = nullptr;
    ctx->scene_clear = clear_value;
    if (!host_depth) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->scene_depth) ctx->alloc_gen++;
        dev_free(ctx->scene_depth);
        return HM_OK;
    }
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    size_t bytes = (size_t)ctx->width * ctx->height * 4;
    if (!ctx->scene_depth) {
        ctx->alloc_gen++;
        HM_HIP(ctx, dev_alloc(&ctx->scene_depth, bytes));
    }
    HM_HIP(ctx, hipMemcpyAsync(ctx->scene_depth, host_depth, bytes, hipMemcpyHostToDevice, ctx->stream));
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

hm_status hm_bind_scene_depth(hm_ctx* ctx, const float* device_depth, float clear_value) {
    if (!ctx) return HM_ERR_INVALID;
    if (!device_depth) return hm_set_scene_depth(ctx, nullptr, clear_value);
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    ctx->scene_depth_ext = device_depth;
    ctx->scene_clear = clear_value;
    return HM_OK;
}

hm_status hm_set_lights(hm_ctx* ctx, const hm_point_light* points, uint32_t np, const hm_directional_light* dirs,
                        uint32_t nd) {
    if (!ctx || (np && !points) || (nd && !dirs)) return fail(ctx, HM_ERR_INVALID, "null light array");
    (void)hipSetDevice(ctx->device);
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (np > ctx->point_cap) {
        dev_free(ctx->points);
        ctx->point_cap = 0;
        ctx->alloc_gen++;
        HM_HIP(ctx, dev_alloc(&ctx->points, (size_t)np * sizeof(hm_point_light)));
        ctx->point_cap = np;
    }
    if (nd > ctx->dir_cap) {
        dev_free(ctx->dirs);
        ctx->dir_cap = 0;
        ctx->alloc_gen++;
        HM_HIP(ctx, dev_alloc(&ctx->dirs, (size_t)nd * sizeof(hm_directional_light)));
        ctx->dir_cap = nd;
    }
    // Point lights first, then spot lights, each kind in the caller's order:
    // the shading loops walk each kind as one range, in the order of Bevy's
    // cluster ranges (shade.wesl:452-480).
    std::vector<hm_point_light> sorted;
    sorted.reserve(np);
    for (uint32_t i = 0; i < np; ++i)
        if (points[i].kind == HM_LIGHT_POINT) sorted.push_back(points[i]);
    const uint32_t spot_first = (uint32_t)sorted.size();
    for (uint32_t i = 0; i < np; ++i)
        if (points[i].kind != HM_LIGHT_POINT) sorted.push_back(points[i]);
    if (np) HM_HIP(ctx, hipMemcpy(ctx->points, sorted.data(), (size_t)np * sizeof(hm_point_light), hipMemcpyHostToDevice));
    if (nd) HM_HIP(ctx, hipMemcpy(ctx->dirs, dirs, (size_t)nd * sizeof(hm_directional_light), hipMemcpyHostToDevice));
    ctx->point_count = np;
    ctx->spot_first = spot_first;
    ctx->dir_count = nd;
    ctx->host_points = std::move(sorted);
    ctx->light_gen++;
    return HM_OK;
}

// Whether culling point and spot lights per wave can pay. The culling costs
// a fixed few hundred instructions per wave and saves the range test of
// each light that reaches none of the wave's pixels, so it pays when, over
// the groom's box (grown by 5 %), the lights miss a good share of it: the
// share of a 3 x 3 x 3 grid of points in the box beyond each light's range,
// summed over the lights, must reach 16. Without a box (shading called on
// its own), from 64 lights on. Either choice gives the same colours.
static bool cull_pays(hm_ctx* ctx) {
    if (ctx->point_count == 0 || ctx->tuning.light_cull == 1u) return false;
    if (ctx->tuning.light_cull == 2u) return true;
    if (!ctx->groom_box_set) return ctx->point_count >= 64u;
    hm_cull_cache& c = ctx->cull_cache;
    if (c.valid && c.light_gen == ctx->light_gen && std::memcmp(c.box, ctx->groom_box, sizeof(c.box)) == 0)
        return c.cull;
    const float* b = ctx->groom_box;
    float missed = 0.0f;
    for (const hm_point_light& L : ctx->host_points) {
        const float inv = L.color_inverse_square_range[3];
        if (!(inv > 0.0f)) continue;  // unbounded range
        int outside = 0;
        for (int i = 0; i < 27; ++i) {
            float d2 = 0.0f;
            for (int k = 0, q = i; k < 3; ++k, q /= 3) {
                const float x = b[k] + 0.5f * (float)(q % 3) * (b[3 + k] - b[k]) - L.position_radius[k];
                d2 += x * x;
            }
            outside += d2 * inv >= 1.0f;
        }
        missed += (float)outside / 27.0f;
    }
    c.valid = true;
    c.light_gen = ctx->light_gen;
    std::memcpy(c.box, ctx->groom_box, sizeof(c.box));
    c.cull = missed >= 16.0f;
    return c.cull;
}

// The groom's world box for cull_pays while hm_frame shades.
struct GroomBox {
    hm_ctx* ctx;
    GroomBox(hm_ctx* c, const hm_groom* g, const hm_params& cam) : ctx(c) {
        const float* m = cam.world_from_model;  // column-major
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int corner = 0; corner < 8; ++corner) {
            const float x = corner & 1 ? g->bmax[0] : g->bmin[0];
            const float y = corner & 2 ? g->bmax[1] : g->bmin[1];
            const float z = corner & 4 ? g->bmax[2] : g->bmin[2];
            for (int k = 0; k < 3; ++k) {
                const float w = m[k] * x + m[4 + k] * y + m[8 + k] * z + m[12 + k];
                lo[k] = std::min(lo[k], w);
                hi[k] = std::max(hi[k], w);
            }
        }
        for (int k = 0; k < 3; ++k) {
            const float grow = 0.05f * (hi[k] - lo[k]) + 1e-4f;
            ctx->groom_box[k] = lo[k] - grow;
            ctx->groom_box[3 + k] = hi[k] + grow;
        }
        ctx->groom_box_set = std::isfinite(lo[0] + lo[1] + lo[2] + hi[0] + hi[1] + hi[2]);
    }
    ~GroomBox() { ctx->groom_box_set = false; }
};

extern "C++" void hm_shade_lights(hm_ctx* ctx, hm::ShadeArgs& a) {
    a.cull = cull_pays(ctx) ? 1u : 0u;
    a.points = ctx->points;
    a.point_count = ctx->point_count;
    a.spot_first = ctx->spot_first;
    a.directionals = ctx->dirs;
    a.directional_count = ctx->dir_count;
    a.rects = ctx->rects;
    a.rect_count = ctx->rect_count;
    a.env = ctx->env;
}

hm_status hm_set_rect_lights(hm_ctx* ctx, const hm_rect_light* rects, uint32_t count) {
    static_assert(sizeof(hm_rect_light) == 80, "hm_rect_light layout");
    if (!ctx || (count && !rects)) return fail(ctx, HM_ERR_INVALID, "null rect light array");
    (void)hipSetDevice(ctx->device);
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (count > ctx->rect_cap) {
        dev_free(ctx->rects);
        ctx->rect_cap = 0;
        ctx->rect_count = 0;
        ctx->alloc_gen++;
        HM_HIP(ctx, dev_alloc(&ctx->rects, (size_t)count * sizeof(hm_rect_light)));
        ctx->rect_cap = count;
    }
    if (count) HM_HIP(ctx, hipMemcpy(ctx->rects, rects, (size_t)count * sizeof(hm_rect_light), hipMemcpyHostToDevice));
    if (count != ctx->rect_count) ctx->shade_gen++;
    ctx->rect_count = count;
    return HM_OK;
}

hm_status hm_set_env_probe(hm_ctx* ctx, const hm_env_probe* probe) {
    if (!ctx) return HM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->shade_gen++;
    if (!probe) {
        ctx->env.on = 0u;
        return HM_OK;
    }
    const uint32_t S = probe->specular_size, M = probe->specular_mips, D = probe->diffuse_size;
    uint32_t levels = 1;
    while ((S >> levels) != 0u) ++levels;  // floor(log2 S) + 1
    if (!probe->specular || !probe->diffuse || S == 0u || S > 16384u || D == 0u || D > 16384u || M == 0u ||
        M > hm::ENV_MAX_MIPS || M > levels)
        return fail(ctx, HM_ERR_INVALID,
                    "hm_set_env_probe: cube maps must be non-null, sizes in [1, 16384], 1 <= mips <= "
                    "min(16, log2(size) + 1)");
    hm::EnvArgs e = {};
    size_t texels = 0;
    for (uint32_t m = 0; m < M; ++m) {
        e.spec_base[m] = (uint32_t)texels;
        const size_t sm = S >> m;
        texels += 6u * sm * sm;
    }
    const size_t spec_texels = texels, diff_texels = 6u * (size_t)D * D;
    if (spec_texels + diff_texels > 0xffffffffull)
        return fail(ctx, HM_ERR_INVALID, "hm_set_env_probe: cube maps too large");
    const size_t bytes = (spec_texels + diff_texels) * 16u;
    if (bytes > ctx->env_bytes) {
        ctx->env.on = 0u;
        dev_free(ctx->env_maps);
        ctx->env_bytes = 0;
        ctx->alloc_gen++;
        HM_HIP(ctx, dev_alloc(&ctx->env_maps, bytes));
        ctx->env_bytes = bytes;
    }
    HM_HIP(ctx, hipMemcpy(ctx->env_maps, probe->specular, spec_texels * 16u, hipMemcpyHostToDevice));
    HM_HIP(ctx, hipMemcpy(ctx->env_maps + spec_texels * 4u, probe->diffuse, diff_texels * 16u,
                          hipMemcpyHostToDevice));
    e.specular = reinterpret_cast<const float4*>(ctx->env_maps);
    e.diffuse = reinterpret_cast<const float4*>(ctx->env_maps + spec_texels * 4u);
    e.spec_size = S;
    e.spec_mips = M;
    e.diff_size = D;
    for (int k = 0; k < 4; ++k) e.rotation[k] = probe->view_rotation[k];
    e.intensity = probe->intensity;
    e.on = 1u;
    ctx->env = e;
    return HM_OK;
}

void hm_key_direction(const hm_directional_light* lights, uint32_t count, float out[3]) {  // pass.rs:711-719
    V3 d = v3(0.35f, 0.82f, 0.45f);
    const hm_directional_light* pick = nullptr;
    for (uint32_t i = 0; i < count && !pick; ++i)
        if (lights[i].flags & HM_LIGHT_FLAG_SHADOWS) pick = &lights[i];
    if (!pick && count) pick = &lights[0];
    if (pick) d = v3(pick->direction_to_light[0], pick->direction_to_light[1], pick->direction_to_light[2]);
    d = normalize_or(d, v3(0, 1, 0));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

hm_status hm_make_params(const hm_view* view, const hm_groom_config* groom, const float bounds_min[3],
                         const float bounds_max[3], uint32_t bundle_count, const float direction_to_light[3],
                         uint32_t width, uint32_t height, uint32_t dom_size, hm_params out[3]) {
    if (!view || !groom || !bounds_min || !bounds_max || !direction_to_light || !out) return HM_ERR_INVALID;
    M4 wfv = load_m4(view->world_from_view);
    V3 camera_pos = transform_point3(wfv, v3(0, 0, 0));
    V3 forward = normalize_or(scale(transform_vector3(wfv, v3(0, 0, 1)), -1.0f), v3(0, 0, -1));
    M4 clip = load_m4(view->clip_from_world);
    M4 wfc = inverse(clip);
    bool reverse = clip_z(clip, add(camera_pos, scale(forward, 0.5f))) > clip_z(clip, add(camera_pos, scale(forward, 30.0f)));
    M4 wfm = load_m4(groom->world_from_model);
    V3 light_eye, light_dir;
    M4 light_clip;
    float light_near, light_far;
    light_fit(bounds_min, bounds_max, wfm, v3(direction_to_light[0], direction_to_light[1], direction_to_light[2]),
              light_eye, light_dir, light_clip, light_near, light_far);
    V3 center;
    float radius;
    world_extent(bounds_min, bounds_max, wfm, center, radius);
    float dist = std::fmax(length(sub(camera_pos, center)), 0.5f);
    const float cam_near = 0.05f;
    float cam_far = std::fmax(dist + radius * 4.0f, cam_near + 1.0f);
    float dom_scale = (light_far - light_near) / (float)HM_LAYER_COUNT;

    ViewArgs cam;
    cam.clip_from_world = clip;
    cam.world_from_clip = wfc;
    cam.world_from_model = wfm;
    cam.light_clip = light_clip;
    cam.camera_pos = camera_pos;
    cam.camera_forward = forward;
    float scr[4] = {(float)width, (float)height, cam_near, cam_far};
    std::memcpy(cam.screen, scr, 16);
    cam.light_forward = light_dir;
    cam.dom_near = light_near;
    cam.dom_far = light_far;
    cam.dom_scale = dom_scale;
    uint32_t pm0[4] = {0u, bundle_count, reverse ? 1u : 0u, 0u};
    std::memcpy(cam.pass_mode, pm0, 16);
    cam.light_eye = light_eye;
    cam.dom_size = dom_size;
    make_params(cam, *groom, out[0]);

    ViewArgs dom = cam;
    dom.clip_from_world = light_clip;
    dom.world_from_clip = inverse(light_clip);
    dom.camera_pos = light_eye;
    dom.camera_forward = light_dir;
    float dscr[4] = {(float)dom_size, (float)dom_size, light_near, light_far};
    std::memcpy(dom.screen, dscr, 16);
    uint32_t pm1[4] = {1u, bundle_count, reverse ? 1u : 0u, 1u};
    std::memcpy(dom.pass_mode, pm1, 16);
    make_params(dom, *groom, out[1]);
    out[2] = out[1];
    out[2].pass_mode[0] = 2u;
    return HM_OK;
}

void hm_shadow_light_from_params(const hm_params* dom_params, hm_shadow_light* out) {  // pass.rs:574-579
    if (!dom_params || !out) return;
    std::memcpy(out->light_clip_from_world, dom_params->clip_from_world, 64);
    float eye[4] = {dom_params->light_eye[0], dom_params->light_eye[1], dom_params->light_eye[2], 1.0f};
    float fwd[4] = {dom_params->light_forward[0], dom_params->light_forward[1], dom_params->light_forward[2], 0.0f};
    float depth[4] = {dom_params->dom_info[2], dom_params->dom_info[3], HM_DOM_SIGMA, 1.0f};
    std::memcpy(out->light_eye, eye, 16);
    std::memcpy(out->light_forward, fwd, 16);
    std::memcpy(out->depth, depth, 16);
}

hm_status hm_clear(hm_ctx* ctx) {
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    Timer t(ctx, HM_K_CLEAR);
    HM_HIP(ctx, hm::launch_fill32(ctx->stream, static_cast<uint32_t*>(ctx->gb_slab), ctx->gb_words, 0xffffffffu,
                                  ctx->cu_count * 8));
    return HM_OK;
}

hm_status hm_clear_dom(hm_ctx* ctx) {
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    Timer t(ctx, HM_K_CLEAR_DOM);
    size_t texels = (size_t)ctx->dom_size * ctx->dom_size;
    HM_HIP(ctx, hm::launch_fill32(ctx->stream, ctx->dom, texels, 0xffffffffu, ctx->cu_count * 8));
    HM_HIP(ctx, hm::launch_fill32(ctx->stream, ctx->dom + texels, texels * HM_DOM_LAYERS, 0u, ctx->cu_count * 8));
    return HM_OK;
}

hm_status hm_lod(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass, const hm_lod_override* ovr) {
    hm_status st = check_pass_params(ctx, g, p, pass);
    if (st != HM_OK) return st;
    Timer t(ctx, pass == HM_PASS_CAMERA ? HM_K_CAM_LOD : HM_K_DOM_LOD);
    uint32_t oc = ovr ? ovr->control_points : 0u;
    float of = ovr ? ovr->strand_fraction : 0.0f;
    HM_HIP(ctx, hm::launch_lod(ctx->stream, *p, g->bundles, g->layers, g->lod[pass], p->pass_mode[1], oc, of));
    // The control points k_lod can give: 65 for the DOM (control_points_for_lod(1)),
    // else the override, else 65 with LOD on and 127 with it off.
    g->ncp_bound[pass] = p->pass_mode[3] == 1u ? 65u
                         : oc                  ? std::min(std::max(oc, 2u), 127u)
                         : p->flags[0] != 0u   ? 65u
                                               : 127u;
    return HM_OK;
}

hm_status hm_scan(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass) {
    hm_status st = check_pass_params(ctx, g, p, pass);
    if (st != HM_OK) return st;
    Timer t(ctx, pass == HM_PASS_CAMERA ? HM_K_CAM_SCAN : HM_K_DOM_SCAN);
    HM_HIP(ctx, hm::launch_scan(ctx->stream, g->lod[pass], p->pass_mode[1], g->capacity, g->prefix[pass],
                                g->indirect[pass], g->refs[pass]));
    return HM_OK;
}

static hm_status raster_args(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass, hm::RasterArgs& a) {
    hm_status st = check_pass_params(ctx, g, p, pass);
    if (st != HM_OK) return st;
    st = check_targets(ctx);
    if (st != HM_OK) return st;
    uint32_t mode = p->pass_mode[0];
    if (mode > 2u || (mode == 0u) != (pass == HM_PASS_CAMERA))
        return fail(ctx, HM_ERR_INVALID, "pass_mode.x does not match the pass");
    if (mode == 0u) {
        if ((uint32_t)p->screen[0] != ctx->width || (uint32_t)p->screen[1] != ctx->height)
            return fail(ctx, HM_ERR_INVALID, "camera params screen size differs from the targets");
    } else if ((uint32_t)p->dom_info[0] != ctx->dom_size || (uint32_t)p->dom_info[1] != ctx->dom_size ||
               (uint32_t)p->screen[0] != ctx->dom_size || (uint32_t)p->screen[1] != ctx->dom_size) {
        return fail(ctx, HM_ERR_INVALID, "DOM params size differs from the targets' DOM size");
    }
    std::memcpy(&a.view, p, sizeof(hm_params));
    a.bundles = g->bundles;
    a.layers = g->layers;
    a.style = g->style;
    a.refs = g->refs[pass];
    a.indirect = g->indirect[pass];
    a.capacity = g->capacity;
    a.center = ctx->center;
    a.cons = ctx->cons;
    a.beta = ctx->beta;
    a.dom = ctx->dom;
    a.scene_depth = ctx->scene_depth_ext ? ctx->scene_depth_ext : ctx->scene_depth;
    a.scene_clear = ctx->scene_clear;
    a.counters = ctx->counters;
    a.opt = ctx->tuning.raster_flags[p->pass_mode[0] < 3u ? p->pass_mode[0] : 0u];
    a.gate = nullptr;
    return HM_OK;
}

hm_status hm_raster(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass) {
    hm::RasterArgs a;
    hm_status st = raster_args(ctx, g, p, pass, a);
    if (st != HM_OK) return st;
    int mode = (int)p->pass_mode[0];
    Timer t(ctx, mode == 0 ? HM_K_CAM_RASTER : (mode == 1 ? HM_K_DOM_DEPTH : HM_K_DOM_OPACITY));
    HM_HIP(ctx, hm::launch_raster(ctx->stream, a, mode, false, raster_blocks(ctx, g->capacity, p->pass_mode[0])));
    return HM_OK;
}

hm_status hm_raster_count(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass, uint64_t counts[4]) {
    if (!counts) return fail(ctx, HM_ERR_INVALID, "null counts");
    hm::RasterArgs a;
    hm_status st = raster_args(ctx, g, p, pass, a);
    if (st != HM_OK) return st;
    HM_HIP(ctx, hm::launch_fill32(ctx->stream, reinterpret_cast<uint32_t*>(ctx->counters), 8, 0u, 1));
    HM_HIP(ctx, hm::launch_raster(ctx->stream, a, (int)p->pass_mode[0], true, raster_blocks(ctx, g->capacity, p->pass_mode[0])));
    HM_HIP(ctx, hipMemcpyAsync(counts, ctx->counters, 32, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

// Binning work buffers and arguments for one pass over a tiles_x x tiles_y grid.
static hipError_t bin_args(hm_ctx* ctx, hm_groom* g, const hm_params* p, int pass, uint32_t T, uint32_t tiles_x,
                           uint32_t tiles_y, hm::BinArgs& b) {
    const uint32_t tiles = tiles_x * tiles_y;
    const size_t slots = g->capacity ? g->capacity : 1u;
    // Every segment (126 per strand at most) spans at most box_limit tiles or
    // goes to the long list, so this many records always fit. The default
    // record buffer is that size; a smaller bin_records can overflow, and then
    // the device switches to the direct kernels.
    const size_t worst = std::min<size_t>(slots * 126u * ctx->tuning.bin_box_limit, 0xffffffffu);
    const size_t want_records = ctx->tuning.bin_records ? std::min<size_t>(ctx->tuning.bin_records, worst) : worst;
    const int bin_blocks = bin_grid(ctx, pass);
    // (tile, rank) list entries: at most one per (counting workgroup, tile) and per record.
    const size_t lists = std::min<size_t>(want_records, (size_t)bin_blocks * tiles);
    // Camera tiles split into chunks: at most tiles + records / chunk of them.
    const uint32_t chunk = pass == HM_PASS_CAMERA ? ctx->tuning.cam_chunk : 0u;
    const size_t chunks = tiles + (chunk ? want_records / chunk + 1u : 0u);
    hipError_t e = ensure_bin(ctx, pass, slots, hm::BIN_META_WORDS + 5u * (size_t)tiles + 2u, want_records, lists,
                              (size_t)bin_blocks, chunks);
    if (e != hipSuccess) return e;
    std::memset(&b, 0, sizeof(b));
    std::memcpy(&b.view, p, sizeof(hm_params));
    b.bundles = g->bundles;
    b.layers = g->layers;
    b.style = g->style;
    b.refs = g->refs[pass];
    b.indirect = g->indirect[pass];
    b.capacity = g->capacity;
    b.cps = ctx->bin.cps;
    b.cpy = ctx->bin.cpy;
    b.dcp = ctx->bin.dcp;
    b.seg_box = ctx->bin.seg_box[pass];
    b.meta = ctx->bin.meta[pass];
    b.records = ctx->bin.records[pass];
    b.record_capacity = (uint32_t)want_records;
    b.can_overflow = want_records < worst ? 1u : 0u;
    b.lists = ctx->bin.lists[pass];
    b.list_capacity = (uint32_t)std::min<size_t>(lists, 0xffffffffu);
    b.wg_list = ctx->bin.wg_list[pass];
    b.chunk = chunk;
    b.chunk_tile = pass == HM_PASS_CAMERA ? ctx->bin.chunk_tile : nullptr;  // DOM tiles are never split
    b.chunk_capacity = (uint32_t)std::min<size_t>(chunks, 0xffffffffu);
    b.gb_slab = static_cast<uint32_t*>(ctx->gb_slab);
    b.gb_words = ctx->gb_words;
    b.fill_clear = pass == HM_PASS_CAMERA ? 1u : 2u;
    b.long_list = ctx->bin.long_list[pass];
    // Segment slots per strand: the pass's most segments, rounded up to a power of two (16 .. 128).
    b.seg_shift = 4u;
    while ((1u << b.seg_shift) + 1u < g->ncp_bound[pass]) ++b.seg_shift;
    b.long_capacity = (uint32_t)std::min<size_t>(slots << b.seg_shift, 0xffffffffu);
    b.tiles_x = tiles_x;
    b.tiles = tiles;
    b.tile_shift = T == 16u ? 4u : 5u;
    b.box_limit = ctx->tuning.bin_box_limit;
    b.dom_size = ctx->dom_size;
    b.dom = ctx->dom;
    b.cpt = ctx->bin.cpt;
    b.sinfo = ctx->bin.sinfo;
    b.width = ctx->width;
    b.height = ctx->height;
    b.center = ctx->center;
    b.cons = ctx->cons;
    b.beta = ctx->beta;
    b.scene_depth = ctx->scene_depth_ext ? ctx->scene_depth_ext : ctx->scene_depth;
    b.scene_clear = ctx->scene_clear;
    return hipSuccess;
}

using TileLists = hm_lists;

// hm_gbuffer; with emit_lists (a frame's own call) the camera tiles also
// write the shade and filter pixel lists, described in *lists.
static hm_status gbuffer_impl(hm_ctx* ctx, hm_groom* g, const hm_params* p, bool emit_lists, TileLists* lists) {
    if (lists) *lists = TileLists();
    hm::RasterArgs a;
    hm_status st = raster_args(ctx, g, p, HM_PASS_CAMERA, a);
    if (st != HM_OK) return st;
    const uint32_t T = ctx->tuning.cam_tile;
    const uint32_t tiles_x = T ? (ctx->width + T - 1u) / T : 0u, tiles_y = T ? (ctx->height + T - 1u) / T : 0u;
    const bool binned = T != 0u && tiles_x * tiles_y <= hm::BIN_MAX_TILES &&
                        (uint64_t)g->capacity * hm::BIN_STRIDE <= 0xffffffffull;
    if (!binned) {
        HM_HIP(ctx, mark_direct(ctx, HM_PASS_CAMERA));
        if ((st = hm_clear(ctx)) != HM_OK) return st;
        return hm_raster(ctx, g, p, HM_PASS_CAMERA);
    }
    hm::BinArgs b;
    HM_HIP(ctx, bin_args(ctx, g, p, HM_PASS_CAMERA, T, tiles_x, tiles_y, b));
    const int bin_blocks = bin_grid(ctx, HM_PASS_CAMERA);
    b.emit_lists = emit_lists ? 1u : 0u;
    b.shade_list = ctx->pix_list;
    b.filter_list = ctx->pix_list2;
    // Tile t writes sub-list t % LIST_SUBS; the fallback compaction, if it
    // runs instead, writes the same layout, so the capacity covers both.
    b.list_cap = std::max((b.tiles + hm::LIST_SUBS - 1u) / hm::LIST_SUBS * T * T,
                          hm::compact_list_cap(ctx->width * ctx->height, true));
    if ((size_t)b.list_cap * hm::LIST_SUBS > ctx->list_len)
        return fail(ctx, HM_ERR_INVALID, "pixel lists too small for the camera tiles");
    b.shaded = ctx->shaded;
    b.shaded_depth = ctx->shaded_depth;
    b.filtered = ctx->filtered;
    b.filtered_depth = ctx->filtered_depth;
    {
        Timer t(ctx, HM_K_CAM_BIN);
        const hipError_t e =
            hm::launch_bin(ctx->stream, b, 0, raster_blocks(ctx, g->capacity, 0u), bin_blocks, ctx->gen_event);
        if (e != hipSuccess) {
            ctx->bin.meta_dirty[HM_PASS_CAMERA] = true;
            return hip_fail(ctx, e, "launch_bin (camera)");
        }
    }
    {
        Timer t(ctx, HM_K_CAM_TILES);
        // One workgroup per chunk for up to 2 records per segment slot.
        const uint64_t usual = (uint64_t)b.tiles + (b.chunk ? (uint64_t)g->capacity * 256u / b.chunk : 0u);
        const int tile_blocks = (int)std::min<uint64_t>(std::max<uint64_t>(usual, 1u), 1u << 20);
        HM_HIP(ctx, hm::launch_cam_tiles(ctx->stream, b, tile_blocks));
    }
    {
        Timer t(ctx, HM_K_CAM_FINISH);
        HM_HIP(ctx, hm::launch_cam_finish(ctx->stream, b, ctx->cu_count));
    }
    const uint32_t* fallback = ctx->bin.meta[HM_PASS_CAMERA] + hm::BIN_FALLBACK;
    if (emit_lists && lists) {
        lists->from_tiles = true;
        lists->cap = b.list_cap;
        lists->gate = b.can_overflow ? fallback : nullptr;
    }
    if (!b.can_overflow) return HM_OK;
    // Direct kernels, switched on by the device only when the records overflowed.
    {
        Timer t(ctx, HM_K_CLEAR);
        HM_HIP(ctx, hm::launch_fill32(ctx->stream, static_cast<uint32_t*>(ctx->gb_slab), ctx->gb_words, 0xffffffffu,
                                      ctx->cu_count, fallback));
    }
    a.gate = fallback;
    {
        Timer t(ctx, HM_K_CAM_RASTER);
        HM_HIP(ctx, hm::launch_raster(ctx->stream, a, 0, false, std::min(raster_blocks(ctx, g->capacity, 0u),
                                                                          ctx->cu_count * 2)));
    }
    return HM_OK;
}

hm_status hm_gbuffer(hm_ctx* ctx, hm_groom* g, const hm_params* p) { return gbuffer_impl(ctx, g, p, false, nullptr); }

// hm_dom; with reuse_camera (a frame whose camera pass ran binned just before,
// with the same model transform) the DOM generation takes the world control
// points of strands the camera work list shares instead of generating them.
static hm_status dom_impl(hm_ctx* ctx, hm_groom* g, const hm_params* dp, const hm_params* op, bool reuse_camera) {
    hm::RasterArgs a1, a2;
    hm_status st = raster_args(ctx, g, dp, HM_PASS_DOM, a1);
    if (st != HM_OK) return st;
    if ((st = raster_args(ctx, g, op, HM_PASS_DOM, a2)) != HM_OK) return st;
    if (dp->pass_mode[0] != 1u || op->pass_mode[0] != 2u)
        return fail(ctx, HM_ERR_INVALID, "hm_dom takes the depth (pass_mode.x 1) then the opacity (2) params");
    const uint32_t T = ctx->tuning.dom_tile;
    const uint32_t S = ctx->dom_size;
    const uint32_t tiles_x = T ? (S + T - 1u) / T : 0u;
    // The tile kernel evaluates both passes with one view: the params must
    // agree except for the mode.
    hm_params same = *op;
    same.pass_mode[0] = dp->pass_mode[0];
    const bool binned = T != 0u && tiles_x * tiles_x <= hm::BIN_MAX_TILES && std::memcmp(&same, dp, sizeof(hm_params)) == 0 &&
                        (uint64_t)g->capacity * hm::BIN_STRIDE <= 0xffffffffull;
    if (!binned) {
        HM_HIP(ctx, mark_direct(ctx, HM_PASS_DOM));
        if ((st = hm_clear_dom(ctx)) != HM_OK) return st;
        if ((st = hm_raster(ctx, g, dp, HM_PASS_DOM)) != HM_OK) return st;
        return hm_raster(ctx, g, op, HM_PASS_DOM);
    }
    hm::BinArgs b;
    const float4* cps_before = ctx->bin.cps;
    HM_HIP(ctx, bin_args(ctx, g, dp, HM_PASS_DOM, T, tiles_x, tiles_x, b));
    if (reuse_camera && ctx->bin.cps == cps_before) {  // not reallocated since the camera pass
        b.reuse_refs = g->refs[HM_PASS_CAMERA];
        b.reuse_indirect = g->indirect[HM_PASS_CAMERA];
    }
    const int bin_blocks = bin_grid(ctx, HM_PASS_DOM);
    {
        Timer t(ctx, HM_K_DOM_BIN);
        const hipError_t e = hm::launch_bin(ctx->stream, b, 1, raster_blocks(ctx, g->capacity, 1u), bin_blocks);
        if (e != hipSuccess) {
            ctx->bin.meta_dirty[HM_PASS_DOM] = true;
            return hip_fail(ctx, e, "launch_bin (DOM)");
        }
    }
    {
        Timer t(ctx, HM_K_DOM_LONG_DEPTH);
        HM_HIP(ctx, hm::launch_dom_long(ctx->stream, b, 1, ctx->cu_count));
    }
    {
        Timer t(ctx, HM_K_DOM_TILES);
        HM_HIP(ctx, hm::launch_dom_tiles(ctx->stream, b, ctx->cu_count * (int)ctx->tuning.dom_busy_blocks_per_cu));
    }
    {
        Timer t(ctx, HM_K_DOM_LONG_OPACITY);
        HM_HIP(ctx, hm::launch_dom_long(ctx->stream, b, 2, ctx->cu_count));
    }
    if (!b.can_overflow) return HM_OK;
    // Direct kernels, switched on by the device only when the records overflowed.
    const uint32_t* fallback = ctx->bin.meta[HM_PASS_DOM] + hm::BIN_FALLBACK;
    const size_t texels = (size_t)S * S;
    {
        Timer t(ctx, HM_K_CLEAR_DOM);
        HM_HIP(ctx, hm::launch_fill32(ctx->stream, ctx->dom, texels, 0xffffffffu, ctx->cu_count, fallback));
        HM_HIP(ctx, hm::launch_fill32(ctx->stream, ctx->dom + texels, texels * HM_DOM_LAYERS, 0u, ctx->cu_count,
                                      fallback));
    }
    a1.gate = fallback;
    a2.gate = fallback;
    {
        Timer t(ctx, HM_K_DOM_DEPTH);
        HM_HIP(ctx, hm::launch_raster(ctx->stream, a1, 1, false, std::min(raster_blocks(ctx, g->capacity, 1u),
                                                                           ctx->cu_count * 2)));
    }
    {
        Timer t(ctx, HM_K_DOM_OPACITY);
        HM_HIP(ctx, hm::launch_raster(ctx->stream, a2, 2, false, std::min(raster_blocks(ctx, g->capacity, 2u),
                                                                           ctx->cu_count * 2)));
    }
    return HM_OK;
}

hm_status hm_dom(hm_ctx* ctx, hm_groom* g, const hm_params* dp, const hm_params* op) {
    return dom_impl(ctx, g, dp, op, false);
}

}  // extern "C"

hm_status hm_shade_list(hm_ctx* ctx, const hm_lists& lists, hm::PixelList* out) {
    const uint32_t px = ctx->width * ctx->height;
    hm::PixelList L{ctx->pix_list, ctx->pix_count, hm::compact_list_cap(px, lists.gate != nullptr)};
    if (lists.from_tiles) L = {ctx->pix_list, ctx->bin.meta[HM_PASS_CAMERA] + hm::BIN_SHADE_SUBS, lists.cap};
    if (!lists.from_tiles || lists.gate) {
        // same layout as the camera tiles' when it replaces them
        if ((size_t)L.cap * hm::LIST_SUBS > ctx->list_len || L.cap < hm::compact_list_cap(px, lists.gate != nullptr))
            return fail(ctx, HM_ERR_INVALID, "pixel lists too small for the compaction");
        HM_HIP(ctx, hm::launch_compact(ctx->stream, 0, ctx->center, ctx->cons, px, L, ctx->shaded, ctx->shaded_depth,
                                       lists.gate));
    }
    *out = L;
    return HM_OK;
}

extern "C" {

// Shading over the covered-pixel list: the camera tiles' (then the
// compaction runs only if the direct kernels may have run instead) or one
// compacted here.
static hm_status shade_impl(hm_ctx* ctx, const hm_params* p, float exposure, const TileLists& lists) {
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    if (!p || (uint32_t)p->screen[0] != ctx->width || (uint32_t)p->screen[1] != ctx->height)
        return fail(ctx, HM_ERR_INVALID, "shade params screen size differs from the targets");
    if ((uint32_t)p->dom_info[0] != ctx->dom_size || (uint32_t)p->dom_info[1] != ctx->dom_size)
        return fail(ctx, HM_ERR_INVALID, "shade params DOM size differs from the targets");
    hm::ShadeArgs a;
    std::memcpy(&a.view, p, sizeof(hm_params));
    a.center = ctx->center;
    a.beta = ctx->beta;
    a.dom = ctx->dom;
    hm_shade_lights(ctx, a);
    a.exposure = exposure;
    a.shaded = ctx->shaded;
    a.shaded_depth = ctx->shaded_depth;
    Timer t(ctx, HM_K_SHADE);
    uint32_t px = ctx->width * ctx->height;
    hm::PixelList L;
    if ((st = hm_shade_list(ctx, lists, &L)) != HM_OK) return st;
    HM_HIP(ctx, hm::launch_shade(ctx->stream, a, L, px, ctx->cu_count * 8));
    return HM_OK;
}

hm_status hm_shade(hm_ctx* ctx, const hm_params* p, float exposure) {
    return shade_impl(ctx, p, exposure, TileLists());
}

static hm_status filter_impl(hm_ctx* ctx, const hm_params* p, const TileLists& lists) {
    hm_status st = check_targets(ctx);
    if (st != HM_OK) return st;
    if (!p || (uint32_t)p->screen[0] != ctx->width || (uint32_t)p->screen[1] != ctx->height)
        return fail(ctx, HM_ERR_INVALID, "filter params screen size differs from the targets");
    hm::FilterArgs a;
    std::memcpy(&a.view, p, sizeof(hm_params));
    a.center = ctx->center;
    a.cons = ctx->cons;
    a.shaded = ctx->shaded;
    a.shaded_depth = ctx->shaded_depth;
    a.filtered = ctx->filtered;
    a.filtered_depth = ctx->filtered_depth;
    Timer t(ctx, HM_K_FILTER);
    uint32_t px = ctx->width * ctx->height;
    hm::PixelList L{ctx->pix_list, ctx->pix_count + hm::LIST_SUBS, hm::compact_list_cap(px, lists.gate != nullptr)};
    if (lists.from_tiles) L = {ctx->pix_list2, ctx->bin.meta[HM_PASS_CAMERA] + hm::BIN_FILTER_SUBS, lists.cap};
    if (!lists.from_tiles || lists.gate) {
        const hm::PixelList& C = L;  // same layout as the camera tiles' when it replaces them
        if ((size_t)C.cap * hm::LIST_SUBS > ctx->list_len || C.cap < hm::compact_list_cap(px, lists.gate != nullptr))
            return fail(ctx, HM_ERR_INVALID, "pixel lists too small for the compaction");
        HM_HIP(ctx, hm::launch_compact(ctx->stream, 1, ctx->center, ctx->cons, px, C, ctx->filtered,
                                       ctx->filtered_depth, lists.gate));
    }
    // With the camera tiles' lists and no overflow gate (the frame's binned
    // path) the tile kernel filters the same pixels, its taps from LDS.
    if (lists.from_tiles && !lists.gate && ctx->tuning.filter_tiles)
        HM_HIP(ctx, hm::launch_filter_tiles(ctx->stream, a, ctx->width, ctx->height));
    else
        HM_HIP(ctx, hm::launch_filter(ctx->stream, a, L, px, ctx->cu_count * 8));
    return HM_OK;
}

hm_status hm_filter(hm_ctx* ctx, const hm_params* p) { return filter_impl(ctx, p, TileLists()); }

hm_status hm_frame(hm_ctx* ctx, hm_groom* g, const hm_params params[3], const hm_frame_opts* opts) {
    if (!ctx || !g || !params) return fail(ctx, HM_ERR_INVALID, "null argument");
    hm_status st;
    float exposure = opts ? opts->exposure : 1.0f;
    const hm_lod_override* ovr = opts ? &opts->camera_lod : nullptr;
    struct FrameSlot {
        hm_ctx* c;
        ~FrameSlot() {
            if (c->slots && c->frame < c->slots) c->frame++;
        }
    } slot{ctx};
    const hm_params& cam = params[0];
    // The DOM pass runs on the side stream beside the camera pass. It starts
    // after the work already on main (an earlier frame's shading reads the
    // map), its generation waits until the camera generation has written the
    // control points it may reuse, and main waits for all of it before
    // shading. Every exit, errors included, joins the side stream back.
    const bool dom_pass = cam.flags[3] != 0u;
    const bool overlap = dom_pass && ctx->tuning.frame_serial == 0u;
    struct Side {
        hm_ctx* c;
        bool forked = false;
        hipError_t fork() {
            hipError_t e = hipEventRecord(c->fork_ev[0], c->main);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->side, c->fork_ev[0], 0);
            forked = e == hipSuccess;
            return e;
        }
        hipError_t join() {
            c->stream = c->main;
            if (!forked) return hipSuccess;
            forked = false;
            hipError_t e = hipEventRecord(c->fork_ev[2], c->side);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->main, c->fork_ev[2], 0);
            return e;
        }
        ~Side() {
            (void)join();
            c->gen_event = nullptr;
        }
    } side{ctx};
    if (overlap) {
        HM_HIP(ctx, side.fork());
        ctx->stream = ctx->side;
        if ((st = hm_lod(ctx, g, &params[1], HM_PASS_DOM, nullptr)) != HM_OK) return st;
        if ((st = hm_scan(ctx, g, &params[1], HM_PASS_DOM)) != HM_OK) return st;
        ctx->stream = ctx->main;
        ctx->gen_event = ctx->fork_ev[1];
    }
    if ((st = hm_lod(ctx, g, &cam, HM_PASS_CAMERA, ovr)) != HM_OK) return st;
    if ((st = hm_scan(ctx, g, &cam, HM_PASS_CAMERA)) != HM_OK) return st;
    TileLists lists;
    if ((st = gbuffer_impl(ctx, g, &cam, true, &lists)) != HM_OK) return st;
    ctx->gen_event = nullptr;
    if (dom_pass) {
        if (overlap) {
            // (the binned camera pass recorded fork_ev[1] after its generation;
            // the direct one reuses nothing, and the event follows all of it)
            if (!lists.from_tiles) HM_HIP(ctx, hipEventRecord(ctx->fork_ev[1], ctx->main));
            ctx->stream = ctx->side;
            HM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev[1], 0));
        } else {
            if ((st = hm_lod(ctx, g, &params[1], HM_PASS_DOM, nullptr)) != HM_OK) return st;
            if ((st = hm_scan(ctx, g, &params[1], HM_PASS_DOM)) != HM_OK) return st;
        }
        const bool same_model = std::memcmp(cam.world_from_model, params[1].world_from_model, 64) == 0;
        if ((st = dom_impl(ctx, g, &params[1], &params[2], lists.from_tiles && same_model)) != HM_OK) return st;
        HM_HIP(ctx, side.join());
    }
    {
        GroomBox box(ctx, g, cam);
        if (opts && opts->shading == HM_SHADING_MLP) {
            if ((st = hm_shade_mlp_impl(ctx, &cam, exposure, opts->mlp_tolerance, nullptr, lists)) != HM_OK) return st;
        } else if ((st = shade_impl(ctx, &cam, exposure, lists)) != HM_OK) {
            return st;
        }
    }
    if (cam.flags[1]) {
        if ((st = filter_impl(ctx, &cam, lists)) != HM_OK) return st;
    }
    return HM_OK;
}

static bool target_buffer(hm_ctx* ctx, hm_groom* g, hm_buffer which, void** ptr, size_t* bytes) {
    size_t px = (size_t)ctx->width * ctx->height;
    size_t dom = (size_t)ctx->dom_size * ctx->dom_size * (1 + HM_DOM_LAYERS) * 4;
    switch (which) {
        case HM_BUF_CENTER: *ptr = ctx->center; *bytes = px * 8; return ctx->center != nullptr;
        case HM_BUF_CONSERVATIVE: *ptr = ctx->cons; *bytes = px * 8; return ctx->cons != nullptr;
        case HM_BUF_BETA: *ptr = ctx->beta; *bytes = px * 4; return ctx->beta != nullptr;
        case HM_BUF_SHADED: *ptr = ctx->shaded; *bytes = px * 16; return ctx->shaded != nullptr;
        case HM_BUF_SHADED_DEPTH: *ptr = ctx->shaded_depth; *bytes = px * 4; return ctx->shaded_depth != nullptr;
        case HM_BUF_FILTERED: *ptr = ctx->filtered; *bytes = px * 16; return ctx->filtered != nullptr;
        case HM_BUF_FILTERED_DEPTH: *ptr = ctx->filtered_depth; *bytes = px * 4; return ctx->filtered_depth != nullptr;
        case HM_BUF_DOM: *ptr = ctx->dom; *bytes = dom; return ctx->dom != nullptr;
        case HM_BUF_SCENE_DEPTH:
            *ptr = ctx->scene_depth_ext ? const_cast<float*>(ctx->scene_depth_ext) : ctx->scene_depth;
            *bytes = px * 4;
            return *ptr != nullptr;
        case HM_BUF_CAM_BIN_STATUS: *ptr = ctx->bin.meta[0]; *bytes = hm::BIN_META_WORDS * 4; return ctx->bin.meta[0] != nullptr;
        case HM_BUF_DOM_BIN_STATUS: *ptr = ctx->bin.meta[1]; *bytes = hm::BIN_META_WORDS * 4; return ctx->bin.meta[1] != nullptr;
        default: break;
    }
    if (!g) return false;
    int pass = (which >= HM_BUF_DOM_LOD) ? 1 : 0;
    switch (which) {
        case HM_BUF_CAMERA_LOD:
        case HM_BUF_DOM_LOD: *ptr = g->lod[pass]; *bytes = (size_t)g->nb * sizeof(hm_lod_record); return true;
        case HM_BUF_CAMERA_REFS:
        case HM_BUF_DOM_REFS: *ptr = g->refs[pass]; *bytes = (size_t)g->capacity * sizeof(hm_strand_ref); return true;
        case HM_BUF_CAMERA_INDIRECT:
        case HM_BUF_DOM_INDIRECT: *ptr = g->indirect[pass]; *bytes = 16; return true;
        default: return false;
    }
}

hm_status hm_read(hm_ctx* ctx, hm_groom* g, hm_buffer which, void* dst, size_t bytes) {
    if (!ctx || !dst) return fail(ctx, HM_ERR_INVALID, "null argument");
    void* src = nullptr;
    size_t size = 0;
    if (!target_buffer(ctx, g, which, &src, &size)) return fail(ctx, HM_ERR_NOT_READY, "buffer not allocated");
    if (bytes > size) return fail(ctx, HM_ERR_INVALID, "read larger than the buffer");
    (void)hipSetDevice(ctx->device);
    HM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

hm_status hm_device_ptr(hm_ctx* ctx, hm_groom* g, hm_buffer which, void** ptr, size_t* bytes) {
    if (!ctx || !ptr || !bytes) return fail(ctx, HM_ERR_INVALID, "null argument");
    if (!target_buffer(ctx, g, which, ptr, bytes)) return fail(ctx, HM_ERR_NOT_READY, "buffer not allocated");
    return HM_OK;
}

hm_status hm_sync(hm_ctx* ctx) {
    if (!ctx) return HM_ERR_INVALID;
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

hm_status hm_timing_enable(hm_ctx* ctx, int frames) {
    if (!ctx || frames < 0) return HM_ERR_INVALID;
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t want = (uint32_t)frames;
    size_t need = (size_t)want * HM_K_COUNT * 2;
    while (ctx->ev.size() < need) {
        hipEvent_t e = nullptr;
        HM_HIP(ctx, hipEventCreate(&e));
        ctx->ev.push_back(e);
    }
    ctx->ran.assign((size_t)want * HM_K_COUNT, 0);
    ctx->slots = want;
    ctx->frame = 0;
    return HM_OK;
}

hm_status hm_timing_read(hm_ctx* ctx, float* ms, uint32_t count) {
    if (!ctx || !ms) return HM_ERR_INVALID;
    HM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t frames = ctx->frame < ctx->slots ? ctx->frame : ctx->slots;
    if (frames == 0 && ctx->slots) frames = 1;  // kernels launched outside hm_frame land in slot 0
    for (uint32_t k = 0; k < count; ++k) {
        ms[k] = 0.0f;
        if (k >= HM_K_COUNT) continue;
        double sum = 0.0;
        uint32_t n = 0;
        for (uint32_t f = 0; f < frames; ++f) {
            size_t i = (size_t)f * HM_K_COUNT + k;
            if (!ctx->ran[i]) continue;
            float v = 0.0f;
            HM_HIP(ctx, hipEventElapsedTime(&v, ctx->ev[i * 2], ctx->ev[i * 2 + 1]));
            sum += v;
            ++n;
        }
        ms[k] = n ? (float)(sum / n) : 0.0f;
    }
    return HM_OK;
}

hm_status hm_get_tuning(hm_ctx* ctx, hm_tuning* out) {
    if (!ctx || !out) return HM_ERR_INVALID;
    *out = ctx->tuning;
    return HM_OK;
}

hm_status hm_set_tuning(hm_ctx* ctx, const hm_tuning* t) {
    if (!ctx || !t) return HM_ERR_INVALID;
    for (int m = 0; m < 3; ++m) {
        if (t->raster_blocks_per_cu[m] == 0 || t->raster_blocks_per_cu[m] > 64)
            return fail(ctx, HM_ERR_INVALID, "raster_blocks_per_cu must be in [1, 64]");
        if (t->raster_flags[m] & ~(HM_RASTER_PRECHECK | HM_RASTER_XCD)) return fail(ctx, HM_ERR_INVALID, "unknown flag");
    }
    if (t->dom_tile != 0u && t->dom_tile != 16u && t->dom_tile != 32u)
        return fail(ctx, HM_ERR_INVALID, "dom_tile must be 0, 16 or 32");
    if (t->cam_tile != 0u && t->cam_tile != 16u && t->cam_tile != 32u)
        return fail(ctx, HM_ERR_INVALID, "cam_tile must be 0, 16 or 32");
    if (t->cam_chunk != 0u && (t->cam_chunk < 64u || t->cam_chunk > (1u << 20)))
        return fail(ctx, HM_ERR_INVALID, "cam_chunk must be 0 or in [64, 1048576]");
    if (t->bin_box_limit == 0u || t->bin_box_limit > 64u)
        return fail(ctx, HM_ERR_INVALID, "bin_box_limit must be in [1, 64]");
    if (t->bin_blocks_per_cu == 0u || t->bin_blocks_per_cu > 64u)
        return fail(ctx, HM_ERR_INVALID, "bin_blocks_per_cu must be in [1, 64]");
    if (t->cam_bin_blocks_per_cu > 64u) return fail(ctx, HM_ERR_INVALID, "cam_bin_blocks_per_cu must be in [0, 64]");
    if (t->light_cull > 2u) return fail(ctx, HM_ERR_INVALID, "light_cull must be 0, 1 or 2");
    if (t->dom_busy_blocks_per_cu > 2u) return fail(ctx, HM_ERR_INVALID, "dom_busy_blocks_per_cu must be 0, 1 or 2");
    if (t->frame_serial > 1u) return fail(ctx, HM_ERR_INVALID, "frame_serial must be 0 or 1");
    if (t->batch_lanes < 1u || t->batch_lanes > 4u) return fail(ctx, HM_ERR_INVALID, "batch_lanes must be in [1, 4]");
    if (t->filter_tiles > 1u) return fail(ctx, HM_ERR_INVALID, "filter_tiles must be 0 or 1");

    ctx->tuning = *t;
    return HM_OK;
}

// The search of hm_autotune; `out` starts as the context's tuning and ends as
// the fastest variant. Leaves ctx->tuning at whatever variant ran last: the
// caller restores or replaces it.
static hm_status autotune_search(hm_ctx* ctx, hm_groom* g, const hm_params params[3], const hm_frame_opts* opts,
                                 uint32_t reps, hm_tuning& out) {
    hm_status st;
    const uint32_t grids[4] = {4u, 8u, 16u, 32u};
    const hm_lod_override* ovr = opts ? &opts->camera_lod : nullptr;
    // Work lists of both passes.
    if ((st = hm_lod(ctx, g, &params[0], HM_PASS_CAMERA, ovr)) != HM_OK) return st;
    if ((st = hm_scan(ctx, g, &params[0], HM_PASS_CAMERA)) != HM_OK) return st;
    if ((st = hm_lod(ctx, g, &params[1], HM_PASS_DOM, nullptr)) != HM_OK) return st;
    if ((st = hm_scan(ctx, g, &params[1], HM_PASS_DOM)) != HM_OK) return st;
    float best_ms[3] = {1e30f, 1e30f, 1e30f};
    hipEvent_t e0 = ctx->frame_ev[0], e1 = ctx->frame_ev[1];
    for (int mode = 0; mode < 3; ++mode) {
        const hm_params* p = &params[mode];
        int pass = mode == 0 ? HM_PASS_CAMERA : HM_PASS_DOM;
        for (uint32_t flags = 0; flags < 4u; ++flags) {
            for (uint32_t gi = 0; gi < 4; ++gi) {
                ctx->tuning = out;
                ctx->tuning.raster_flags[mode] = flags;
                ctx->tuning.raster_blocks_per_cu[mode] = grids[gi];
                // Each repetition starts from the state a frame sees: cleared
                // targets (and, for the opacity pass, a finished depth pass).
                float total = 0.0f;
                for (uint32_t r = 0; r <= reps; ++r) {
                    if (mode == 0) {
                        if ((st = hm_clear(ctx)) != HM_OK) return st;
                    } else {
                        if ((st = hm_clear_dom(ctx)) != HM_OK) return st;
                        if (mode == 2 && (st = hm_raster(ctx, g, &params[1], HM_PASS_DOM)) != HM_OK) return st;
                    }
                    HM_HIP(ctx, hipEventRecord(e0, ctx->stream));
                    if ((st = hm_raster(ctx, g, p, pass)) != HM_OK) return st;
                    HM_HIP(ctx, hipEventRecord(e1, ctx->stream));
                    HM_HIP(ctx, hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) total += ms;  // repetition 0 warms up
                }
                if (total < best_ms[mode]) {
                    best_ms[mode] = total;
                    out.raster_flags[mode] = flags;
                    out.raster_blocks_per_cu[mode] = grids[gi];
                }
            }
        }
    }
    // Camera: direct kernels against binned tiles of either size with 64 to
    // 2048 records (segments) per tile workgroup, whole hm_gbuffer timed.
    {
        constexpr uint32_t kVariants = 13;
        const uint32_t variants[kVariants][2] = {{0u, 0u},     {16u, 64u},   {16u, 128u},  {16u, 256u}, {16u, 512u},
                                                 {16u, 1024u}, {16u, 2048u}, {32u, 64u},   {32u, 128u}, {32u, 256u},
                                                 {32u, 512u},  {32u, 1024u}, {32u, 2048u}};
        float best_cam = 1e30f;
        for (uint32_t ti = 0; ti < kVariants; ++ti) {
            ctx->tuning = out;
            ctx->tuning.cam_tile = variants[ti][0];
            if (variants[ti][1]) ctx->tuning.cam_chunk = variants[ti][1];
            float total = 0.0f;
            for (uint32_t r = 0; r <= reps; ++r) {
                HM_HIP(ctx, hipEventRecord(e0, ctx->stream));
                if ((st = hm_gbuffer(ctx, g, &params[0])) != HM_OK) return st;
                HM_HIP(ctx, hipEventRecord(e1, ctx->stream));
                HM_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.0f;
                HM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total += ms;
            }
            if (total < best_cam) {
                best_cam = total;
                out.cam_tile = ctx->tuning.cam_tile;
                out.cam_chunk = ctx->tuning.cam_chunk;
            }
        }
    }
    // DOM: direct kernels against binned tiles of either size (16-texel tiles
    // also with the persistent tile workgroups, 1 or 2 per CU, or one
    // workgroup per tile), whole hm_dom timed.
    if (params[1].flags[3]) {
        constexpr uint32_t kVariants = 11;
        const uint32_t tile_sizes[kVariants][3] = {{0u, 0u, 0u},  {16u, 1u, 2u}, {16u, 2u, 2u}, {16u, 4u, 2u},
                                                   {16u, 2u, 1u}, {16u, 2u, 0u}, {16u, 1u, 0u}, {16u, 4u, 0u},
                                                   {32u, 1u, 0u}, {32u, 2u, 0u}, {32u, 4u, 0u}};
        float best_dom = 1e30f;
        for (uint32_t ti = 0; ti < kVariants; ++ti) {
            ctx->tuning = out;
            ctx->tuning.dom_tile = tile_sizes[ti][0];
            if (tile_sizes[ti][1]) ctx->tuning.bin_blocks_per_cu = tile_sizes[ti][1];
            ctx->tuning.dom_busy_blocks_per_cu = tile_sizes[ti][2];
            float total = 0.0f;
            for (uint32_t r = 0; r <= reps; ++r) {
                HM_HIP(ctx, hipEventRecord(e0, ctx->stream));
                if ((st = hm_dom(ctx, g, &params[1], &params[2])) != HM_OK) return st;
                HM_HIP(ctx, hipEventRecord(e1, ctx->stream));
                HM_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.0f;
                HM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total += ms;
            }
            if (total < best_dom) {
                best_dom = total;
                out.dom_tile = ctx->tuning.dom_tile;
                out.bin_blocks_per_cu = ctx->tuning.bin_blocks_per_cu;
                out.dom_busy_blocks_per_cu = ctx->tuning.dom_busy_blocks_per_cu;
            }
        }
    }
    // Frame: the DOM pass beside the camera pass or after it, whole hm_frame
    // timed. The two differ by a few percent, so they alternate frame by
    // frame (a slow drift of the clock counts against both) over more frames.
    if (params[0].flags[3]) {
        float total[2] = {0.0f, 0.0f};
        const uint32_t frames = reps < 4u ? 8u : 2u * reps;
        for (uint32_t r = 0; r <= frames; ++r) {
            for (uint32_t serial = 0; serial < 2u; ++serial) {
                ctx->tuning = out;
                ctx->tuning.frame_serial = serial;
                HM_HIP(ctx, hipEventRecord(e0, ctx->stream));
                if ((st = hm_frame(ctx, g, params, opts)) != HM_OK) return st;
                HM_HIP(ctx, hipEventRecord(e1, ctx->stream));
                HM_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.0f;
                HM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total[serial] += ms;  // the first round warms up
            }
        }
        out.frame_serial = total[1] < total[0] ? 1u : 0u;
    }
    // Filter on the binned path: over the camera tiles' pixel lists, or by
    // screen tiles with the taps in LDS (ahead on small views, where the
    // per-pixel chain's latency dominates); whole frames, alternating.
    if (params[0].flags[1] && out.cam_tile != 0u) {
        float total[2] = {0.0f, 0.0f};
        const uint32_t frames = reps < 4u ? 8u : 2u * reps;
        for (uint32_t r = 0; r <= frames; ++r) {
            for (uint32_t tiles = 0; tiles < 2u; ++tiles) {
                ctx->tuning = out;
                ctx->tuning.filter_tiles = tiles;
                HM_HIP(ctx, hipEventRecord(e0, ctx->stream));
                if ((st = hm_frame(ctx, g, params, opts)) != HM_OK) return st;
                HM_HIP(ctx, hipEventRecord(e1, ctx->stream));
                HM_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.0f;
                HM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total[tiles] += ms;
            }
        }
        out.filter_tiles = total[1] < total[0] ? 1u : 0u;
    }

    return HM_OK;
}

hm_status hm_autotune(hm_ctx* ctx, hm_groom* g, const hm_params params[3], const hm_frame_opts* opts, uint32_t reps,
                      hm_tuning* best) {
    if (!ctx || !g || !params || reps == 0) return fail(ctx, HM_ERR_INVALID, "null argument");
    const hm_tuning keep = ctx->tuning;
    const uint32_t saved_slots = ctx->slots;
    ctx->slots = 0;  // no per-kernel events while tuning
    hm_tuning out = keep;
    const hm_status st = autotune_search(ctx, g, params, opts, reps, out);
    // On any error the context gets back the tuning and timing it had.
    ctx->slots = saved_slots;
    ctx->tuning = st == HM_OK ? out : keep;
    if (st == HM_OK && best) *best = out;
    return st;
}

hm_status hm_bench_copy(hm_ctx* ctx, size_t bytes, uint32_t reps, float* gbps) {
    if (!ctx || !gbps || bytes < 16 || reps == 0) return HM_ERR_INVALID;
    bytes &= ~(size_t)15;
    (void)hipSetDevice(ctx->device);
    void *a = nullptr, *b = nullptr;
    HM_HIP(ctx, hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hm::launch_fill32(ctx->stream, (uint32_t*)a, bytes / 4, 1u, ctx->cu_count * 8);
    if (e == hipSuccess) e = hm::launch_copy(ctx->stream, a, b, bytes, ctx->cu_count * 8);  // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[0], ctx->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = hm::launch_copy(ctx->stream, a, b, bytes, ctx->cu_count * 8);
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->frame_ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->frame_ev[0], ctx->frame_ev[1]);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(ctx, e, "hm_bench_copy");
    *gbps = (float)(2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9);  // read + write
    return HM_OK;
}

hm_status hm_bench_atomics(hm_ctx* ctx, size_t target_bytes, uint32_t kind, uint64_t count, float* per_second) {
    if (!ctx || !per_second || kind > 2 || count == 0) return HM_ERR_INVALID;
    size_t word = kind == 0 ? 8 : 4;
    size_t words = target_bytes / word;
    if (words == 0) return HM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    void* buf = nullptr;
    HM_HIP(ctx, hipMalloc(&buf, words * word));
    uint32_t threads = (uint32_t)ctx->cu_count * 2048u;
    uint32_t per = (uint32_t)((count + threads - 1) / threads);
    hipError_t e = hm::launch_fill32(ctx->stream, (uint32_t*)buf, words * word / 4, 0xffffffffu, ctx->cu_count * 8);
    if (e == hipSuccess) e = hm::launch_atomics(ctx->stream, buf, words, threads, per, kind);  // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[0], ctx->stream);
    if (e == hipSuccess) e = hm::launch_atomics(ctx->stream, buf, words, threads, per, kind);
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->frame_ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->frame_ev[0], ctx->frame_ev[1]);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(ctx, e, "hm_bench_atomics");
    *per_second = (float)((double)threads * per / (ms * 1e-3));
    return HM_OK;
}

hm_status hm_bench_lds_atomics(hm_ctx* ctx, uint32_t lds_bytes, uint32_t kind, uint64_t count, float* per_second) {
    if (!ctx || !per_second || kind > 2 || count == 0) return HM_ERR_INVALID;
    const uint32_t word = kind == 0 ? 8u : 4u;
    if (lds_bytes < 1024u * word || lds_bytes > 65536u) return fail(ctx, HM_ERR_INVALID, "lds_bytes out of range");
    uint32_t words = 1u;
    while (words * 2u * word <= lds_bytes) words *= 2u;  // largest power of two that fits
    (void)hipSetDevice(ctx->device);
    const uint32_t blocks = (uint32_t)ctx->cu_count * 8u;
    const uint64_t threads = (uint64_t)blocks * 256u;
    const uint32_t per = (uint32_t)std::max<uint64_t>(1u, (count + threads - 1) / threads);
    uint32_t* sink = nullptr;
    HM_HIP(ctx, hipMalloc(&sink, blocks * 4u));
    hipError_t e = hm::launch_lds_atomics(ctx->stream, sink, blocks, words, per, kind);  // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[0], ctx->stream);
    if (e == hipSuccess) e = hm::launch_lds_atomics(ctx->stream, sink, blocks, words, per, kind);
    if (e == hipSuccess) e = hipEventRecord(ctx->frame_ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->frame_ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->frame_ev[0], ctx->frame_ev[1]);
    (void)hipFree(sink);
    if (e != hipSuccess) return hip_fail(ctx, e, "hm_bench_lds_atomics");
    *per_second = (float)((double)threads * per / (ms * 1e-3));
    return HM_OK;
}

hm_status hm_time_frames(hm_ctx* ctx, hm_groom* g, const hm_params params[3], const hm_frame_opts* opts,
                         uint32_t frames, float* total_ms) {
    if (!ctx || !total_ms) return HM_ERR_INVALID;
    HM_HIP(ctx, hipEventRecord(ctx->frame_ev[0], ctx->stream));
    for (uint32_t f = 0; f < frames; ++f) {
        hm_status st = hm_frame(ctx, g, params, opts);
        if (st != HM_OK) return st;
    }
    HM_HIP(ctx, hipEventRecord(ctx->frame_ev[1], ctx->stream));
    HM_HIP(ctx, hipEventSynchronize(ctx->frame_ev[1]));
    HM_HIP(ctx, hipEventElapsedTime(total_ms, ctx->frame_ev[0], ctx->frame_ev[1]));
    return HM_OK;
}

}  // extern "C"
