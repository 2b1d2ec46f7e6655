// gfx950 kernels of the deferred strand rasterizer.
//
// Reference kernels (paths relative to the reference repository):
//   k_fill32          clear.wgsl:7-16, clear_dom.wgsl:8-21 (one fill per slab)
//   k_lod             lod.wgsl:6-59         one wave per bundle, lane = cage corner
//   k_scan/k_expand   scan.wgsl:6-24        parallel scan, same StrandRef order
//   k_raster<MODE>    raster.wgsl:225-337   one wave per strand, lane = segment
//   k_shade           shade.wesl:523-593    one thread per pixel
//   k_filter          filter.wgsl:20-135    one thread per pixel
#include "hm_kernels.h"

#include <cstring>

#include "hm_device.h"

namespace hm {

static_assert(sizeof(View) == sizeof(hm_params), "View must mirror hm_params byte for byte");

// ---- fills ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill32(uint32_t* __restrict__ dst, size_t words, uint32_t value) {
    size_t quads = words / 4;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    uint4 v = make_uint4(value, value, value, value);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < quads; i += stride) d4[i] = v;
    size_t tail = quads * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tail < words) dst[tail] = value;
}

// ---- lod.wgsl -----------------------------------------------------------------

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = mn(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = mx(v, __shfl_xor(v, off));
    return v;
}

__global__ __launch_bounds__(256) void k_lod(View p, const hm_bundle* __restrict__ bundles,
                                             const hm_layer* __restrict__ layers, hm_lod_record* __restrict__ out,
                                             uint32_t nb, uint32_t ovr_cps, float ovr_frac) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t bundle_id = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (bundle_id >= nb) return;
    const hm_bundle bundle = bundles[bundle_id];
    // Lane = layer * 4 + corner, the order lod.wgsl:320-322 walks.
    F4 v = ld4(layers[bundle.layer_offset + lane].pos_ao);
    F3 world = to_world(p, xyz(v));
    float cw = mat_row(p.clip, 3, world.x, world.y, world.z, 1.0f);
    float dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
    bool front = cw > 0.0f && dist > 0.0f;
    Projected pr = project_world(p, world);
    float minx = wave_min(pr.screen.x), miny = wave_min(pr.screen.y);
    float maxx = wave_max(pr.screen.x), maxy = wave_max(pr.screen.y);
    bool any_front = __any(front);
    if (lane != 0) return;
    bool culled = !any_front || maxx < 0.0f || maxy < 0.0f || minx > p.screen.x || miny > p.screen.y;
    uint32_t n_lod = 0u, cps = HM_LAYER_COUNT;
    if (!culled) {
        if (p.pass_mode[3] == 1u) {
            n_lod = bundle.strand_count;
            cps = control_points_for_lod(1.0f);
        } else if (p.flags[0] == 0u) {
            n_lod = bundle.strand_count;
            cps = 127u;
        } else {
            F2 d = {maxx - minx, maxy - miny};
            float l = clampf(rdiv(len2(d), mx(p.screen.y, 1.0f)) * p.appearance.z, 0.0f, 1.0f);
            float delta = u2f(pcg(bundle_id ^ 0x9e3779b9u)) * (1.0f / 4294967296.0f);
            float n = u2f(bundle.strand_count);
            float raw = ceilf(l * (n + delta));
            n_lod = to_u32(clampf(raw, 1.0f, n));
            cps = control_points_for_lod(l);
        }
        if (p.pass_mode[3] == 0u) {
            if (ovr_cps) cps = ovr_cps < 2u ? 2u : (ovr_cps > 127u ? 127u : ovr_cps);
            if (ovr_frac > 0.0f) {
                float n = u2f(bundle.strand_count);
                n_lod = to_u32(clampf(ceilf(ovr_frac * n), 1.0f, n));
            }
        }
    }
    hm_lod_record rec = {n_lod, cps, bundle.strand_count, 0u};
    out[bundle_id] = rec;
}

// ---- scan.wgsl ------------------------------------------------------------------

// Exclusive scan of n_lod over bundles in one 1024-thread block. prefix[b] is
// the first StrandRef index of bundle b; prefix[nb] the uncapped total.
__global__ __launch_bounds__(1024) void k_scan(const hm_lod_record* __restrict__ lod, uint32_t nb,
                                               uint32_t capacity, uint32_t* __restrict__ prefix,
                                               uint32_t* __restrict__ indirect) {
    __shared__ uint32_t wave_sums[16];
    __shared__ uint32_t carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024u) {
        uint32_t i = base + tid;
        uint32_t v = i < nb ? lod[i].n_lod : 0u;
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t up = __shfl_up(incl, off);
            if (lane >= (uint32_t)off) incl += up;
        }
        if (lane == 63) wave_sums[wid] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t w = 0; w < wid; ++w) before += wave_sums[w];
        if (i < nb) prefix[i] = before + incl - v;
        __syncthreads();
        if (tid == 1023) carry = before + incl;
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t total = carry;
        prefix[nb] = total;
        indirect[0] = total < capacity ? total : capacity;
        indirect[1] = 1u;
        indirect[2] = 1u;
        indirect[3] = 0u;
    }
}

// StrandRef i belongs to the last bundle whose prefix is <= i (empty bundles
// share their successor's prefix, so the last one is the non-empty owner).
__global__ __launch_bounds__(256) void k_expand(const hm_lod_record* __restrict__ lod,
                                                const uint32_t* __restrict__ prefix, uint32_t nb,
                                                const uint32_t* __restrict__ indirect,
                                                hm_strand_ref* __restrict__ refs) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t total = indirect[0];
    if (i >= total) return;
    uint32_t lo = 0, hi = nb;  // invariant: prefix[lo] <= i < prefix[hi]
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid;
        else hi = mid;
    }
    hm_lod_record r = lod[lo];
    hm_strand_ref ref = {lo, i - prefix[lo], r.n_lod, r.control_points};
    refs[i] = ref;
}

// ---- raster.wgsl ----------------------------------------------------------------

__device__ __forceinline__ void gmin64(uint64_t* a, uint64_t v) {
    // Targets only decrease during a pass, so a value read here is an upper
    // bound of the current one and a payload not below it cannot win.
    if (v < *a) atomicMin((unsigned long long*)a, (unsigned long long)v);
}
__device__ __forceinline__ void gmin32(uint32_t* a, uint32_t v) {
    if (v < *a) atomicMin(a, v);
}

__device__ __forceinline__ F3 shfl3(F3 v, int src) { return {__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)}; }
__device__ __forceinline__ F4 shfl4(F4 v, int src) {
    return {__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src)};
}

struct Counters {
    uint64_t cam_writes, cam_centers, dom_writes, segments;
};

template <int MODE, bool COUNT>
struct Drawer {
    const View& p;
    const RasterArgs& a;
    Counters c;

    __device__ __forceinline__ float scene_at(int32_t x, int32_t y) const {
        if (!a.scene_depth) return a.scene_clear;
        return a.scene_depth[(size_t)(uint32_t)y * (uint32_t)to_i32(p.screen.x) + (uint32_t)x];
    }

    // One camera write of a given payload (raster.wgsl:93-115 after the
    // payload is built); the payload depends only on t, not on the pixel.
    __device__ __forceinline__ void put(int32_t x, int32_t y, int32_t width, float clip_depth, uint64_t payload,
                                        bool center_hit, uint32_t beta_word) {
        float scene = scene_at(x, y);
        if (p.pass_mode[2] == 1u) {
            if (clip_depth < scene) return;
        } else if (clip_depth > scene) {
            return;
        }
        size_t pix = (size_t)(uint32_t)y * (uint32_t)width + (uint32_t)x;
        if (COUNT) {
            c.cam_writes++;
            if (center_hit) c.cam_centers++;
            return;
        }
        gmin64(a.cons + pix, payload);
        if (center_hit) {
            gmin64(a.center + pix, payload);
            gmin32(a.beta + pix, beta_word);
        }
    }

    __device__ void segment(F3 p0, F3 p1, F3 t0, F3 t1, F3 uvw0, F3 uvw1, float ao0, float ao1, float beta) {
        // raster.wgsl:192-223
        Projected A = project_world(p, p0);
        Projected B = project_world(p, p1);
        if (!A.in_front && !B.in_front) return;
        if (COUNT) c.segments++;
        F2 delta = B.screen - A.screen;
        int32_t steps = to_i32(ceilf(mx(fabsf(delta.x), fabsf(delta.y))));
        int32_t n = steps < 1 ? 1 : (steps > 96 ? 96 : steps);
        float diameter = mx(p.appearance.w, 0.05f);
        const int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        F2 seg = B.screen - A.screen;
        float denom = mx(dot2(seg, seg), 1e-6f);
        uint32_t beta_q = to_u32(clampf(beta, 0.0f, 1.0f) * 255.0f + 0.5f);
        int32_t ppx = 0, ppy = 0;
        for (int32_t i = 0; i <= n; ++i) {
            float t = rdiv(i2f(i), i2f(n));
            F2 screen = mix2(A.screen, B.screen, t);
            int32_t px = to_i32(floorf(screen.x)), py = to_i32(floorf(screen.y));
            bool diag = i > 0 && ppx != px && ppy != py;
            bool in0 = px >= 0 && py >= 0 && px < width && py < height;
            bool in1 = diag && ppx >= 0 && py >= 0 && ppx < width && py < height;
            bool in2 = diag && px >= 0 && ppy >= 0 && px < width && ppy < height;
            if (in0 || in1 || in2) {
                F3 world = mix3(p0, p1, t);
                float vd = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
                if (vd > p.screen.z) {
                    float cz = mat_row(p.clip, 2, world.x, world.y, world.z, 1.0f);
                    float cw = mat_row(p.clip, 3, world.x, world.y, world.z, 1.0f);
                    float clip_depth = rdiv(cz, mx(cw, 1e-5f));
                    F3 tangent = safe_normalize(mix3(t0, t1, t));
                    F3 uvw = mix3(uvw0, uvw1, t);
                    float ao = mixs(ao0, ao1, t);
                    uint32_t depth = quantize_depth(vd, p.screen.z, p.screen.w);
                    uint32_t ox, oy;
                    oct_encode(tangent, ox, oy);
                    uint32_t ao_q = to_u32(clampf(ao, 0.0f, 1.0f) * 63.0f + 0.5f);
                    if (p.flags[2] == 0u) ao_q = 63u;
                    uint64_t payload = pack_gb(depth, ox, oy, quant63(uvw.x), quant63(uvw.y), quant63(uvw.z), ao_q);
                    uint32_t beta_word = (depth << 8u) | beta_q;
                    if (in0) {
                        F2 center = {i2f(px) + 0.5f, i2f(py) + 0.5f};
                        float closest = clampf(rdiv(dot2(center - A.screen, seg), denom), 0.0f, 1.0f);
                        float dist = len2(center - (A.screen + seg * closest));
                        bool hit = dist <= diameter * 0.5f;
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

    __device__ __forceinline__ void dom_pixel(int32_t x, int32_t y, F3 world) {  // raster.wgsl:118-148
        int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
        if (x < 0 || y < 0 || x >= width || y >= height) return;
        float near_ = p.screen.z, far_ = p.screen.w;
        float dist = dot3(world - xyz(p.camera_pos), xyz(p.camera_forward));
        if (dist <= near_ || dist >= far_) return;
        uint32_t dom_w = to_u32(p.dom_info.x);
        size_t pix = (size_t)(uint32_t)y * dom_w + (uint32_t)x;
        uint32_t q = quantize_depth(dist, near_, far_);
        if (MODE == 1) {
            if (COUNT) c.dom_writes++;
            else gmin32(a.dom + pix, q);
            return;
        }
        uint32_t front_q = a.dom[pix];  // final: written by the previous (depth) pass
        if (front_q == HM_DOM_EMPTY) return;
        float z_front = dequantize_depth(front_q, near_, far_);
        float thickness = rdiv(mx(far_ - near_, 1e-3f), 16.0f);
        float t = rdiv(mx(dist - z_front, 0.0f), thickness);
        uint32_t slice = to_u32(t);
        slice = slice < 15u ? slice : 15u;
        size_t pixels = (size_t)dom_w * to_u32(p.dom_info.y);
        if (COUNT) c.dom_writes++;
        else atomicAdd(a.dom + pixels + pix * HM_DOM_LAYERS + slice, 1u);
    }

    __device__ void dom_segment(F3 p0, F3 p1, bool include_start) {  // raster.wgsl:150-190
        Projected A = project_world(p, p0);
        Projected B = project_world(p, p1);
        if (!A.in_front && !B.in_front) return;
        if (COUNT) c.segments++;
        int32_t x0 = to_i32(floorf(A.screen.x)), y0 = to_i32(floorf(A.screen.y));
        int32_t x1 = to_i32(floorf(B.screen.x)), y1 = to_i32(floorf(B.screen.y));
        int32_t dx = wabs(wsub(x1, x0)), dy = wabs(wsub(y1, y0));
        int32_t sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
        int32_t err = wsub(dx, dy);
        F2 seg = B.screen - A.screen;
        float seg_len2 = mx(dot2(seg, seg), 1e-6f);
        bool skip = !include_start && (x0 != x1 || y0 != y1);
        int32_t ndy = wsub(0, dy);
        for (int step = 0; step < 1024; ++step) {
            if (!skip) {
                F2 center = {i2f(x0) + 0.5f, i2f(y0) + 0.5f};
                float t = clampf(rdiv(dot2(center - A.screen, seg), seg_len2), 0.0f, 1.0f);
                dom_pixel(x0, y0, mix3(p0, p1, t));
            }
            skip = false;
            if (x0 == x1 && y0 == y1) break;
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

template <int MODE, bool COUNT>
__global__ __launch_bounds__(256) void k_raster(RasterArgs a) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t waves_per_block = blockDim.x >> 6;
    const uint32_t wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + (threadIdx.x >> 6));
    const uint32_t nwaves = gridDim.x * waves_per_block;
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    Drawer<MODE, COUNT> dr{p, a, {0, 0, 0, 0}};

    for (uint32_t s = wave0; s < total; s += nwaves) {
        const hm_strand_ref ref = a.refs[s];
        const hm_bundle bundle = a.bundles[ref.bundle];
        const F2 uv = pcg_uv(ref.strand ^ (ref.bundle * 0x85ebca6bu));

        // Lanes 0..15 sample layer `lane` of the cage (raster.wgsl:22-32;
        // only pos_ao is kept: the layer tangent is recomputed below).
        // Lanes 16..31 sample level `lane-16` of the style volume at this uv
        // (the u/v part of raster.wgsl:34-60; the w blend happens per point).
        F4 lay = {0, 0, 0, 0};
        F3 sty = {0, 0, 0};
        if (lane < 16) {
            const hm_layer* L = a.layers + bundle.layer_offset + (uint32_t)lane * 4u;
            lay = bilinear4(ld4(L[0].pos_ao), ld4(L[1].pos_ao), ld4(L[2].pos_ao), ld4(L[3].pos_ao), uv);
        } else if (lane < 32) {
            uint32_t k = (uint32_t)lane - 16u;
            float u = clampf(uv.x, 0.0f, 1.0f) * 7.0f;
            float v = clampf(uv.y, 0.0f, 1.0f) * 7.0f;
            uint32_t u0 = to_u32(floorf(u)), v0 = to_u32(floorf(v));
            uint32_t u1 = u0 + 1u < 7u ? u0 + 1u : 7u;
            uint32_t v1 = v0 + 1u < 7u ? v0 + 1u : 7u;
            float tu = u - u2f(u0), tv = v - u2f(v0);
            const float* S = a.style;
            uint32_t off = bundle.style_offset + k * 64u;
            F3 c000 = xyz(ld4(S + 4u * (off + v0 * 8u + u0)));
            F3 c100 = xyz(ld4(S + 4u * (off + v0 * 8u + u1)));
            F3 c010 = xyz(ld4(S + 4u * (off + v1 * 8u + u0)));
            F3 c110 = xyz(ld4(S + 4u * (off + v1 * 8u + u1)));
            sty = mix3(mix3(c000, c100, tu), mix3(c010, c110, tu), tv);
        }
        sty = shfl3(sty, (lane + 16) & 63);  // style level k now sits in lane k

        // Layer tangents (raster.wgsl:244-267).
        F3 lp = xyz(lay);
        F3 prev_l = shfl3(lp, lane > 0 ? lane - 1 : 0);
        F3 next_l = shfl3(lp, lane < 15 ? lane + 1 : 15);
        F3 ltan = calc_tangent((uint32_t)lane, HM_LAYER_COUNT, prev_l, lp, next_l);

        // Control points idx = lane and lane + 64 (raster.wgsl:270-284).
        const uint32_t ncp = ref.control_points < 2u ? 2u : (ref.control_points > 127u ? 127u : ref.control_points);
        const float inv_den = u2f(ncp - 1u > 1u ? ncp - 1u : 1u);
        F3 cp[2], cw[2];
        float cp_w[2], cp_ao[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint32_t idx = (uint32_t)lane + 64u * (uint32_t)j;
            float w = rdiv(u2f(idx), inv_den);
            float layers_f = w * 15.0f;
            uint32_t i0 = to_u32(floorf(layers_f));
            i0 = i0 < 15u ? i0 : 15u;
            uint32_t i1 = i0 + 1u < 15u ? i0 + 1u : 15u;
            float t = layers_f - u2f(i0);
            F4 a0 = shfl4(lay, (int)i0), a1 = shfl4(lay, (int)i1);
            F3 m0 = shfl3(ltan, (int)i0), m1 = shfl3(ltan, (int)i1);
            F3 base = hermite(xyz(a0), xyz(a1), m0, m1, t);
            float ao = mixs(a0.w, a1.w, t);
            // Style: w in [0, 1], so its level index and blend equal i0/i1/t.
            F3 s0 = shfl3(sty, (int)i0), s1 = shfl3(sty, (int)i1);
            F3 residual = mix3(s0, s1, t);
            cp[j] = base + residual;
            cp_w[j] = w;
            cp_ao[j] = ao;
        }

        // Control-point tangents (raster.wgsl:286-317).
        F3 up0 = shfl3(cp[0], (lane + 63) & 63), up1 = shfl3(cp[1], (lane + 63) & 63);
        F3 dn0 = shfl3(cp[0], (lane + 1) & 63), dn1 = shfl3(cp[1], (lane + 1) & 63);
        F3 ctan[2];
        {
            uint32_t idx = (uint32_t)lane;
            F3 prev = lane > 0 ? up0 : cp[0];
            F3 next = idx + 1u < ncp ? (lane < 63 ? dn0 : dn1) : cp[0];
            ctan[0] = safe_normalize(calc_tangent(idx, ncp, prev, cp[0], next));
            idx += 64u;
            prev = lane > 0 ? up1 : up0;
            next = idx + 1u < ncp ? dn1 : cp[1];
            ctan[1] = safe_normalize(calc_tangent(idx, ncp, prev, cp[1], next));
        }

        const float beta = rdiv(u2f(ref.n_lod), u2f(bundle.strand_count > 1u ? bundle.strand_count : 1u));
#pragma unroll
        for (int j = 0; j < 2; ++j) cw[j] = to_world(p, cp[j]);
        // Segment idx -> idx+1: the end point lives in the next lane (slot 0 of
        // lane 63 continues in slot 1 of lane 0).
        F3 n0 = shfl3(cw[0], (lane + 1) & 63), n1 = shfl3(cw[1], (lane + 1) & 63);
        F3 e0 = lane < 63 ? n0 : n1;
        if (MODE != 0) {
            if ((uint32_t)lane + 1u < ncp) dr.dom_segment(cw[0], e0, lane == 0);
            if ((uint32_t)lane + 65u < ncp) dr.dom_segment(cw[1], n1, false);
        } else {
            F3 wt0 = dir_world(p, ctan[0]), wt1 = dir_world(p, ctan[1]);
            F3 nt0 = shfl3(wt0, (lane + 1) & 63), nt1 = shfl3(wt1, (lane + 1) & 63);
            float nw0 = __shfl(cp_w[0], (lane + 1) & 63), nw1 = __shfl(cp_w[1], (lane + 1) & 63);
            float na0 = __shfl(cp_ao[0], (lane + 1) & 63), na1 = __shfl(cp_ao[1], (lane + 1) & 63);
            if ((uint32_t)lane + 1u < ncp) {
                bool wrap = lane == 63;
                dr.segment(cw[0], e0, wt0, wrap ? nt1 : nt0, F3{uv.x, uv.y, cp_w[0]},
                           F3{uv.x, uv.y, wrap ? nw1 : nw0}, cp_ao[0], wrap ? na1 : na0, beta);
            }
            if ((uint32_t)lane + 65u < ncp)
                dr.segment(cw[1], n1, wt1, nt1, F3{uv.x, uv.y, cp_w[1]}, F3{uv.x, uv.y, nw1}, cp_ao[1], na1, beta);
        }
    }
    if (COUNT) {
        unsigned long long v[4] = {dr.c.cam_writes, dr.c.cam_centers, dr.c.dom_writes, dr.c.segments};
        for (int k = 0; k < 4; ++k)
            if (v[k]) atomicAdd((unsigned long long*)a.counters + k, v[k]);
    }
}

// ---- shade.wesl -----------------------------------------------------------------

struct Lobes { F3 r, tt, trt; };
struct Shape { F3 alpha, beta; };
struct DualScatter { Lobes fwd; F3 a_f, beta_f2, a_back, delta_b, sigma_b2; };

__device__ __forceinline__ F3 vexp(F3 a) { return {expf(a.x), expf(a.y), expf(a.z)}; }
__device__ __forceinline__ F3 vsqrt(F3 a) { return {sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }

__device__ __forceinline__ F3 gaussian3(F3 b, F3 x) {
    F3 bb = vmax(b, splat3(1e-3f));
    const float s2pi = sqrtf(2.0f * kPi);
    return vexp(-0.5f * x * x / (bb * bb)) / (s2pi * bb);
}

__device__ __forceinline__ float fresnel(float cos_t) {
    const float ior = 1.55f;
    float r0 = (1.0f - ior) / (1.0f + ior);
    float f0 = r0 * r0;
    float x = 1.0f - clampf(cos_t, 0.0f, 1.0f);
    float x2 = x * x;
    return f0 + (1.0f - f0) * (x2 * x2 * x);
}

__device__ __forceinline__ F3 absorption_from_color(F3 color, float roughness) {
    float b = clampf(roughness, 0.0f, 1.0f);
    float b2 = b * b, b3 = b2 * b;
    float d = 5.969f - 0.215f * b + 2.532f * b2 - 10.73f * b3 + 5.574f * b2 * b2 + 0.245f * b3 * b2;
    F3 c = vclamp(color, 1e-4f, 1.0f);
    F3 l = {logf(c.x) / d, logf(c.y) / d, logf(c.z) / d};
    return l * l;
}

__device__ __forceinline__ Lobes azimuthal(float cos_theta_d, float cos_phi, F3 sigma_a) {
    float cos_half_phi = sqrtf(clampf(0.5f + 0.5f * cos_phi, 0.0f, 1.0f));
    float n_r = 0.25f * cos_half_phi * fresnel(cos_theta_d * cos_half_phi);
    float eta = 1.19f / cos_theta_d + 0.36f * cos_theta_d;
    float a = 1.0f / eta;
    float h = cos_half_phi * (1.0f + a * (0.6f - 0.8f * cos_phi));
    float f_tt = fresnel(cos_theta_d * sqrtf(clampf(1.0f - h * h, 0.0f, 1.0f)));
    float s_tt = sqrtf(clampf(1.0f - h * h * a * a, 0.0f, 1.0f));
    F3 t_tt = vexp(sigma_a * -1.0f * 2.0f * s_tt / cos_theta_d);
    F3 n_tt = t_tt * (expf(-3.65f * cos_phi - 3.98f) * (1.0f - f_tt) * (1.0f - f_tt));
    float f_trt = fresnel(cos_theta_d * 0.5f);
    F3 t_trt = vexp(sigma_a * -1.0f * 3.2f / cos_theta_d);
    F3 n_trt = t_trt * (expf(17.0f * cos_phi - 16.78f) * (1.0f - f_trt) * (1.0f - f_trt) * f_trt);
    return {splat3(n_r), n_tt, n_trt};
}

__device__ __forceinline__ F3 lobe_mean(const Lobes& l, F3 total, F3 q) {
    return (l.r * q.x + l.tt * q.y + l.trt * q.z) / vmax(total, splat3(1e-5f));
}

__device__ DualScatter dual_scatter(float cos_theta_d, F3 sigma_a, const Shape& shape) {
    Lobes fwd = {splat3(0), splat3(0), splat3(0)}, back = fwd;
    for (uint32_t i = 0; i < 8u; ++i) {
        float u = (u2f(i) + 0.5f) / 8.0f;
        Lobes b = azimuthal(cos_theta_d, cosf(u * 0.5f * kPi), sigma_a);
        Lobes f = azimuthal(cos_theta_d, cosf((1.0f + u) * 0.5f * kPi), sigma_a);
        back = {back.r + b.r, back.tt + b.tt, back.trt + b.trt};
        fwd = {fwd.r + f.r, fwd.tt + f.tt, fwd.trt + f.trt};
    }
    float w = kPi / 8.0f;
    fwd = {fwd.r * w, fwd.tt * w, fwd.trt * w};
    back = {back.r * w, back.tt * w, back.trt * w};
    F3 a_f = vmin(fwd.r + fwd.tt + fwd.trt, splat3(0.95f));
    F3 a_b = back.r + back.tt + back.trt;
    F3 beta2 = shape.beta * shape.beta;
    F3 beta_f2 = lobe_mean(fwd, a_f, beta2);
    F3 alpha_b = lobe_mean(back, a_b, shape.alpha);
    F3 beta_b2 = lobe_mean(back, a_b, beta2);
    F3 af2 = a_f * a_f;
    F3 rest = splat3(1.0f) - af2;
    F3 ab2 = a_b * a_b, ab3 = ab2 * a_b;
    F3 rest3 = rest * rest * rest;
    F3 a_back = a_b * af2 / rest + ab3 * af2 / rest3;
    F3 delta_b = alpha_b * (splat3(1.0f) - 2.0f * ab2 / (rest * rest)) +
                 alpha_b * (2.0f * rest * rest + 4.0f * af2 * ab2) / rest3;
    F3 beta_f = vsqrt(beta_f2), beta_b = vsqrt(beta_b2);
    F3 sigma_b = (splat3(1.0f) + 0.7f * af2) *
                 (a_b * vsqrt(2.0f * beta_f2 + beta_b2) + ab3 * vsqrt(2.0f * beta_f2 + 3.0f * beta_b2)) /
                 vmax(a_b + ab3 * (2.0f * beta_f + 3.0f * beta_b), splat3(1e-5f));
    return {fwd, a_f, beta_f2, a_back, delta_b, sigma_b * sigma_b};
}

__device__ F3 hair_bsdf(F3 t, F3 v, F3 light, F3 sigma_a, const Shape& shape, const DualScatter& ds, float depth) {
    // t and v arrive normalized (safe_normalize is idempotent on them up to rounding).
    t = safe_normalize(t);
    v = safe_normalize(v);
    F3 l = safe_normalize(light);
    float sin_i = clampf(dot3(t, l), -1.0f, 1.0f);
    float sin_o = clampf(dot3(t, v), -1.0f, 1.0f);
    float theta_i = asinf(sin_i), theta_o = asinf(sin_o);
    float theta_h = theta_i + theta_o;
    float cos_theta_d = mx(cosf(0.5f * fabsf(theta_i - theta_o)), 1e-3f);
    F3 t_i = l - t * sin_i;
    F3 t_o = v - t * sin_o;
    float proj = sqrtf(mx(dot3(t_i, t_i) * dot3(t_o, t_o), 1e-8f));
    float cos_phi = clampf(dot3(t_i, t_o) / proj, -1.0f, 1.0f);
    Lobes n = azimuthal(cos_theta_d, cos_phi, sigma_a);
    F3 m = gaussian3(shape.beta, splat3(theta_h) - shape.alpha);
    F3 single = m.x * n.r + m.y * n.tt + m.z * n.trt;
    F3 back = 2.0f * ds.a_back * gaussian3(vsqrt(ds.sigma_b2), splat3(theta_h) - ds.delta_b) / kPi;
    float direct = expf(-depth);
    F3 color = direct * (single + 0.7f * back);
    if (depth > 1e-3f) {
        F3 sigma_f2 = depth * ds.beta_f2;
        F3 transmit = {powf(ds.a_f.x, depth), powf(ds.a_f.y, depth), powf(ds.a_f.z, depth)};
        F3 beta2 = shape.beta * shape.beta;
        F3 x = splat3(theta_h);
        F3 spread = gaussian3(vsqrt(splat3(beta2.x) + sigma_f2), x - splat3(shape.alpha.x)) * ds.fwd.r +
                    gaussian3(vsqrt(splat3(beta2.y) + sigma_f2), x - splat3(shape.alpha.y)) * ds.fwd.tt +
                    gaussian3(vsqrt(splat3(beta2.z) + sigma_f2), x - splat3(shape.alpha.z)) * ds.fwd.trt;
        F3 back_spread = 2.0f * ds.a_back * gaussian3(vsqrt(ds.sigma_b2 + sigma_f2), x - ds.delta_b) / kPi;
        color = color + (1.0f - direct) * transmit * 0.7f * (spread / kPi + kPi * 0.7f * back_spread);
    }
    return color;
}

__device__ __forceinline__ float dom_texel_visibility(const View& p, const uint32_t* __restrict__ dom, int32_t x,
                                                      int32_t y, float dist, float near_, float far_, uint32_t dom_w,
                                                      uint32_t dom_h) {
    if (x < 0 || y < 0 || x >= (int32_t)dom_w || y >= (int32_t)dom_h) return 1.0f;
    uint32_t pix = (uint32_t)y * dom_w + (uint32_t)x;
    uint32_t front_q = dom[pix];
    if (front_q == HM_DOM_EMPTY) return 1.0f;
    float z_front = dequantize_depth(front_q, near_, far_);
    float thickness = rdiv(mx(far_ - near_, 1e-3f), 16.0f);
    float rel = mx(dist - z_front - thickness * 0.05f, 0.0f);
    float t = rdiv(rel, thickness);
    const uint4* c4 = reinterpret_cast<const uint4*>(dom + (size_t)dom_w * dom_h + (size_t)pix * HM_DOM_LAYERS);
    float occ = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint4 c = c4[q];
        uint32_t cs[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float i = (float)(q * 4 + k);
            float count = u2f(cs[k]);
            if (i + 1.0f <= t) occ = occ + count;
            else if (i < t) occ = occ + count * (t - i);
        }
    }
    return expf(-occ * p.filter_params.w);
}

__device__ float dom_visibility(const View& p, const uint32_t* __restrict__ dom, F3 world, float beta) {
    if (p.flags[3] == 0u) return 1.0f;
    float cx = mat_row(p.lclip, 0, world.x, world.y, world.z, 1.0f);
    float cy = mat_row(p.lclip, 1, world.x, world.y, world.z, 1.0f);
    float cw = mat_row(p.lclip, 3, world.x, world.y, world.z, 1.0f);
    if (cw <= 0.0f) return 1.0f;
    float ndc_x = rdiv(cx, cw), ndc_y = rdiv(cy, cw);
    float ux = ndc_x * 0.5f + 0.5f;
    float uy = 1.0f - (ndc_y * 0.5f + 0.5f);
    if (ux < 0.0f || uy < 0.0f || ux > 1.0f || uy > 1.0f) return 1.0f;
    uint32_t dom_w = to_u32(p.dom_info.x), dom_h = to_u32(p.dom_info.y);
    float near_ = p.dom_info.z, far_ = p.dom_info.w;
    float dist = dot3(world - xyz(p.light_eye), xyz(p.light_forward)) - (-logf(mx(beta, 1e-4f)) * p.filter_params.z);
    float fx = ux * u2f(dom_w) - 0.5f;
    float fy = uy * u2f(dom_h) - 0.5f;
    float flx = floorf(fx), fly = floorf(fy);
    int32_t x0 = to_i32(flx), y0 = to_i32(fly);
    float tx = fx - flx, ty = fy - fly;
    float v00 = dom_texel_visibility(p, dom, x0, y0, dist, near_, far_, dom_w, dom_h);
    float v10 = dom_texel_visibility(p, dom, x0 + 1, y0, dist, near_, far_, dom_w, dom_h);
    float v01 = dom_texel_visibility(p, dom, x0, y0 + 1, dist, near_, far_, dom_w, dom_h);
    float v11 = dom_texel_visibility(p, dom, x0 + 1, y0 + 1, dist, near_, far_, dom_w, dom_h);
    return mixs(mixs(v00, v10, tx), mixs(v01, v11, tx), ty);
}

__device__ __forceinline__ float distance_attenuation(float dist2, float inv_range2) {
    float factor = dist2 * inv_range2;
    float smooth = clampf(1.0f - factor * factor, 0.0f, 1.0f);
    return smooth * smooth * 1.0f / mx(dist2, 1e-4f);
}

__device__ __forceinline__ float spot_mask(const hm_point_light& L, F3 world) {
    F3 sd = {L.light_custom_data[0], 0.0f, L.light_custom_data[1]};
    sd.y = sqrtf(mx(0.0f, 1.0f - sd.x * sd.x - sd.z * sd.z));
    if (L.flags & HM_LIGHT_FLAG_SPOT_Y_NEGATIVE) sd.y = -sd.y;
    F3 ltf = F3{L.position_radius[0], L.position_radius[1], L.position_radius[2]} - world;
    float cd = dot3(sd * -1.0f, safe_normalize(ltf));
    float att = clampf(cd * L.light_custom_data[2] + L.light_custom_data[3], 0.0f, 1.0f);
    return att * att;
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    uint32_t width = to_u32(p.screen.x), height = to_u32(p.screen.y);
    if (x >= width || y >= height) return;
    size_t pix = (size_t)y * width + x;
    uint64_t word = a.center[pix];
    if (gb_empty(word)) {
        a.shaded[pix] = make_float4(0, 0, 0, 0);
        a.shaded_depth[pix] = 0.0f;
        return;
    }
    uint32_t low = (uint32_t)word;
    float view_dist = dequantize_depth(gb_depth(word), p.screen.z, p.screen.w);
    F3 world = reconstruct_world(p, (int32_t)x, (int32_t)y, view_dist);
    F3 tangent = oct_decode(gb_tx(word), gb_ty(word));
    float uvw_z = u2f((low >> 6u) & 0x3fu) / 63.0f;
    float beta = u2f(a.beta[pix] & 0xffu) / 255.0f;
    float roughness = mx(p.appearance.x * (0.65f + 0.7f * uvw_z), 0.04f);
    float tilt = p.appearance.y;
    F3 albedo = xyz(p.albedo) * (0.65f + 0.35f * uvw_z);
    bool ortho = p.clip[3] == 0.0f && p.clip[7] == 0.0f && p.clip[11] == 0.0f && p.clip[15] == 1.0f;
    F3 view = ortho ? safe_normalize(F3{p.clip[2], p.clip[6], p.clip[10]}) : safe_normalize(xyz(p.camera_pos) - world);

    float bb = clampf(roughness * roughness, 0.02f, 1.0f);
    Shape shape = {{tilt, -0.5f * tilt, -1.5f * tilt}, {bb, 0.5f * bb, 2.0f * bb}};
    float tv = dot3(tangent, view);
    float cos_theta_o = sqrtf(mx(1.0f - tv * tv, 1e-6f));
    F3 sigma_a = absorption_from_color(albedo, roughness);
    DualScatter ds = dual_scatter(cos_theta_o, sigma_a, shape);
    F3 color = splat3(0.0f);
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t i = 0; i < a.point_count; ++i) {
            const hm_point_light& L = a.points[i];
            if ((pass == 0) != (L.kind == HM_LIGHT_POINT)) continue;
            F3 ltf = F3{L.position_radius[0], L.position_radius[1], L.position_radius[2]} - world;
            float dist2 = dot3(ltf, ltf);
            float att = distance_attenuation(dist2, L.color_inverse_square_range[3]);
            if (att == 0.0f) continue;  // outside the light's range: the cluster would not list it
            F3 incident = safe_normalize(ltf);
            F3 lobe = hair_bsdf(tangent, view, incident, sigma_a, shape, ds, 0.0f);
            F3 c = lobe * F3{L.color_inverse_square_range[0], L.color_inverse_square_range[1],
                             L.color_inverse_square_range[2]} * att;
            if (pass == 1) c = c * spot_mask(L, world);
            color = color + c;
        }
    }
    bool dom_left = true;
    for (uint32_t i = 0; i < a.directional_count; ++i) {
        const hm_directional_light& L = a.directionals[i];
        bool casts = (L.flags & HM_LIGHT_FLAG_SHADOWS) != 0u;
        float depth = 0.0f;
        if (dom_left && casts) {
            depth = -logf(mx(dom_visibility(p, a.dom, world, beta), 1e-6f));
            dom_left = false;
        }
        F3 lobe = hair_bsdf(tangent, view, F3{L.direction_to_light[0], L.direction_to_light[1], L.direction_to_light[2]},
                            sigma_a, shape, ds, depth);
        color = color + lobe * F3{L.color[0], L.color[1], L.color[2]};
    }
    color = color * a.exposure;
    a.shaded[pix] = make_float4(color.x, color.y, color.z, 1.0f);
    a.shaded_depth[pix] = view_dist;
}

// ---- filter.wgsl ----------------------------------------------------------------

__global__ __launch_bounds__(256) void k_filter(FilterArgs a) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    int32_t x = (int32_t)(blockIdx.x * 16u + (threadIdx.x & 15u));
    int32_t y = (int32_t)(blockIdx.y * 16u + (threadIdx.x >> 4));
    int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
    if (x >= width || y >= height) return;
    size_t pix = (size_t)y * (uint32_t)width + (uint32_t)x;
    uint64_t center_word = a.center[pix];
    uint64_t cons_word = a.cons[pix];
    bool center_hit = !gb_empty(center_word), cons_hit = !gb_empty(cons_word);
    if (!center_hit && !cons_hit) {
        a.filtered[pix] = make_float4(0, 0, 0, 0);
        a.filtered_depth[pix] = 0.0f;
        return;
    }
    uint64_t w0 = center_hit ? center_word : cons_word;
    F3 tangent = oct_decode(gb_tx(w0), gb_ty(w0));
    uint32_t depth_q = gb_depth(w0);
    float view_dist = dequantize_depth(depth_q, p.screen.z, p.screen.w);
    F3 world = reconstruct_world(p, x, y, view_dist);
    // screen_tangent (filter.wgsl:9-18)
    F2 ts;
    {
        Projected A = project_world(p, world);
        Projected B = project_world(p, world + tangent);
        F2 d = B.screen - A.screen;
        float l = len2(d);
        ts = l < 1e-3f ? F2{1.0f, 0.0f} : F2{rdiv(d.x, l), rdiv(d.y, l)};
    }
    float4 cc = a.shaded[pix];
    const int radius = 5;
    float s_par = mx(p.filter_params.x, 0.05f), s_perp = mx(p.filter_params.y, 0.02f);
    float sigma_par = 5.0f * s_par, sigma_perp = 5.0f * s_perp;
    float spar2 = sigma_par * sigma_par, sperp2 = sigma_perp * sigma_perp;
    const float sigma_c2 = 0.9f * 0.9f;
    float p_norm = rdiv(u2f(depth_q), 16777214.0f);
    bool promote = !center_hit;
    bool use_color = !promote && cc.w > 0.0f;
    F3 accum = splat3(0.0f);
    float weight = 0.0f, h_center = 0.0f, h_cons = 0.0f;
    for (int dy = -radius; dy <= radius; ++dy) {
        int32_t qy = y + dy;
        if (qy < 0 || qy >= height) continue;
        for (int dx = -radius; dx <= radius; ++dx) {
            int32_t qx = x + dx;
            if (qx < 0 || qx >= width) continue;
            size_t qpix = (size_t)qy * (uint32_t)width + (uint32_t)qx;
            F2 d = {i2f(dx), i2f(dy)};
            float d_par = dot2(d, ts);
            float d_perp = len2(d - ts * d_par);
            uint64_t q_center = a.center[qpix];
            bool qc_empty = gb_empty(q_center);
            if (d_perp <= 0.5f) {
                if (!qc_empty) h_center = h_center + 1.0f;
                else if (!gb_empty(a.cons[qpix])) h_cons = h_cons + 1.0f;
            }
            if (qc_empty) continue;
            float q_norm = rdiv(u2f(gb_depth(q_center)), 16777214.0f);
            if (fabsf(q_norm - p_norm) > 1.45e-3f) continue;
            float spatial = expf(-d_par * d_par / spar2 - d_perp * d_perp / sperp2);
            float4 qc = a.shaded[qpix];
            F3 q_color = {qc.x, qc.y, qc.z};
            float color_w = 1.0f;
            if (use_color) {
                F3 dc = F3{cc.x, cc.y, cc.z} - q_color;
                float chroma = len3(dc);
                color_w = expf(-chroma * chroma / sigma_c2);
            }
            float w = spatial * color_w;
            accum = accum + q_color * w;
            weight = weight + w;
        }
    }
    F3 color = {cc.x, cc.y, cc.z};
    if (weight > 1e-5f) color = accum / weight;
    float coverage = h_center / mx(h_center + h_cons, 1.0f);
    float cons_depth = cons_hit ? rdiv(u2f(gb_depth(cons_word)), 16777214.0f) : 1.0f;
    bool closer = cons_hit && center_hit && (p_norm - cons_depth) >= 5e-4f;
    float alpha = closer ? 1.0f : clampf(coverage, 0.0f, 1.0f);
    if (promote && weight <= 1e-5f) {
        alpha = 0.0f;
        color = splat3(0.0f);
    }
    a.filtered[pix] = make_float4(color.x, color.y, color.z, alpha);
    a.filtered_depth[pix] = view_dist;
}

}  // namespace hm

// ---- launchers ----------------------------------------------------------------

namespace hm {

hipError_t launch_fill32(hipStream_t s, uint32_t* dst, size_t words, uint32_t value, int max_blocks) {
    if (words == 0) return hipSuccess;
    size_t quads = (words + 3) / 4;
    size_t blocks = (quads + 255) / 256;
    if (blocks > (size_t)max_blocks) blocks = (size_t)max_blocks;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)blocks), dim3(256), 0, s, dst, words, value);
    return hipGetLastError();
}

hipError_t launch_lod(hipStream_t s, const hm_params& prm, const hm_bundle* bundles, const hm_layer* layers,
                      hm_lod_record* out, uint32_t nb, uint32_t ovr_cps, float ovr_frac) {
    if (nb == 0) return hipSuccess;
    View v;
    std::memcpy(&v, &prm, sizeof(v));
    hipLaunchKernelGGL(k_lod, dim3((nb + 3) / 4), dim3(256), 0, s, v, bundles, layers, out, nb, ovr_cps, ovr_frac);
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t s, const hm_lod_record* lod, uint32_t nb, uint32_t capacity, uint32_t* prefix,
                       uint32_t* indirect, hm_strand_ref* refs) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, lod, nb, capacity, prefix, indirect);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || capacity == 0 || nb == 0) return e;
    hipLaunchKernelGGL(k_expand, dim3((capacity + 255) / 256), dim3(256), 0, s, lod, prefix, nb, indirect, refs);
    return hipGetLastError();
}

hipError_t launch_raster(hipStream_t s, const RasterArgs& a, int mode, bool count, int blocks) {
    if (blocks < 1) blocks = 1;
    dim3 g((unsigned)blocks), b(256);
    if (count) {
        if (mode == 0) hipLaunchKernelGGL((k_raster<0, true>), g, b, 0, s, a);
        else if (mode == 1) hipLaunchKernelGGL((k_raster<1, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_raster<2, true>), g, b, 0, s, a);
    } else {
        if (mode == 0) hipLaunchKernelGGL((k_raster<0, false>), g, b, 0, s, a);
        else if (mode == 1) hipLaunchKernelGGL((k_raster<1, false>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_raster<2, false>), g, b, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const ShadeArgs& a, uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_filter(hipStream_t s, const FilterArgs& a, uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) return hipSuccess;
    hipLaunchKernelGGL(k_filter, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace hm
