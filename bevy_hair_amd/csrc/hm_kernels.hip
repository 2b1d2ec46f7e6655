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
#include "hm_shade.h"

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

// ---- microbenchmarks (roofline denominators) ----------------------------------

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Random-address atomics over `words` targets: kind 0 u64 min, 1 u32 min, 2 u32 add.
__global__ __launch_bounds__(256) void k_atomics(void* __restrict__ buf, size_t words, uint32_t per_thread,
                                                 uint32_t kind, uint32_t seed) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = pcg(id ^ seed);
    for (uint32_t k = 0; k < per_thread; ++k) {
        h = pcg(h);
        size_t i = (size_t)(((uint64_t)h * words) >> 32);
        if (kind == 0) atomicMin((unsigned long long*)buf + i, (unsigned long long)h);
        else if (kind == 1) atomicMin((uint32_t*)buf + i, h);
        else atomicAdd((uint32_t*)buf + i, 1u);
    }
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

// With `check`, a plain load screens the atomic: targets only decrease during
// a pass, so a value read here bounds the current one from above and a payload
// that is not below it cannot change the word.
__device__ __forceinline__ void gmin64(uint64_t* a, uint64_t v, bool check) {
    if (!check || v < *a) atomicMin((unsigned long long*)a, (unsigned long long)v);
}
__device__ __forceinline__ void gmin32(uint32_t* a, uint32_t v, bool check) {
    if (!check || v < *a) atomicMin(a, v);
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
        const bool chk = (a.opt & RASTER_PRECHECK) != 0u;
        gmin64(a.cons + pix, payload, chk);
        if (center_hit) {
            gmin64(a.center + pix, payload, chk);
            gmin32(a.beta + pix, beta_word, chk);
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
            else gmin32(a.dom + pix, q, (a.opt & RASTER_PRECHECK) != 0u);
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
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    // Strand schedule. Default: wave w takes strands w, w + W, ... With
    // RASTER_XCD, workgroup b runs on XCD b % 8 and each XCD takes one
    // contiguous eighth of the bundle-major strand list, so strands that hit
    // the same screen region share that XCD's L2.
    uint32_t first = 0, end = total;
    uint32_t wave0 = blockIdx.x * waves_per_block + (threadIdx.x >> 6);
    uint32_t nwaves = gridDim.x * waves_per_block;
    if ((a.opt & RASTER_XCD) && gridDim.x >= 8u) {
        uint32_t xcd = blockIdx.x & 7u;
        uint32_t blocks_here = (gridDim.x - xcd + 7u) / 8u;
        first = (uint32_t)((uint64_t)total * xcd / 8u);
        end = (uint32_t)((uint64_t)total * (xcd + 1u) / 8u);
        wave0 = first + (blockIdx.x >> 3) * waves_per_block + (threadIdx.x >> 6);
        nwaves = blocks_here * waves_per_block;
    }
    wave0 = __builtin_amdgcn_readfirstlane(wave0);
    Drawer<MODE, COUNT> dr{p, a, {0, 0, 0, 0}};

    for (uint32_t s = wave0; s < end; s += nwaves) {
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

// Covered-pixel lists: one pass over the G-buffer writes the empty result of
// every uncovered pixel and appends covered pixels to a list (one atomic per
// wave), so the shading and filter kernels run on full waves.
// mode 0: shade list (centre word present); mode 1: filter list (centre or
// conservative word present).
template <int MODE>
__global__ __launch_bounds__(256) void k_compact(const uint64_t* __restrict__ center, const uint64_t* __restrict__ cons,
                                                 uint32_t pixels, uint32_t* __restrict__ list,
                                                 uint32_t* __restrict__ count, float4* __restrict__ color_out,
                                                 float* __restrict__ depth_out) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    bool inb = pix < pixels;
    bool hit = false;
    if (inb) {
        hit = !gb_empty(center[pix]);
        if (MODE == 1 && !hit) hit = !gb_empty(cons[pix]);
        if (!hit) {
            color_out[pix] = make_float4(0, 0, 0, 0);
            depth_out[pix] = 0.0f;
        }
    }
    uint64_t mask = __ballot(hit);
    uint32_t lane = threadIdx.x & 63u;
    uint32_t total = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (total) {
        if (lane == 0) base = atomicAdd(count, total);
        base = __shfl(base, 0);
    }
    if (hit) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = pix;
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a, const uint32_t* __restrict__ list,
                                               const uint32_t* __restrict__ count) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t n = *count;
    uint32_t width = to_u32(p.screen.x);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t pix = list[i];
        uint32_t x = pix % width, y = pix / width;
        PixelShade px = pixel_setup(p, a.center[pix], a.beta[pix], x, y);
        F3 color = plain_pixel_color(a, p, px);
        a.shaded[pix] = make_float4(color.x, color.y, color.z, 1.0f);
        a.shaded_depth[pix] = px.view_dist;
    }
}

// ---- filter.wgsl ----------------------------------------------------------------

__device__ __forceinline__ void filter_pixel(const FilterArgs& a, const View& p, int32_t x, int32_t y) {
    int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
    size_t pix = (size_t)y * (uint32_t)width + (uint32_t)x;
    uint64_t center_word = a.center[pix];
    uint64_t cons_word = a.cons[pix];
    bool center_hit = !gb_empty(center_word), cons_hit = !gb_empty(cons_word);
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

__global__ __launch_bounds__(256) void k_filter(FilterArgs a, const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ count) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t n = *count;
    uint32_t width = to_u32(p.screen.x);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t pix = list[i];
        filter_pixel(a, p, (int32_t)(pix % width), (int32_t)(pix / width));
    }
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

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes, int blocks) {
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, bytes / 16);
    return hipGetLastError();
}

hipError_t launch_atomics(hipStream_t s, void* buf, size_t words, uint32_t threads, uint32_t per_thread, uint32_t kind) {
    hipLaunchKernelGGL(k_atomics, dim3((threads + 255) / 256), dim3(256), 0, s, buf, words, per_thread, kind, 12345u);
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

hipError_t launch_compact(hipStream_t s, int mode, const uint64_t* center, const uint64_t* cons, uint32_t pixels,
                          uint32_t* list, uint32_t* count, float4* color_out, float* depth_out) {
    if (pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill32, dim3(1), dim3(64), 0, s, count, (size_t)1, 0u);
    dim3 g((pixels + 255) / 256), b(256);
    if (mode == 0) hipLaunchKernelGGL(k_compact<0>, g, b, 0, s, center, cons, pixels, list, count, color_out, depth_out);
    else hipLaunchKernelGGL(k_compact<1>, g, b, 0, s, center, cons, pixels, list, count, color_out, depth_out);
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const ShadeArgs& a, const uint32_t* list, const uint32_t* count,
                        uint32_t pixels, int blocks) {
    if (pixels == 0) return hipSuccess;
    uint32_t want = (pixels + 255) / 256;
    hipLaunchKernelGGL(k_shade, dim3(want < (uint32_t)blocks ? want : (uint32_t)blocks), dim3(256), 0, s, a, list, count);
    return hipGetLastError();
}

hipError_t launch_filter(hipStream_t s, const FilterArgs& a, const uint32_t* list, const uint32_t* count,
                         uint32_t pixels, int blocks) {
    if (pixels == 0) return hipSuccess;
    uint32_t want = (pixels + 255) / 256;
    hipLaunchKernelGGL(k_filter, dim3(want < (uint32_t)blocks ? want : (uint32_t)blocks), dim3(256), 0, s, a, list,
                       count);
    return hipGetLastError();
}

}  // namespace hm
