// gfx950 kernels of the deferred strand rasterizer.
//
// Reference kernels (paths relative to the reference repository):
//   k_fill32          clear.wgsl:7-16, clear_dom.wgsl:8-21 (one fill per slab)
//   k_lod             lod.wgsl:6-59         one wave per bundle, lane = cage corner
//   k_scan/k_expand   scan.wgsl:6-24        parallel scan, same StrandRef order
//   k_raster<MODE>    raster.wgsl:225-337   one wave per strand, lane = segment
//   k_shade           shade.wesl:523-593    one thread per pixel
//   k_filter          filter.wgsl:20-135    one thread per listed pixel
//   k_filter_tiles    filter.wgsl:20-135    one workgroup per 16 x 16 tile, taps from LDS
#include "hm_kernels.h"

#include <cstring>

#include "hm_device.h"
#include "hm_lists.h"
#include "hm_raster.h"
#include "hm_shade.h"

#include <type_traits>

namespace hm {

static_assert(sizeof(View) == sizeof(hm_params), "View must mirror hm_params byte for byte");

// ---- fills ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill32(uint32_t* __restrict__ dst, size_t words, uint32_t value,
                                                const uint32_t* gate) {
    if (gate && *gate == 0u) return;
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

__global__ __launch_bounds__(256) void k_copy32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                size_t n) {
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

// Random-address LDS atomics within one workgroup's array of `words` targets
// (a power of two): KIND 0 u64 min, 1 u32 min, 2 u32 add. Each lane starts at
// a hashed target and steps by a fixed odd stride, so the addresses of a
// wave's lanes stay random (bank conflicts as in the tile kernels) while the
// address arithmetic costs two VALU operations per atomic.
template <int KIND>
__global__ __launch_bounds__(256) void k_lds_atomics(uint32_t* __restrict__ sink, uint32_t words, uint32_t per_thread) {
    extern __shared__ uint32_t lds[];
    const uint32_t n32 = KIND == 0 ? 2u * words : words;
    for (uint32_t i = threadIdx.x; i < n32; i += blockDim.x) lds[i] = KIND == 2 ? 0u : 0xffffffffu;
    __syncthreads();
    const uint32_t mask = words - 1u;
    uint32_t i = pcg(blockIdx.x * blockDim.x + threadIdx.x) & mask;
    const uint32_t v = pcg(i ^ 0x9e3779b9u);
    for (uint32_t k = 0; k < per_thread; ++k) {
        if (KIND == 0) atomicMin(reinterpret_cast<unsigned long long*>(lds) + i, (unsigned long long)v << 8);
        else if (KIND == 1) atomicMin(lds + i, v);
        else atomicAdd(lds + i, 1u);
        i = (i + 613u) & mask;
    }
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = lds[0] ^ lds[mask];
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

template <int MODE, bool COUNT>
__global__ __launch_bounds__(256) void k_raster(RasterArgs a) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t waves_per_block = blockDim.x >> 6;
    if (a.gate && *a.gate == 0u) return;
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    // Strand schedule. Default: wave w takes strands w, w + W, ... With
    // RASTER_XCD, workgroup b runs on XCD b % 8 and each XCD takes one
    // contiguous eighth of the bundle-major strand list, so strands that hit
    // the same screen region share that XCD's L2.
    uint32_t end = total;
    uint32_t wave0 = blockIdx.x * waves_per_block + (threadIdx.x >> 6);
    uint32_t nwaves = gridDim.x * waves_per_block;
    if ((a.opt & RASTER_XCD) && gridDim.x >= 8u) {
        uint32_t xcd = blockIdx.x & 7u;
        uint32_t blocks_here = (gridDim.x - xcd + 7u) / 8u;
        uint32_t first = (uint32_t)((uint64_t)total * xcd / 8u);
        end = (uint32_t)((uint64_t)total * (xcd + 1u) / 8u);
        wave0 = first + (blockIdx.x >> 3) * waves_per_block + (threadIdx.x >> 6);
        nwaves = blocks_here * waves_per_block;
    }
    wave0 = __builtin_amdgcn_readfirstlane(wave0);
    typedef typename std::conditional<COUNT, CountSink, GlobalSink>::type Sink;
    Sink sink;
    if constexpr (COUNT) {
        sink.dom = a.dom;
    } else {
        sink.center = a.center;
        sink.cons = a.cons;
        sink.beta = a.beta;
        sink.dom = a.dom;
        sink.dom_pixels = (size_t)to_u32(p.dom_info.x) * to_u32(p.dom_info.y);
        sink.check = (a.opt & RASTER_PRECHECK) != 0u;
    }
    Drawer<MODE, Sink> dr{p, a.scene_depth, a.scene_clear, sink, {0, 0, 0, 0}};

    for (uint32_t s = wave0; s < end; s += nwaves) {
        const hm_strand_ref ref = a.refs[s];
        StrandCps sc = gen_strand(p, ref, a.bundles, a.layers, a.style, lane);
        const uint32_t ncp = sc.ncp;
        // Segment idx -> idx + 1: the end point lives in the next lane (slot 0
        // of lane 63 continues in slot 1 of lane 0).
        F3 n0 = shfl3(sc.cw[0], (lane + 1) & 63), n1 = shfl3(sc.cw[1], (lane + 1) & 63);
        F3 e0 = lane < 63 ? n0 : n1;
        if (MODE != 0) {
            if ((uint32_t)lane + 1u < ncp) dr.dom_segment(sc.cw[0], e0, lane == 0);
            if ((uint32_t)lane + 65u < ncp) dr.dom_segment(sc.cw[1], n1, false);
        } else {
            F3 wt0 = dir_world(p, sc.ct[0]), wt1 = dir_world(p, sc.ct[1]);
            F3 nt0 = shfl3(wt0, (lane + 1) & 63), nt1 = shfl3(wt1, (lane + 1) & 63);
            float nw0 = __shfl(sc.w[0], (lane + 1) & 63), nw1 = __shfl(sc.w[1], (lane + 1) & 63);
            float na0 = __shfl(sc.ao[0], (lane + 1) & 63), na1 = __shfl(sc.ao[1], (lane + 1) & 63);
            const F2 uv = sc.uv;
            if ((uint32_t)lane + 1u < ncp) {
                bool wrap = lane == 63;
                dr.segment(sc.cw[0], e0, wt0, wrap ? nt1 : nt0, F3{uv.x, uv.y, sc.w[0]},
                           F3{uv.x, uv.y, wrap ? nw1 : nw0}, sc.ao[0], wrap ? na1 : na0, sc.beta);
            }
            if ((uint32_t)lane + 65u < ncp)
                dr.segment(sc.cw[1], n1, wt1, nt1, F3{uv.x, uv.y, sc.w[1]}, F3{uv.x, uv.y, nw1}, sc.ao[1], na1,
                           sc.beta);
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
                                                 uint32_t pixels, PixelList out, float4* __restrict__ color_out,
                                                 float* __restrict__ depth_out, const uint32_t* gate) {
    if (gate && *gate == 0u) return;
    __shared__ uint32_t scratch[5];
    const uint32_t sub = blockIdx.x % LIST_SUBS;
    // Grid-stride over whole blocks, so every thread runs the same trip count.
    for (uint32_t first = blockIdx.x * blockDim.x; first < pixels; first += gridDim.x * blockDim.x) {
        uint32_t pix = first + threadIdx.x;
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
        block_append(out, sub, hit, pix, scratch);
    }
}

// F: SHADE_* variant bits (pixel_color). The plain variant fits four waves
// per SIMD (128 VGPRs, a few spills outside the light loop); the others spill
// too much there and keep three.
template <uint32_t F>
__global__ __launch_bounds__(256, F == 0u ? 4 : 3) void k_shade(ShadeArgs a, PixelList L) {
    __shared__ uint32_t pre[LIST_SUBS + 1];
    list_prefix(L, pre);
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const uint32_t width = to_u32(p.screen.x), pixels = width * to_u32(p.screen.y);
    const uint32_t n = pre[LIST_SUBS];
    if (!(F & SHADE_CULL)) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            const uint32_t pix = list_entry(L, pre, i);
            if (pix >= pixels) continue;
            const PixelShade px = pixel_setup(p, a.center[pix], a.beta[pix], pix % width, pix / width);
            const F3 color = plain_pixel_color<F>(a, p, px, true);
            a.shaded[pix] = make_float4(color.x, color.y, color.z, 1.0f);
            a.shaded_depth[pix] = px.view_dist;
        }
        return;
    }
    // Whole waves step together (the light culling works across the wave):
    // the grid stride and the block are multiples of 64.
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t first = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); first < n; first += gridDim.x * blockDim.x) {
        const uint32_t i = first + lane;
        const uint32_t pix = i < n ? list_entry(L, pre, i) : pixels;
        const bool valid = pix < pixels;
        PixelShade px;
        px.world = splat3(0.0f);
        if (valid) px = pixel_setup(p, a.center[pix], a.beta[pix], pix % width, pix / width);
        const F3 color = plain_pixel_color<F>(a, p, px, valid);
        if (valid) {
            a.shaded[pix] = make_float4(color.x, color.y, color.z, 1.0f);
            a.shaded_depth[pix] = px.view_dist;
        }
    }
}

// ---- filter.wgsl ----------------------------------------------------------------

// Where filter_pixel takes its taps from: the centre word and colour of the
// in-screen pixel (qx, qy).
struct GlobalTaps {
    const uint64_t* center;
    const float4* shaded;
    uint32_t width;
    __device__ __forceinline__ size_t at(int32_t qx, int32_t qy) const {
        return (size_t)(uint32_t)qy * width + (uint32_t)qx;
    }
    __device__ __forceinline__ uint64_t word(int32_t qx, int32_t qy) const { return center[at(qx, qy)]; }
    __device__ __forceinline__ float4 color(int32_t qx, int32_t qy) const { return shaded[at(qx, qy)]; }
};

template <class Taps>
__device__ __forceinline__ void filter_pixel(const FilterArgs& a, const View& p, int32_t x, int32_t y,
                                             const Taps& taps) {
    int32_t width = to_i32(p.screen.x), height = to_i32(p.screen.y);
    size_t pix = (size_t)y * (uint32_t)width + (uint32_t)x;
    uint64_t center_word = taps.word(x, y);
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
    float4 cc = taps.color(x, y);
    const int radius = 5;
    float s_par = mx(p.filter_params.x, 0.05f), s_perp = mx(p.filter_params.y, 0.02f);
    float sigma_par = 5.0f * s_par, sigma_perp = 5.0f * s_perp;
    float spar2 = sigma_par * sigma_par, sperp2 = sigma_perp * sigma_perp;
    // Tap weights are continuous: the spatial and colour Gaussians
    // (filter.wgsl:99-115) as one exp2 of the summed exponents, the reciprocal
    // widths with log2(e) folded in.
    const float k_par = frcp(spar2) * 1.44269504088896341f, k_perp = frcp(sperp2) * 1.44269504088896341f;
    const float k_col = 1.44269504088896341f / (0.9f * 0.9f);
    float p_norm = depth_code_fraction(depth_q);
    // filter.wgsl's tap depth test |rdiv(q, 16777214) - p_norm| <= 1.45e-3 as
    // an interval of depth codes: the quotient and the difference are
    // monotone in q, and the pixel's own code passes. The bounds start from a
    // float estimate and move to the exact edge with the test itself.
    uint32_t q_lo, q_hi;
    {
        auto pass = [&](uint32_t q) { return !(fabsf(depth_code_fraction(q) - p_norm) > 1.45e-3f); };
        q_lo = to_u32((p_norm - 1.45e-3f) * 16777214.0f);
        q_lo = q_lo < depth_q ? q_lo : depth_q;
        if (pass(q_lo)) {
            while (q_lo > 0u && pass(q_lo - 1u)) --q_lo;
        } else {
            do { ++q_lo; } while (!pass(q_lo));  // ends at depth_q at the latest
        }
        q_hi = to_u32((p_norm + 1.45e-3f) * 16777214.0f);
        q_hi = q_hi > depth_q ? (q_hi < 0xFFFFFEu ? q_hi : 0xFFFFFEu) : depth_q;
        if (pass(q_hi)) {
            while (q_hi < 0xFFFFFEu && pass(q_hi + 1u)) ++q_hi;
        } else {
            do { --q_hi; } while (!pass(q_hi));
        }
    }
    bool promote = !center_hit;
    bool use_color = !promote && cc.w > 0.0f;
    F3 accum = splat3(0.0f);
    float weight = 0.0f, h_center = 0.0f, h_cons = 0.0f;
    // One row of 11 taps at a time: the row's centre words and colours are
    // fetched together before the taps are summed in the reference's order.
    constexpr int K = 2 * radius + 1;
    for (int dy = -radius; dy <= radius; ++dy) {
        const int32_t qy = y + dy;
        const bool row_in = qy >= 0 && qy < height;
        const int32_t ry = row_in ? qy : y;
        uint64_t qw[K];
        float4 qc[K];
        bool in[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int32_t qx = x + j - radius;
            in[j] = row_in && qx >= 0 && qx < width;
            const int32_t rx = in[j] ? qx : x;
            qw[j] = taps.word(rx, ry);
            qc[j] = taps.color(rx, ry);
        }
        float d_par[K], perp2[K];
        bool use[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const F2 d = {i2f(j - radius), i2f(dy)};
            d_par[j] = dot2(d, ts);
            const F2 perp = d - ts * d_par[j];
            // d_perp = sqrt(perp2); correctly rounded sqrt(x) <= 0.5 exactly when
            // x <= 0x3E800001 (0.25 + 2^-25)
            perp2[j] = dot2(perp, perp);
            const bool qc_empty = gb_empty(qw[j]);
            if (in[j] && perp2[j] <= __uint_as_float(0x3E800001u)) {
                if (!qc_empty) h_center = h_center + 1.0f;
                else if (!gb_empty(a.cons[(size_t)(uint32_t)ry * (uint32_t)width + (uint32_t)(x + j - radius)]))
                    h_cons = h_cons + 1.0f;
            }
            const uint32_t qd = gb_depth(qw[j]);
            use[j] = in[j] && !qc_empty && qd >= q_lo && qd <= q_hi;
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (!use[j]) continue;
            float e = __builtin_fmaf(d_par[j] * d_par[j], k_par, perp2[j] * k_perp);
            if (use_color) {
                const float dr = cc.x - qc[j].x, dg = cc.y - qc[j].y, db = cc.z - qc[j].z;
                e = __builtin_fmaf(__builtin_fmaf(dr, dr, __builtin_fmaf(dg, dg, db * db)), k_col, e);
            }
            const float w = __builtin_amdgcn_exp2f(-e);
            accum = {__builtin_fmaf(qc[j].x, w, accum.x), __builtin_fmaf(qc[j].y, w, accum.y),
                     __builtin_fmaf(qc[j].z, w, accum.z)};
            weight = weight + w;
        }
    }
    F3 color = {cc.x, cc.y, cc.z};
    if (weight > 1e-5f) color = accum * frcp(weight);
    float coverage = h_center / mx(h_center + h_cons, 1.0f);
    float cons_depth = cons_hit ? depth_code_fraction(gb_depth(cons_word)) : 1.0f;
    bool closer = cons_hit && center_hit && (p_norm - cons_depth) >= 5e-4f;
    float alpha = closer ? 1.0f : clampf(coverage, 0.0f, 1.0f);
    if (promote && weight <= 1e-5f) {
        alpha = 0.0f;
        color = splat3(0.0f);
    }
    a.filtered[pix] = make_float4(color.x, color.y, color.z, alpha);
    a.filtered_depth[pix] = view_dist;
}

__global__ __launch_bounds__(256) void k_filter(FilterArgs a, PixelList L) {
    __shared__ uint32_t pre[LIST_SUBS + 1];
    list_prefix(L, pre);
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const uint32_t width = to_u32(p.screen.x), pixels = width * to_u32(p.screen.y);
    const uint32_t n = pre[LIST_SUBS];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t pix = list_entry(L, pre, i);
        if (pix >= pixels) continue;
        filter_pixel(a, p, (int32_t)(pix % width), (int32_t)(pix / width), GlobalTaps{a.center, a.shaded, width});
    }
}

// The filter by 16 x 16 screen tiles, one workgroup per tile and one thread
// per pixel: the pixels with a conservative word (the filter list's pixels: a
// centre hit writes the conservative word too) are gathered, in pixel order,
// into the first waves, and their taps read from an LDS copy of the tile's
// 26 x 26 neighbourhood (centre words and colours; empty words off the
// screen). The same filter_pixel as k_filter: the same results.
constexpr int FILTER_TILE = 16, FILTER_HALO = FILTER_TILE + 10;

struct LdsTaps {
    const uint64_t* words;
    const float4* colors;
    int32_t bx, by;  // screen position of halo entry 0
    __device__ __forceinline__ int at(int32_t qx, int32_t qy) const { return (qy - by) * FILTER_HALO + (qx - bx); }
    __device__ __forceinline__ uint64_t word(int32_t qx, int32_t qy) const { return words[at(qx, qy)]; }
    __device__ __forceinline__ float4 color(int32_t qx, int32_t qy) const { return colors[at(qx, qy)]; }
};

__global__ __launch_bounds__(256) void k_filter_tiles(FilterArgs a) {
    static_assert(FILTER_TILE * FILTER_TILE == 256, "one thread per tile pixel");
    __shared__ uint64_t s_word[FILTER_HALO * FILTER_HALO];
    __shared__ float4 s_color[FILTER_HALO * FILTER_HALO];
    __shared__ uint8_t s_pix[FILTER_TILE * FILTER_TILE];
    __shared__ uint32_t s_wave[4];
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const int32_t W = to_i32(p.screen.x), H = to_i32(p.screen.y);
    const uint32_t tiles_x = ((uint32_t)W + FILTER_TILE - 1u) / FILTER_TILE;
    const uint32_t tiles = tiles_x * (((uint32_t)H + FILTER_TILE - 1u) / FILTER_TILE);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int32_t ox = (int32_t)(tile % tiles_x) * FILTER_TILE, oy = (int32_t)(tile / tiles_x) * FILTER_TILE;
        const int i = (int)threadIdx.x;
        const int32_t x = ox + (i & (FILTER_TILE - 1)), y = oy + i / FILTER_TILE;
        const bool cov = x < W && y < H && !gb_empty(a.cons[(size_t)(uint32_t)y * (uint32_t)W + (uint32_t)x]);
        const uint64_t bal = __ballot(cov);
        if (lane == 0u) s_wave[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = 0u;
        for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
        const uint32_t n = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        if (cov) s_pix[before + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = (uint8_t)i;
        if (n == 0u) {  // the same for every thread
            __syncthreads();  // s_wave is written again for the next tile
            continue;
        }
        for (int e = i; e < FILTER_HALO * FILTER_HALO; e += 256) {
            const int32_t qx = ox - 5 + e % FILTER_HALO, qy = oy - 5 + e / FILTER_HALO;
            const bool in = qx >= 0 && qy >= 0 && qx < W && qy < H;
            const size_t q = in ? (size_t)(uint32_t)qy * (uint32_t)W + (uint32_t)qx : 0u;
            s_word[e] = in ? a.center[q] : ~0ull;
            s_color[e] = in ? a.shaded[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        __syncthreads();
        if ((uint32_t)i < n) {
            const uint32_t k = s_pix[i];
            const LdsTaps taps{s_word, s_color, ox - 5, oy - 5};
            filter_pixel(a, p, ox + (int32_t)(k & (FILTER_TILE - 1)), oy + (int32_t)(k / FILTER_TILE), taps);
        }
        __syncthreads();  // the next tile reuses the LDS copies
    }
}

}  // namespace hm

// ---- launchers ----------------------------------------------------------------

namespace hm {

hipError_t launch_fill32(hipStream_t s, uint32_t* dst, size_t words, uint32_t value, int max_blocks,
                         const uint32_t* gate) {
    if (words == 0) return hipSuccess;
    size_t quads = (words + 3) / 4;
    size_t blocks = (quads + 255) / 256;
    if (blocks > (size_t)max_blocks) blocks = (size_t)max_blocks;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)blocks), dim3(256), 0, s, dst, words, value, gate);
    return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes, int blocks) {
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, bytes / 16);
    return hipGetLastError();
}

hipError_t launch_copy32(hipStream_t s, const void* src, void* dst, size_t words, int max_blocks) {
    if (words == 0) return hipSuccess;
    size_t blocks = (words + 255) / 256;
    if (blocks > (size_t)max_blocks) blocks = (size_t)max_blocks;
    hipLaunchKernelGGL(k_copy32, dim3((unsigned)blocks), dim3(256), 0, s, (const uint32_t*)src, (uint32_t*)dst, words);
    return hipGetLastError();
}

hipError_t launch_atomics(hipStream_t s, void* buf, size_t words, uint32_t threads, uint32_t per_thread, uint32_t kind) {
    hipLaunchKernelGGL(k_atomics, dim3((threads + 255) / 256), dim3(256), 0, s, buf, words, per_thread, kind, 12345u);
    return hipGetLastError();
}

hipError_t launch_lds_atomics(hipStream_t s, uint32_t* sink, uint32_t blocks, uint32_t words, uint32_t per_thread,
                              uint32_t kind) {
    const size_t lds = (size_t)words * (kind == 0 ? 8u : 4u);
    if (kind == 0) hipLaunchKernelGGL(k_lds_atomics<0>, dim3(blocks), dim3(256), lds, s, sink, words, per_thread);
    else if (kind == 1) hipLaunchKernelGGL(k_lds_atomics<1>, dim3(blocks), dim3(256), lds, s, sink, words, per_thread);
    else hipLaunchKernelGGL(k_lds_atomics<2>, dim3(blocks), dim3(256), lds, s, sink, words, per_thread);
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

static uint32_t compact_blocks(uint32_t pixels, bool gated) {
    const uint32_t want = (pixels + 255) / 256;
    return gated ? (want < 256u ? want : 256u) : want;  // gated: usually switched off, keep it small
}

uint32_t compact_list_cap(uint32_t pixels, bool gated) {
    const uint32_t blocks = compact_blocks(pixels, gated);
    const uint32_t iters = blocks ? (pixels + blocks * 256u - 1u) / (blocks * 256u) : 0u;
    return ((blocks + LIST_SUBS - 1u) / LIST_SUBS) * iters * 256u;  // sub-list of block b: b % LIST_SUBS
}

hipError_t launch_compact(hipStream_t s, int mode, const uint64_t* center, const uint64_t* cons, uint32_t pixels,
                          const PixelList& out, float4* color_out, float* depth_out, const uint32_t* gate) {
    if (pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill32, dim3(1), dim3(64), 0, s, out.counts, (size_t)LIST_SUBS, 0u, gate);
    dim3 g(compact_blocks(pixels, gate != nullptr)), b(256);
    if (mode == 0)
        hipLaunchKernelGGL(k_compact<0>, g, b, 0, s, center, cons, pixels, out, color_out, depth_out, gate);
    else
        hipLaunchKernelGGL(k_compact<1>, g, b, 0, s, center, cons, pixels, out, color_out, depth_out, gate);
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const ShadeArgs& a, const PixelList& list, uint32_t pixels, int blocks) {
    if (pixels == 0) return hipSuccess;
    uint32_t want = (pixels + 255) / 256;
    const dim3 grid(want < (uint32_t)blocks ? want : (uint32_t)blocks);
    switch (shade_variant(a)) {
        case 0u: hipLaunchKernelGGL(k_shade<0u>, grid, dim3(256), 0, s, a, list); break;
        case SHADE_CULL: hipLaunchKernelGGL(k_shade<SHADE_CULL>, grid, dim3(256), 0, s, a, list); break;
        case SHADE_EXTRAS: hipLaunchKernelGGL(k_shade<SHADE_EXTRAS>, grid, dim3(256), 0, s, a, list); break;
        default: hipLaunchKernelGGL(k_shade<SHADE_CULL | SHADE_EXTRAS>, grid, dim3(256), 0, s, a, list); break;
    }
    return hipGetLastError();
}

hipError_t launch_filter(hipStream_t s, const FilterArgs& a, const PixelList& list, uint32_t pixels, int blocks) {
    if (pixels == 0) return hipSuccess;
    uint32_t want = (pixels + 255) / 256;
    hipLaunchKernelGGL(k_filter, dim3(want < (uint32_t)blocks ? want : (uint32_t)blocks), dim3(256), 0, s, a, list);
    return hipGetLastError();
}

hipError_t launch_filter_tiles(hipStream_t s, const FilterArgs& a, uint32_t width, uint32_t height) {
    const uint32_t tiles = ((width + FILTER_TILE - 1u) / FILTER_TILE) * ((height + FILTER_TILE - 1u) / FILTER_TILE);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_filter_tiles, dim3(tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace hm
