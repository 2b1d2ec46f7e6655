// What the frame's results feed after hair_pass:
//  - hm_shadow_factors: HairShadowMaterial's deep-opacity lookup for other
//    surfaces (hair_shadow.wesl:30-75), for a batch of world points;
//  - hm_composite: the composite draw (composite.wgsl:27-46 with the
//    pipeline state of pipeline.rs:299-339 and the compare of
//    pass.rs:891-922) into caller-owned view targets, which makes hair of
//    one groom occlude the next through the view depth;
//  - hm_bind_scene_depth: the view's depth buffer as the raster's scene
//    depth with no copy (raster.wgsl:93, pass.rs:365-370).
#include <algorithm>
#include <cstring>

#include "hm_ctx.h"
#include "hm_device.h"
#include "hm_kernels.h"

namespace hm {
namespace {

struct ReceiverLight {
    float clip[16];
    F4 eye, forward, depth;  // depth: near, far, extinction per strand fragment, 1 once a map exists
};

// hair_shadow.wesl:30-50 with the map size at run time. Float operations in
// the shader's order (the library is built without contraction).
__device__ __forceinline__ float receiver_texel(const ReceiverLight& L, const uint32_t* __restrict__ dom, uint32_t S,
                                                int32_t x, int32_t y, float dist) {
    if (x < 0 || y < 0 || x >= (int32_t)S || y >= (int32_t)S) return 1.0f;
    const uint32_t pix = (uint32_t)y * S + (uint32_t)x;
    const uint32_t front_q = dom[pix];
    if (front_q == HM_DOM_EMPTY) return 1.0f;
    const float near_ = L.depth.x, far_ = L.depth.y;
    const float z_front = mixs(near_, far_, rdiv(u2f(front_q), 16777214.0f));
    const float thickness = rdiv(mx(far_ - near_, 1e-3f), 16.0f);
    const float t = rdiv(mx(dist - z_front, 0.0f), thickness);
    const uint4* c = reinterpret_cast<const uint4*>(dom + (size_t)S * S + (size_t)pix * HM_DOM_LAYERS);
    float hits = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = c[q];
        const uint32_t n[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) hits = hits + u2f(n[k]) * clampf(t - (float)(q * 4 + k), 0.0f, 1.0f);
    }
    return __expf(-hits * L.depth.z);
}

__device__ __forceinline__ float receiver_visibility(const ReceiverLight& L, const uint32_t* __restrict__ dom,
                                                     uint32_t S, F3 w) {  // hair_shadow.wesl:52-75
    if (L.depth.w == 0.0f) return 1.0f;
    const float cx = mat_row(L.clip, 0, w.x, w.y, w.z, 1.0f);
    const float cy = mat_row(L.clip, 1, w.x, w.y, w.z, 1.0f);
    const float cw = mat_row(L.clip, 3, w.x, w.y, w.z, 1.0f);
    if (cw <= 0.0f) return 1.0f;
    const float ndc_x = rdiv(cx, cw), ndc_y = rdiv(cy, cw);
    const float ux = ndc_x * 0.5f + 0.5f;
    const float uy = 1.0f - (ndc_y * 0.5f + 0.5f);
    if (ux < 0.0f || uy < 0.0f || ux > 1.0f || uy > 1.0f) return 1.0f;
    const float dist = dot3(w - xyz(L.eye), xyz(L.forward));
    const float fx = ux * u2f(S) - 0.5f, fy = uy * u2f(S) - 0.5f;
    const float flx = floorf(fx), fly = floorf(fy);
    const int32_t x0 = to_i32(flx), y0 = to_i32(fly);
    const float wx = fx - flx, wy = fy - fly;
    const float v00 = receiver_texel(L, dom, S, x0, y0, dist);
    const float v10 = receiver_texel(L, dom, S, x0 + 1, y0, dist);
    const float v01 = receiver_texel(L, dom, S, x0, y0 + 1, dist);
    const float v11 = receiver_texel(L, dom, S, x0 + 1, y0 + 1, dist);
    return mixs(mixs(v00, v10, wx), mixs(v01, v11, wx), wy);
}

__global__ __launch_bounds__(256) void k_shadow_factors(ReceiverLight L, const uint32_t* __restrict__ dom, uint32_t S,
                                                        const float4* __restrict__ points, uint32_t count,
                                                        float* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float4 q = points[i];
        out[i] = receiver_visibility(L, dom, S, F3{q.x, q.y, q.z});
    }
}

// One thread per pixel of the view targets (composite.wgsl:27-46 and its
// blend and depth state). Each pixel is read and written by its own thread
// only, so the blend needs no ordering beyond the stream's.
__global__ __launch_bounds__(256) void k_composite(View p, const float4* __restrict__ color,
                                                   const float* __restrict__ depth_buf, float4* __restrict__ tc,
                                                   float* __restrict__ td, uint32_t width, uint32_t pixels) {
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < pixels; pix += gridDim.x * blockDim.x) {
        const float4 c = color[pix];
        if (c.w < 0.001f) continue;  // discard
        const int32_t x = (int32_t)(pix % width), y = (int32_t)(pix / width);
        const F3 w = reconstruct_world(p, x, y, depth_buf[pix]);
        const float cz = mat_row(p.clip, 2, w.x, w.y, w.z, 1.0f);
        const float cw = mx(mat_row(p.clip, 3, w.x, w.y, w.z, 1.0f), 1e-5f);
        const float d = clampf(rdiv(cz, cw), 0.0f, 1.0f);
        const float cur = td[pix];
        const bool pass = p.pass_mode[2] == 1u ? d >= cur : d <= cur;
        if (!pass) continue;
        td[pix] = d;
        const float4 o = tc[pix];
        const float a = c.w, ia = 1.0f - a;
        tc[pix] = make_float4(c.x * a + o.x * ia, c.y * a + o.y * ia, c.z * a + o.z * ia, a + o.w * ia);
    }
}

}  // namespace
}  // namespace hm

extern "C" {

hm_status hm_shadow_factors(hm_ctx* ctx, const hm_shadow_light* light, const uint32_t* dom, uint32_t dom_size,
                            const float* points, uint32_t count, float* factors) {
    if (!ctx) return HM_ERR_INVALID;
    if (!light || (count && (!points || !factors || !dom)))
        return hm_fail(ctx, HM_ERR_INVALID, "hm_shadow_factors: null buffer");
    // hm_targets' map sizes: even, so the counts after the S^2 fronts start on
    // a 16-byte boundary (read as uint4)
    if (dom_size == 0 || (dom_size & 1u) || dom_size > 16384u || (reinterpret_cast<uintptr_t>(dom) & 15u) ||
        (reinterpret_cast<uintptr_t>(points) & 15u))
        return hm_fail(ctx, HM_ERR_INVALID, "hm_shadow_factors: map size must be even in [2, 16384], buffers 16-byte aligned");
    if (count == 0) return HM_OK;
    hm::ReceiverLight L;
    static_assert(sizeof(hm::ReceiverLight) == sizeof(hm_shadow_light), "ShadowLight layout");
    std::memcpy(&L, light, sizeof(L));
    const uint32_t blocks = std::min<uint32_t>((count + 255u) / 256u, (uint32_t)ctx->cu_count * 16u);
    hipLaunchKernelGGL(hm::k_shadow_factors, dim3(blocks), dim3(256), 0, ctx->stream, L, dom, dom_size,
                       reinterpret_cast<const float4*>(points), count, factors);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hm_hip_fail(ctx, e, "hm_shadow_factors");
    return HM_OK;
}

hm_status hm_composite(hm_ctx* ctx, const hm_params* camera, float* target_color, float* target_depth) {
    if (!ctx) return HM_ERR_INVALID;
    if (!camera || !target_color || !target_depth) return hm_fail(ctx, HM_ERR_INVALID, "hm_composite: null argument");
    if (reinterpret_cast<uintptr_t>(target_color) & 15u)
        return hm_fail(ctx, HM_ERR_INVALID, "hm_composite: target_color must be 16-byte aligned");
    if (!ctx->gb_slab) return hm_fail(ctx, HM_ERR_NOT_READY, "hm_composite: targets not allocated (hm_targets)");
    if ((uint32_t)camera->screen[0] != ctx->width || (uint32_t)camera->screen[1] != ctx->height ||
        camera->pass_mode[0] != 0u)
        return hm_fail(ctx, HM_ERR_INVALID, "hm_composite: camera params do not match the targets");
    hm::View v;
    static_assert(sizeof(hm::View) == sizeof(hm_params), "View layout");
    std::memcpy(&v, camera, sizeof(v));
    const bool filtered = camera->flags[1] != 0u;
    const float4* color = filtered ? ctx->filtered : ctx->shaded;
    const float* depth = filtered ? ctx->filtered_depth : ctx->shaded_depth;
    const uint32_t pixels = ctx->width * ctx->height;
    const uint32_t blocks = std::min<uint32_t>((pixels + 255u) / 256u, (uint32_t)ctx->cu_count * 16u);
    hipLaunchKernelGGL(hm::k_composite, dim3(blocks), dim3(256), 0, ctx->stream, v, color, depth,
                       reinterpret_cast<float4*>(target_color), target_depth, ctx->width, pixels);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hm_hip_fail(ctx, e, "hm_composite");
    return HM_OK;
}

}  // extern "C"
