// Batched multi-view frames: a fixed list of views dealt in turn to one or
// more lanes and replayed from one captured HIP graph. A lane is a chain of
// whole frames on streams and work buffers of its own (the first lane is the
// context itself, the others are lane contexts that read its lights, probe
// and network weights); the lanes fork from the context's stream at the start
// of the graph and join it at the end, so one view's low-occupancy phases
// (LOD, scans, long-segment and finish launches, the tail of a tile kernel)
// overlap another view's wide kernels. The views' buffers are grown by one
// eager pass before the capture, so the captured sequence holds kernel
// launches only (no allocation, no synchronization, no memset nodes: every
// clear is a fill kernel).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "hm_ctx.h"
#include "hm_kernels.h"

struct hm_batch {
    std::vector<hm_batch_view> views;
    std::vector<void*> out;  // per view: rgba float4[px], then depth float[px]
    std::vector<hm_ctx*> lanes;  // lanes[0] is the context; the others are owned here
    // The groom each view renders: its own on lane 0, on the other lanes a
    // shadow with per-pass buffers of the lane's own (owned here).
    std::vector<hm_groom*> grooms;
    std::vector<hm_groom*> shadows;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr;        // the graph's start, on the context's stream
    std::vector<hipEvent_t> join;     // per lane after the first: its last copy
    // the context's buffers the captured launches read: their allocation
    // generation, and the light counts baked into the launches
    uint64_t alloc_gen = 0, shade_gen = 0;
    uint32_t point_count = 0, dir_count = 0;
};

namespace {

#define BATCH_HIP(ctx, call)                                            \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return hm_hip_fail((ctx), e_, #call);     \
    } while (0)

void release(hm_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (size_t k = 1; k < b->lanes.size(); ++k) (void)hipStreamSynchronize(b->lanes[k]->stream);
    for (hm_groom* g : b->shadows) hm_groom_free(nullptr, g);
    for (size_t k = 1; k < b->lanes.size(); ++k) hm_destroy(b->lanes[k]);
    for (void* p : b->out)
        if (p) (void)hipFree(p);
    delete b;
}

// One view on lane `c` with groom `g` (the view's own or the lane's shadow of
// it): targets, frame, copy of the final colour and depth.
hm_status record_view(hm_ctx* c, hm_groom* g, const hm_batch_view& v, void* out) {
    hm_status st = hm_targets(c, v.width, v.height, v.dom_size);
    if (st != HM_OK) return st;
    if ((st = hm_frame(c, g, v.params, &v.opts)) != HM_OK) return st;
    const size_t px = (size_t)v.width * v.height;
    const bool filtered = v.params[0].flags[1] != 0u;
    const void* rgba = filtered ? (const void*)c->filtered : (const void*)c->shaded;
    const void* depth = filtered ? (const void*)c->filtered_depth : (const void*)c->shaded_depth;
    char* dst = static_cast<char*>(out);
    BATCH_HIP(c, hm::launch_copy32(c->stream, rgba, dst, px * 4, c->cu_count * 8));
    BATCH_HIP(c, hm::launch_copy32(c->stream, depth, dst + px * 16, px, c->cu_count * 8));
    return HM_OK;
}

// Every view once, view i on lane i % lanes, between the fork from and the
// join back into the context's stream (the join is enqueued on every exit,
// so that a failed capture still ends with every lane joined).
hm_status record(hm_ctx* ctx, hm_batch* b) {
    const size_t nl = b->lanes.size();
    if (nl > 1) {
        BATCH_HIP(ctx, hipEventRecord(b->fork, ctx->stream));
        for (size_t k = 1; k < nl; ++k) BATCH_HIP(ctx, hipStreamWaitEvent(b->lanes[k]->stream, b->fork, 0));
    }
    hm_status st = HM_OK;
    for (size_t i = 0; i < b->views.size() && st == HM_OK; ++i) {
        hm_ctx* c = b->lanes[i % nl];
        st = record_view(c, b->grooms[i], b->views[i], b->out[i]);
        if (st != HM_OK && c != ctx) ctx->err = "batch lane " + std::to_string(i % nl) + ": " + c->err;
    }
    for (size_t k = 1; k < nl; ++k) {
        hipError_t e = hipEventRecord(b->join[k - 1], b->lanes[k]->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, b->join[k - 1], 0);
        if (e != hipSuccess && st == HM_OK) st = hm_hip_fail(ctx, e, "batch lane join");
    }
    return st;
}

}  // namespace

extern "C" {

hm_status hm_batch_create(hm_ctx* ctx, const hm_batch_view* views, uint32_t count, hm_batch** out) {
    if (!ctx || !views || !out || count == 0) return hm_fail(ctx, HM_ERR_INVALID, "null argument or empty batch");
    *out = nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        const hm_batch_view& v = views[i];
        if (!v.groom || v.width == 0 || v.height == 0)
            return hm_fail(ctx, HM_ERR_INVALID, "batch view " + std::to_string(i) + ": no groom or zero size");
        if ((uint32_t)v.params[0].screen[0] != v.width || (uint32_t)v.params[0].screen[1] != v.height ||
            (uint32_t)v.params[0].dom_info[0] != v.dom_size)
            return hm_fail(ctx, HM_ERR_INVALID, "batch view " + std::to_string(i) + ": params differ from its size");
    }
    (void)hipSetDevice(ctx->device);
    hm_batch* b = new hm_batch();
    b->views.assign(views, views + count);
    b->out.assign(count, nullptr);
    b->lanes.push_back(ctx);
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
        e = hipMalloc(&b->out[i], (size_t)views[i].width * views[i].height * 20);
    for (hipEvent_t& ev : b->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        release(b);
        return hm_hip_fail(ctx, e, "hm_batch_create");
    }
    // More lanes than views would idle; a scene depth belongs to one target
    // size in one context and keeps the views in one chain.
    const bool depth = ctx->scene_depth || ctx->scene_depth_ext;
    const uint32_t lanes = depth ? 1u : std::min(ctx->tuning.batch_lanes, count);
    for (uint32_t k = 1; k < lanes; ++k) {
        hm_ctx* lane = nullptr;
        hm_status lst = hm_lane_create(ctx, &lane);
        if (lst != HM_OK) {
            release(b);
            return lst;
        }
        b->lanes.push_back(lane);
        hipEvent_t j = nullptr;
        e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        b->join.push_back(j);
        if (e == hipSuccess && !b->fork) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
        if (e != hipSuccess) {
            release(b);
            return hm_hip_fail(ctx, e, "hm_batch_create (lane events)");
        }
    }
    // Lanes other than the first render through shadows of the grooms: a
    // groom's LOD records and strand refs are rewritten by every frame of it,
    // and two lanes may render the same groom at the same time.
    b->grooms.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const size_t lane = i % b->lanes.size();
        b->grooms[i] = views[i].groom;
        if (lane == 0) continue;
        hm_groom* shadow = nullptr;
        for (uint32_t j = 0; j < i && !shadow; ++j)
            if (j % b->lanes.size() == lane && views[j].groom == views[i].groom) shadow = b->grooms[j];
        if (!shadow) {
            hm_status gst = hm_groom_shadow(ctx, views[i].groom, &shadow);
            if (gst != HM_OK) {
                release(b);
                return gst;
            }
            b->shadows.push_back(shadow);
        }
        b->grooms[i] = shadow;
    }
    // Per-kernel timing records events on the stream: keep it out of the graph.
    const uint32_t slots = ctx->slots;
    ctx->slots = 0;
    hm_status st = record(ctx, b);  // eager pass: grows every buffer the views need
    for (hm_ctx* c : b->lanes) {
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess && st == HM_OK) st = hm_hip_fail(ctx, e, "hm_batch_create (eager pass)");
    }
    if (st == HM_OK) {
        e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            st = hm_hip_fail(ctx, e, "hipStreamBeginCapture");
        } else {
            st = record(ctx, b);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(ctx->stream, &g);
            b->graph = g;
            if (st == HM_OK && e != hipSuccess) st = hm_hip_fail(ctx, e, "hipStreamEndCapture");
        }
    }
    if (st == HM_OK) {
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) st = hm_hip_fail(ctx, e, "hipGraphInstantiate");
    }
    ctx->slots = slots;
    if (st != HM_OK) {
        release(b);
        return st;
    }
    b->alloc_gen = ctx->alloc_gen;
    b->shade_gen = ctx->shade_gen;
    b->point_count = ctx->point_count;
    b->dir_count = ctx->dir_count;
    *out = b;
    return HM_OK;
}

hm_status hm_batch_run(hm_ctx* ctx, hm_batch* b, uint32_t reps) {
    if (!ctx || !b) return hm_fail(ctx, HM_ERR_INVALID, "null argument");
    if (b->alloc_gen != ctx->alloc_gen)
        return hm_fail(ctx, HM_ERR_NOT_READY, "context buffers were reallocated since the batch was created; create it again");
    if (b->point_count != ctx->point_count || b->dir_count != ctx->dir_count)
        return hm_fail(ctx, HM_ERR_NOT_READY, "the light counts changed since the batch was created; create it again");
    if (b->shade_gen != ctx->shade_gen)
        return hm_fail(ctx, HM_ERR_NOT_READY,
                       "the rect lights or the environment probe changed since the batch was created; create it again");
    for (uint32_t r = 0; r < reps; ++r) BATCH_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    return HM_OK;
}

hm_status hm_batch_time(hm_ctx* ctx, hm_batch* b, uint32_t reps, float* total_ms) {
    if (!ctx || !b || !total_ms) return hm_fail(ctx, HM_ERR_INVALID, "null argument");
    BATCH_HIP(ctx, hipEventRecord(b->ev[0], ctx->stream));
    hm_status st = hm_batch_run(ctx, b, reps);
    if (st != HM_OK) return st;
    BATCH_HIP(ctx, hipEventRecord(b->ev[1], ctx->stream));
    BATCH_HIP(ctx, hipEventSynchronize(b->ev[1]));
    BATCH_HIP(ctx, hipEventElapsedTime(total_ms, b->ev[0], b->ev[1]));
    return HM_OK;
}

hm_status hm_batch_read(hm_ctx* ctx, hm_batch* b, uint32_t view, float* rgba, float* depth) {
    if (!ctx || !b || view >= b->views.size()) return hm_fail(ctx, HM_ERR_INVALID, "bad batch view");
    const size_t px = (size_t)b->views[view].width * b->views[view].height;
    const char* src = static_cast<const char*>(b->out[view]);
    if (rgba) BATCH_HIP(ctx, hipMemcpyAsync(rgba, src, px * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) BATCH_HIP(ctx, hipMemcpyAsync(depth, src + px * 16, px * 4, hipMemcpyDeviceToHost, ctx->stream));
    BATCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

void hm_batch_free(hm_ctx* ctx, hm_batch* b) {
    if (!b) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);  // the lanes join it in every replay
    release(b);
}

}  // extern "C"
