// Private state behind the opaque hm_ctx / hm_groom handles.
#pragma once

#include <string>
#include <vector>

#include "hm_kernels.h"

struct hm_mlp_state;  // learned shading (hm_mlp_host.cpp)

// The last per-wave culling choice, for the light list and groom box it
// was made for.
struct hm_cull_cache {
    bool valid = false, cull = false;
    uint64_t light_gen = 0;
    float box[6] = {0, 0, 0, 0, 0, 0};
};

struct hm_groom {
    uint32_t nb = 0, nl = 0, ns = 0, strands = 0, capacity = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    hm_bundle* bundles = nullptr;
    hm_layer* layers = nullptr;
    float* style = nullptr;
    hm_lod_record* lod[2] = {nullptr, nullptr};
    hm_strand_ref* refs[2] = {nullptr, nullptr};
    uint32_t* prefix[2] = {nullptr, nullptr};
    uint32_t* indirect[2] = {nullptr, nullptr};
    // Per pass: the most control points the last hm_lod can have given a
    // strand (lod.wgsl:38-55 with the override); 127 before any hm_lod.
    uint32_t ncp_bound[2] = {127u, 127u};
    // A batch lane's shadow (hm_groom_shadow): bundles, layers and style are
    // the groom's own, never freed here.
    bool borrowed = false;
};

// Work buffers of the binned DOM (sized for the largest groom and map seen).
struct hm_bin_state {
    float4* cps = nullptr;          // [slot * 128 + idx]: world xyz, light-space x
    float* cpy = nullptr;           // light-space (or screen) y
    float* dcp = nullptr;           // DOM: 5 floats per [slot * 128 + idx]: world xyz, light-space x, y
    float4* cpt = nullptr;          // camera: world tangent xyz, ao
    float4* sinfo = nullptr;        // camera: per strand uv, beta, control points
    size_t slots = 0;               // strand slots cps / cpy / dcp / cpt / sinfo hold
    // Per pass (camera, DOM), each grown only by its own pass, so that the
    // two passes of a frame can run at the same time: status words (the
    // camera's shade / filter list lengths are read after the DOM pass),
    // per-tile count / offset / cursor / chunk offset, segment boxes,
    // long-segment lists, tile records and the counting workgroups' lists.
    uint32_t* meta[2] = {nullptr, nullptr};
    size_t meta_words[2] = {0, 0};
    // A pass's launch sequence failed part way: its status words (the
    // long-segment cursor among them) are cleared before the pass runs again.
    bool meta_dirty[2] = {false, false};
    uint32_t* seg_box[2] = {nullptr, nullptr};    // per segment id
    uint32_t* long_list[2] = {nullptr, nullptr};
    size_t seg_slots[2] = {0, 0};                 // strand slots seg_box / long_list hold
    uint32_t* records[2] = {nullptr, nullptr};
    size_t record_capacity[2] = {0, 0};
    uint2* lists[2] = {nullptr, nullptr};         // (tile, rank) entries of the counting workgroups
    size_t list_capacity[2] = {0, 0};
    uint2* wg_list[2] = {nullptr, nullptr};       // one per binning workgroup
    size_t wg_capacity[2] = {0, 0};
    uint32_t* chunk_tile = nullptr; // camera: tile of each tile-workgroup chunk
    size_t chunk_capacity = 0;
};

struct hm_ctx {
    int device = 0;
    int cu_count = 256;
    // The stream every call enqueues on. hm_frame points it at `side` for the
    // DOM pass, which then runs beside the camera pass and joins `main`
    // before shading; both streams have the default priority.
    hipStream_t stream = nullptr;
    hipStream_t main = nullptr, side = nullptr;
    // hm_frame's fork and join: frame start on main, camera control points
    // written (the DOM generation reads them), DOM pass done on side.
    hipEvent_t fork_ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t gen_event = nullptr;  // recorded after the camera generation when set
    uint32_t width = 0, height = 0, dom_size = 0;
    // Target allocations only grow: a smaller size reuses them (the layouts
    // below follow the current size), so views of several sizes can follow
    // one another without allocating.
    size_t cap_gb_words = 0, cap_px = 0, cap_dom_words = 0;
    void* gb_slab = nullptr;  // center | conservative | beta, cleared by one fill
    size_t gb_words = 0;
    uint64_t* center = nullptr;
    uint64_t* cons = nullptr;
    uint32_t* beta = nullptr;
    float4* shaded = nullptr;
    float* shaded_depth = nullptr;
    float4* filtered = nullptr;
    float* filtered_depth = nullptr;
    uint32_t* dom = nullptr;
    uint32_t* pix_list = nullptr;   // covered-pixel list (shade, then filter)
    uint32_t* pix_list2 = nullptr;  // filter list when the camera tiles write both lists
    size_t list_len = 0;            // entries in pix_list / pix_list2 (capacity)
    uint32_t* pix_count = nullptr;  // sub-list counters: shade, then filter
    float* scene_depth = nullptr;          // library-owned copy (hm_set_scene_depth)
    const float* scene_depth_ext = nullptr; // caller-owned device buffer (hm_bind_scene_depth), used first
    float scene_clear = 0.0f;
    hm_point_light* points = nullptr;
    uint32_t point_count = 0, point_cap = 0;
    uint32_t spot_first = 0;  // the point lights come first in `points`, then the spot lights
    hm_directional_light* dirs = nullptr;
    uint32_t dir_count = 0, dir_cap = 0;
    std::vector<hm_point_light> host_points;  // the point/spot list as last set
    uint64_t light_gen = 0;                     // bumped by every hm_set_lights
    hm_cull_cache cull_cache;
    // World box of the groom being shaded (hm_frame), for the choice to cull
    // lights per wave; unset outside hm_frame.
    bool groom_box_set = false;
    float groom_box[6] = {0, 0, 0, 0, 0, 0};
    hm_rect_light* rects = nullptr;
    uint32_t rect_count = 0, rect_cap = 0;
    float* env_maps = nullptr;  // the probe's specular levels, then its diffuse map (RGBA float)
    size_t env_bytes = 0;
    hm::EnvArgs env = {};       // env.on == 0: no environment probe
    // Bumped when a value the shading launches take by value changes (the
    // rect light count, the probe's sizes, rotation or intensity).
    uint64_t shade_gen = 0;
    uint64_t* counters = nullptr;
    hm_tuning tuning = {{HM_RASTER_PRECHECK, HM_RASTER_PRECHECK, HM_RASTER_PRECHECK}, {16u, 16u, 16u}, 16u, 16u, 8u, 0u, 2u, 2048u, 4u, 0u, 2u, 0u, 2u, 0u};
    hm_bin_state bin;
    // Per-kernel event pairs for `slots` frames: [slot][kernel][begin/end].
    uint32_t slots = 0, frame = 0;
    std::vector<hipEvent_t> ev;
    std::vector<uint8_t> ran;
    hipEvent_t frame_ev[2] = {nullptr, nullptr};
    hm_mlp_state* mlp = nullptr;
    // Bumped whenever a context-owned device buffer is (re)allocated or freed:
    // a captured batch graph holds their addresses.
    uint64_t alloc_gen = 0;
    // A batch lane (hm_lane_create): the light, rect light and probe buffers
    // and the network weights are the parent context's, never freed here.
    bool borrowed = false;
    std::string err;
};


// Where the covered-pixel lists of a frame are: written by the camera tiles
// (sub-list counters in the camera status words, `cap` entries per sub-list;
// `gate` nonzero on the device when the direct kernels ran instead, null when
// that cannot happen), or compacted by the shading call itself.
struct hm_lists {
    bool from_tiles = false;
    uint32_t cap = 0;
    const uint32_t* gate = nullptr;
};

// Shared helpers of the host translation units.
hm_status hm_fail(hm_ctx* ctx, hm_status st, const std::string& msg);
hm_status hm_hip_fail(hm_ctx* ctx, hipError_t e, const char* what);
void hm_mlp_release(hm_ctx* ctx);
// A batch lane of `parent`: streams, targets, binning buffers, counters and
// learned-shading flags of its own; the parent's lights, rect lights, probe
// and network weights read in place (the parent must outlive the lane and
// keep those buffers; its allocation generation tells when it did not).
// No scene depth, no timing slots; freed with hm_destroy.
hm_status hm_lane_create(hm_ctx* parent, hm_ctx** out);
// A batch lane's shadow of groom g: g's bundles, layers and style (read
// only), and per-pass LOD records, strand refs, prefix and indirect words of
// its own, so that lanes rendering g at the same time do not share them.
// Freed with hm_groom_free; g must outlive it.
hm_status hm_groom_shadow(hm_ctx* ctx, const hm_groom* g, hm_groom** out);
// The lane's learned-shading state: the parent's weights, its own flags.
hm_status hm_mlp_borrow(hm_ctx* lane, const hm_ctx* parent);
// Looks up the allocation-free scan's scratch address on `device` (current).
hipError_t hm_noalloc_setup(int device);
// The light lists and the probe of the context, into shading arguments.
void hm_shade_lights(hm_ctx* ctx, hm::ShadeArgs& a);
// The shade list for `lists`, compacting it first where needed (enqueued on the stream).
hm_status hm_shade_list(hm_ctx* ctx, const hm_lists& lists, hm::PixelList* out);
hm_status hm_shade_mlp_impl(hm_ctx* ctx, const hm_params* p, float exposure, float tolerance, uint32_t* fallback,
                            const hm_lists& lists);
