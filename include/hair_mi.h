/*
 * hair_mi.h — C ABI of the MI355X (gfx950) deferred strand rasterizer.
 *
 * The library replaces the GPU side of bevy_hair's `hair_pass`
 * (src/render/pass.rs:259-657): the WGSL compute dispatches clear, clear_dom,
 * lod, scan, raster (x3), the fullscreen shade pass and the reconnection
 * filter. A Rust host links this header through `extern "C"` and keeps its own
 * Bevy plugin surface (HairPlugin, HairGroom, HairMesh, HairShadowMaterial).
 *
 * Conventions
 *  - All structs are plain C, little-endian, with the byte layout of the
 *    reference's `#[repr(C)]` Pod types; sizes are checked with static asserts.
 *  - Matrices are column-major float[16]: m[col*4 + row] (glam `to_cols_array`).
 *  - Every entry point returns hm_status; none aborts. The reference has no
 *    error codes and silently skips (pass.rs:290-322); a Rust caller maps any
 *    non-OK status onto that same skip.
 *  - One hm_ctx owns one HIP stream and all device allocations made through
 *    it. A context is single-threaded; use one context per host thread.
 */
#ifndef HAIR_MI_H
#define HAIR_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_LAYER_COUNT 16u  /* mesh.rs:10, common.wgsl:1 */
#define HM_STYLE_U 8u       /* mesh.rs:11 */
#define HM_STYLE_V 8u       /* mesh.rs:12 */
#define HM_STYLE_W 16u      /* mesh.rs:13 */
#define HM_STYLE_TEXELS (HM_STYLE_U * HM_STYLE_V * HM_STYLE_W)
#define HM_DOM_LAYERS 16u   /* pass.rs:24, common.wgsl:7 */
#define HM_DOM_SIGMA 0.05f  /* pass.rs:26 */
#define HM_DOM_EMPTY 0xffffffffu
#define HM_GB_EMPTY 0xffffffffffffffffull

typedef enum hm_status {
    HM_OK = 0,
    HM_ERR_INVALID = 1,     /* bad argument, size mismatch, null handle */
    HM_ERR_OOM = 2,         /* device allocation failed */
    HM_ERR_HIP = 3,         /* a HIP runtime call failed */
    HM_ERR_NO_DEVICE = 4,   /* no gfx950 device at that index */
    HM_ERR_NOT_READY = 5,   /* targets or groom missing (the reference's silent skip) */
    HM_ERR_CAPACITY = 6     /* a fixed-size limit was exceeded */
} hm_status;

/* HairParams, 432 bytes (pass.rs:42-62, common.wgsl:10-28). */
typedef struct hm_params {
    float clip_from_world[16];
    float world_from_clip[16];
    float world_from_model[16];
    float light_clip_from_world[16];
    float camera_pos[4];
    float camera_forward[4];
    float screen[4];        /* width, height, near, far */
    float appearance[4];    /* roughness, tilt, lambda, center_diameter (px) */
    float filter_params[4]; /* s_par, s_perp, dom_scale, dom_sigma */
    float albedo[4];        /* linear rgb, 1 */
    uint32_t flags[4];      /* lod, filter, ambient_occlusion, deep_opacity */
    uint32_t pass_mode[4];  /* mode (0 camera, 1 DOM depth, 2 DOM opacity), bundle count, reverse-Z, is_dom */
    float light_eye[4];
    float light_forward[4];
    float dom_info[4];      /* dom width, dom height, light near, light far */
} hm_params;

/* GpuBundle, 16 bytes (pass.rs:71-78). */
typedef struct hm_bundle {
    uint32_t layer_offset;  /* index of the bundle's first hm_layer */
    uint32_t style_offset;  /* index of the bundle's first style texel */
    uint32_t strand_count;
    uint32_t _pad;
} hm_bundle;

/* GpuLayer, 32 bytes (pass.rs:64-69): one cage corner. 16 layers x 4 corners
 * per bundle, layer-major, corner order c00, c10, c01, c11 (mesh.rs:34-36). */
typedef struct hm_layer {
    float pos_ao[4];
    float tangent[4];
} hm_layer;

/* LodRecord, 16 bytes (common.wgsl:42-47). */
typedef struct hm_lod_record {
    uint32_t n_lod;
    uint32_t control_points;
    uint32_t strand_count;
    uint32_t _pad;
} hm_lod_record;

/* StrandRef, 16 bytes (common.wgsl:49-54). */
typedef struct hm_strand_ref {
    uint32_t bundle;
    uint32_t strand;
    uint32_t n_lod;
    uint32_t control_points;
} hm_strand_ref;

/* ShadowLight, 112 bytes (pass.rs:31-40): the DOM's light view for receivers. */
typedef struct hm_shadow_light {
    float light_clip_from_world[16];
    float light_eye[4];
    float light_forward[4];
    float depth[4];         /* near, far, extinction per strand fragment, 1 once a map exists */
} hm_shadow_light;

/* A clusterable light as shade.wesl:452-480 reads it from Bevy's
 * `clustered_lights` (64 bytes). kind 0 = point, 1 = spot. */
#define HM_LIGHT_POINT 0u
#define HM_LIGHT_SPOT 1u
#define HM_LIGHT_FLAG_SHADOWS 1u            /* POINT/DIRECTIONAL_LIGHT_FLAGS_SHADOWS_ENABLED_BIT */
#define HM_LIGHT_FLAG_SPOT_Y_NEGATIVE 2u    /* POINT_LIGHT_FLAGS_SPOT_LIGHT_Y_NEGATIVE */
typedef struct hm_point_light {
    float position_radius[4];
    float color_inverse_square_range[4];   /* rgb radiance scale, w = 1 / range^2 */
    float light_custom_data[4];            /* spot: dir.x, dir.z, cone scale, cone offset */
    uint32_t flags;
    uint32_t kind;
    uint32_t _pad[2];
} hm_point_light;

/* A directional light as shade.wesl:484-501 reads it (48 bytes). */
typedef struct hm_directional_light {
    float direction_to_light[4];
    float color[4];
    uint32_t flags;         /* HM_LIGHT_FLAG_SHADOWS marks a shadow caster */
    uint32_t _pad[3];
} hm_directional_light;

/* A rectangular area light as shade.wesl:503-518 reads Bevy's `rect_lights`
 * (80 bytes): centre, the rectangle's right and up directions (its normal is
 * normalize(up x right)), linear colour, width, height and range. */
typedef struct hm_rect_light {
    float position[4];
    float right[4];
    float up[4];
    float color[4];
    float width, height, range, _pad;
} hm_rect_light;

/* The view's environment light probe (shade.wesl:341-427, Bevy's
 * light_probes view bindings): RGBA float32 cube maps, faces +x, -x, +y, -y,
 * +z, -z, each face row-major from its top-left texel (Vulkan's cube map
 * layout). The specular map holds `specular_mips` levels one after another,
 * level m with (specular_size >> m)^2 texels per face; it is sampled at level
 * roughness * (specular_mips - 1) (smallest_specular_mip_level_for_view). */
typedef struct hm_env_probe {
    const float* specular;
    uint32_t specular_size;
    uint32_t specular_mips;
    const float* diffuse;
    uint32_t diffuse_size;
    float view_rotation[4];     /* quaternion x, y, z, w */
    float intensity;            /* intensity_for_view */
} hm_env_probe;

/* HairGroom (lib.rs:32-64) after extract_hair (mod.rs:126-139). */
typedef struct hm_groom_config {
    float world_from_model[16];
    float lambda;
    float albedo_linear[3];
    float roughness;
    float tilt;
    float center_diameter;
    uint32_t lod;
    uint32_t filter;
    uint32_t ambient_occlusion;
    uint32_t deep_opacity;
} hm_groom_config;

/* The camera view (ExtractedView) hair_pass reads (pass.rs:342-349). */
typedef struct hm_view {
    float clip_from_world[16];
    float world_from_view[16];
} hm_view;

/* Cage + style of one HairMesh asset (mesh.rs:31-41). */
typedef struct hm_groom_desc {
    const hm_bundle* bundles;
    uint32_t bundle_count;
    const hm_layer* layers;
    uint32_t layer_count;       /* 64 per bundle */
    const float* style;         /* 4 floats per texel */
    uint32_t style_count;       /* 1024 per bundle */
    float bounds_min[3];
    float bounds_max[3];
} hm_groom_desc;

/* Harness-level overrides of the LOD record (camera pass only). They exist
 * because the measured workloads ask for fixed strand and control-point
 * counts that the reference's screen-space LOD would not pick; a zero value
 * keeps the reference behaviour of lod.wgsl:41-58. Culled bundles stay culled. */
typedef struct hm_lod_override {
    uint32_t control_points;   /* 0 = Eqs. 4-5; else forced count (2..127) */
    float strand_fraction;     /* 0 = Eq. 1; else n_lod = clamp(ceil(f * N), 1, N) */
} hm_lod_override;

typedef struct hm_ctx hm_ctx;
typedef struct hm_groom hm_groom;

/* Which per-pass buffers of a groom an entry point uses. */
#define HM_PASS_CAMERA 0
#define HM_PASS_DOM 1

/* Buffers that hm_read / hm_device_ptr expose. */
typedef enum hm_buffer {
    HM_BUF_CENTER = 0,          /* u64[W*H] */
    HM_BUF_CONSERVATIVE = 1,    /* u64[W*H] */
    HM_BUF_BETA = 2,            /* u32[W*H] */
    HM_BUF_SHADED = 3,          /* float4[W*H] */
    HM_BUF_SHADED_DEPTH = 4,    /* float[W*H] */
    HM_BUF_FILTERED = 5,        /* float4[W*H] */
    HM_BUF_FILTERED_DEPTH = 6,  /* float[W*H] */
    HM_BUF_DOM = 7,             /* u32[S*S*17]: S*S fronts, then 16 counts per texel */
    HM_BUF_SCENE_DEPTH = 8,     /* float[W*H] */
    HM_BUF_CAMERA_LOD = 16,     /* hm_lod_record[B] of the groom */
    HM_BUF_CAMERA_REFS = 17,    /* hm_strand_ref[N] */
    HM_BUF_CAMERA_INDIRECT = 18,/* u32[4]: x, 1, 1, 0 */
    HM_BUF_DOM_LOD = 19,
    HM_BUF_DOM_REFS = 20,
    HM_BUF_DOM_INDIRECT = 21,
    HM_BUF_CAM_BIN_STATUS = 24, /* u32[144] of the last hm_gbuffer: [0] records, [1] binned, [2] fallback
                                   (direct kernels after a record overflow), [3] long segments present,
                                   [4] long segment count; [1] and [2] both 0 when the direct kernels
                                   ran by choice; [16..79] / [80..143] the shade / filter sub-list
                                   lengths of a frame's camera tiles */
    HM_BUF_DOM_BIN_STATUS = 25  /* u32[144] of the last hm_dom, same header */
} hm_buffer;

/* Kernel ids for hm_timing_read. */
typedef enum hm_kernel {
    HM_K_CLEAR = 0,
    HM_K_CAM_LOD = 1,
    HM_K_CAM_SCAN = 2,
    HM_K_CAM_RASTER = 3,
    HM_K_CLEAR_DOM = 4,
    HM_K_DOM_LOD = 5,
    HM_K_DOM_SCAN = 6,
    HM_K_DOM_DEPTH = 7,
    HM_K_DOM_OPACITY = 8,
    HM_K_SHADE = 9,
    HM_K_FILTER = 10,
    HM_K_DOM_BIN = 11,          /* binned DOM: control points, tile counts, scan, records */
    HM_K_DOM_TILES = 12,        /* binned DOM: the tile kernel (both passes per tile) */
    HM_K_CAM_BIN = 13,          /* binned camera pass: control points, tile counts, scan, records */
    HM_K_CAM_TILES = 14,        /* binned camera pass: the tile kernel (per tile G-buffer) */
    HM_K_CAM_FINISH = 15,       /* binned camera pass: pixel lists of split tiles, long segments */
    HM_K_DOM_LONG_DEPTH = 16,   /* binned DOM: long segments, depth pass (before the tiles) */
    HM_K_DOM_LONG_OPACITY = 17, /* binned DOM: long segments, opacity pass (after the tiles) */
    HM_K_COUNT = 18
} hm_kernel;

#define HM_SHADING_PLAIN 0u
#define HM_SHADING_MLP 1u

/* Per-frame options of hm_frame (pass.rs:375-643). */
typedef struct hm_frame_opts {
    float exposure;                 /* view.exposure (shade.wesl:588) */
    hm_lod_override camera_lod;     /* see hm_lod_override */
    uint32_t shading;               /* HM_SHADING_PLAIN or HM_SHADING_MLP */
    float mlp_tolerance;            /* per-tile probe bound for HM_SHADING_MLP */
} hm_frame_opts;

/* ---- lifetime ----------------------------------------------------------- */

/* Opens `device` (a gfx950 GPU) and creates the context's stream. */
hm_status hm_create(int device, hm_ctx** out);
void hm_destroy(hm_ctx* ctx);
const char* hm_status_string(hm_status status);
/* Detail of the last failure on this context (static storage, never NULL). */
const char* hm_last_error(const hm_ctx* ctx);

/* Replaces GpuGroom::upload (pass.rs:98-167): copies the cage, the style
 * volume and allocates the per-pass LOD/ref/indirect buffers. */
hm_status hm_groom_upload(hm_ctx* ctx, const hm_groom_desc* desc, hm_groom** out);
void hm_groom_free(hm_ctx* ctx, hm_groom* groom);
uint32_t hm_groom_bundle_count(const hm_groom* groom);
uint32_t hm_groom_strand_count(const hm_groom* groom);

/* Replaces ensure_targets (pass.rs:789-815) with a runtime DOM size (the
 * reference hard-codes 512, pass.rs:22). Allocates only when a size exceeds
 * what was allocated before: the allocations only grow, so target sizes can
 * alternate without allocating or synchronizing. A change of the target size
 * drops the scene depth (hm_set_scene_depth). */
hm_status hm_targets(hm_ctx* ctx, uint32_t width, uint32_t height, uint32_t dom_size);

/* The scene depth texture raster tests against (raster.wgsl:93). NULL means
 * no scene geometry: every texel reads `clear_value` (0 for reverse-Z). */
hm_status hm_set_scene_depth(hm_ctx* ctx, const float* host_depth, float clear_value);

/* Binds a caller-owned device buffer float[W*H] as that scene depth, with no
 * copy and no synchronization: the view's depth texture as the reference
 * binds it (pass.rs:365-370), e.g. imported with hipImportExternalMemory, and
 * the buffer hm_composite writes, so that one groom's composited hair
 * occludes the next groom's raster. Holds until hm_set_scene_depth, another
 * binding or a change of the target size; NULL is hm_set_scene_depth(NULL).
 * The buffer must stay valid while frames that read it run; a batch created
 * with a binding in force replays that binding. */
hm_status hm_bind_scene_depth(hm_ctx* ctx, const float* device_depth, float clear_value);

/* The light list shade reads in place of Bevy's view bindings. Point and spot
 * lights are culled per wave of 64 pixels where none of them is in the
 * light's range, as Bevy's clusters list only the lights that reach a froxel
 * (the range window is zero beyond the range, so the colours are unchanged). */
hm_status hm_set_lights(hm_ctx* ctx, const hm_point_light* points, uint32_t point_count,
                        const hm_directional_light* directionals, uint32_t directional_count);
/* Rectangular area lights (the AREA_LIGHT_LUTS loop, shade.wesl:503-518);
 * count 0 removes them. Copied to the device. */
hm_status hm_set_rect_lights(hm_ctx* ctx, const hm_rect_light* rects, uint32_t count);
/* The environment probe shade adds (probe_radiance * ao, shade.wesl:586);
 * NULL removes it (the !ENVIRONMENT_MAP shader, which adds nothing). The
 * cube maps are copied to the device. */
hm_status hm_set_env_probe(hm_ctx* ctx, const hm_env_probe* probe);

/* ---- host-side parameter construction ----------------------------------- */

/* The three HairParams of one groom: camera, DOM depth, DOM opacity
 * (pass.rs:399-453 with make_params pass.rs:676-709, light_fit pass.rs:747-776,
 * world_extent pass.rs:737-745, depth_is_reverse pass.rs:778-787).
 * `direction_to_light` is key_direction()'s result (pass.rs:711-719). */
hm_status hm_make_params(const hm_view* view, const hm_groom_config* groom,
                         const float bounds_min[3], const float bounds_max[3],
                         uint32_t bundle_count, const float direction_to_light[3],
                         uint32_t width, uint32_t height, uint32_t dom_size,
                         hm_params out[3]);

/* key_direction (pass.rs:711-719): first shadow caster, else first light,
 * else (0.35, 0.82, 0.45); normalized. */
void hm_key_direction(const hm_directional_light* lights, uint32_t count, float out[3]);

/* The ShadowLight published with the DOM (pass.rs:574-579). */
void hm_shadow_light_from_params(const hm_params* dom_params, hm_shadow_light* out);

/* ---- kernel-level entry points (one per WGSL dispatch) ------------------ */

hm_status hm_clear(hm_ctx* ctx);                                        /* clear.wgsl */
hm_status hm_clear_dom(hm_ctx* ctx);                                    /* clear_dom.wgsl */
hm_status hm_lod(hm_ctx* ctx, hm_groom* groom, const hm_params* params,
                 int pass, const hm_lod_override* override_or_null);    /* lod.wgsl */
hm_status hm_scan(hm_ctx* ctx, hm_groom* groom, const hm_params* params, int pass); /* scan.wgsl */
hm_status hm_raster(hm_ctx* ctx, hm_groom* groom, const hm_params* params, int pass); /* raster.wgsl */
hm_status hm_shade(hm_ctx* ctx, const hm_params* params, float exposure); /* shade.wesl:fs */
hm_status hm_filter(hm_ctx* ctx, const hm_params* params);              /* filter.wgsl */

/* The G-buffer of one groom in one call: clear (clear.wgsl) and the camera
 * raster (raster.wgsl, pass_mode.x = 0) over the camera work list that
 * hm_lod/hm_scan built (pass.rs:379-520 after the lod and scan dispatches).
 * The result equals hm_clear + hm_raster bit for bit. With hm_tuning.cam_tile
 * set and at most 16384 tiles it runs binned: one workgroup per screen tile
 * rasterizes into LDS and writes the tile once; when the tile records do not
 * fit their buffer the direct kernels run instead, decided on the device
 * (HM_BUF_CAM_BIN_STATUS says which ran). Asynchronous. */
hm_status hm_gbuffer(hm_ctx* ctx, hm_groom* groom, const hm_params* camera_params);

/* The deep opacity map of one groom in one call: clear_dom, the depth pass
 * (`depth_params`, pass_mode.x = 1) and the opacity pass (`opacity_params`,
 * pass_mode.x = 2) over the DOM work list that hm_lod/hm_scan built
 * (pass.rs:524-583 after the lod and scan dispatches). The result equals
 * hm_clear_dom + hm_raster + hm_raster bit for bit. With hm_tuning.dom_tile
 * set, at most 16384 tiles and the two params differing only in pass_mode
 * it runs binned: one
 * workgroup per light-space tile evaluates both passes in LDS; when the tile
 * records do not fit their buffer the direct kernels run instead, decided on
 * the device (HM_BUF_DOM_BIN_STATUS says which ran). Asynchronous. */
hm_status hm_dom(hm_ctx* ctx, hm_groom* groom, const hm_params* depth_params,
                 const hm_params* opacity_params);

/* Counts the raster's algorithmic work for one pass without writing any
 * target: counts[0] = camera write_pixel calls that reach the atomics,
 * counts[1] = of those, centre hits, counts[2] = DOM writes that pass the
 * depth range, counts[3] = segments drawn. Synchronous. */
hm_status hm_raster_count(hm_ctx* ctx, hm_groom* groom, const hm_params* params,
                          int pass, uint64_t counts[4]);

/* ---- offline baker ------------------------------------------------------ */

/* bake_hair_mesh (bake.rs:89-213): explicit strands -> hair-mesh cages with
 * style volumes and ambient occlusion, the arrays hm_groom_upload takes.
 * `points` float[point_count][3]; `strands` (first point, point count) pairs
 * per strand, as HairStrands holds them (hair_file.rs:9-15). Roots are
 * clustered in a 3D grid (the reference's scalp-mesh clustering is not part
 * of this entry point). Host computation, no device and no context. */
typedef struct hm_bake_result hm_bake_result;
hm_status hm_bake(const float* points, uint32_t point_count, const uint32_t* strands, uint32_t strand_count,
                  hm_bake_result** out);
/* The baked mesh as an hm_groom_desc whose pointers stay valid until
 * hm_bake_free (bundle_count 0 for the empty mesh). */
hm_status hm_bake_get(const hm_bake_result* result, hm_groom_desc* desc);
void hm_bake_free(hm_bake_result* result);

/* ---- frame driver ------------------------------------------------------- */

/* One groom of hair_pass (pass.rs:379-619): camera lod/scan, hm_gbuffer,
 * [DOM lod/scan, hm_dom if flags.w], shade, [filter if flags.y].
 * `params` are hm_make_params's three records. Asynchronous on the stream. */
hm_status hm_frame(hm_ctx* ctx, hm_groom* groom, const hm_params params[3],
                   const hm_frame_opts* opts);

/* ---- after the frame: composite and shadow receivers ---------------------- */

/* The composite draw of one groom (composite.wgsl:27-46 with the blend and
 * depth state of pipeline.rs:299-339): the frame's final colour (filtered
 * when camera->flags[1], else shaded) and depth into the caller's view
 * targets, target_color float4[W*H] and target_depth float[W*H] (device
 * pointers). A pixel with alpha >= 0.001 takes the clip depth of its
 * reconstructed point, clamped to [0, 1]; it passes when >= the stored depth
 * with reverse Z (camera->pass_mode[2] = 1, pass.rs:916-921), else <=; a
 * passing pixel writes its depth and blends its colour (rgb src * a + dst *
 * (1 - a), alpha src + dst * (1 - a)). Asynchronous on the context's stream. */
hm_status hm_composite(hm_ctx* ctx, const hm_params* camera, float* target_color, float* target_depth);

/* HairShadowMaterial's lookup (hair_shadow.wesl:30-75): the key light's
 * attenuation by the hair for `count` world points of other surfaces,
 * factors[i] = visibility at points[i] (float4 per point, w unused). `dom` is
 * a map of dom_size^2 texels in hm_dom's layout (HM_BUF_DOM), `light` its
 * ShadowLight (hm_shadow_light_from_params of the DOM params; depth.w = 0
 * gives 1 everywhere). Device pointers; asynchronous on the context's stream. */
hm_status hm_shadow_factors(hm_ctx* ctx, const hm_shadow_light* light, const uint32_t* dom, uint32_t dom_size,
                            const float* points, uint32_t count, float* factors);

/* ---- data access and timing --------------------------------------------- */

/* Copies a buffer to host memory (synchronizes the stream). Groom buffers
 * (HM_BUF_*_LOD/REFS/INDIRECT) need `groom`; target buffers ignore it. */
hm_status hm_read(hm_ctx* ctx, hm_groom* groom, hm_buffer which, void* dst, size_t bytes);
/* Device address and size of a buffer, for zero-copy sharing (the DOM for
 * HairShadow receivers, pass.rs:571-573). */
hm_status hm_device_ptr(hm_ctx* ctx, hm_groom* groom, hm_buffer which, void** ptr, size_t* bytes);
hm_status hm_sync(hm_ctx* ctx);

/* With frames > 0, the next `frames` hm_frame calls record a HIP event pair
 * around each kernel on the context's stream (0 turns recording off).
 * hm_timing_read returns each kernel's mean milliseconds over the recorded
 * frames (HM_K_COUNT floats, 0 for kernels that did not run). */
hm_status hm_timing_enable(hm_ctx* ctx, int frames);
hm_status hm_timing_read(hm_ctx* ctx, float* ms, uint32_t count);

/* Times `frames` back-to-back hm_frame calls with one event pair around all
 * of them; returns total milliseconds. Per-kernel recording, if enabled,
 * continues inside. */
hm_status hm_time_frames(hm_ctx* ctx, hm_groom* groom, const hm_params params[3],
                         const hm_frame_opts* opts, uint32_t frames, float* total_ms);

/* ---- tuning ------------------------------------------------------------- */

#define HM_RASTER_PRECHECK 1u   /* screen each atomic min with a plain load */
#define HM_RASTER_XCD 2u        /* one contiguous strand range per XCD (shared L2) */

/* Launch choices of the raster passes; results do not depend on them. */
typedef struct hm_tuning {
    uint32_t raster_flags[3];      /* per pass mode (camera, DOM depth, DOM opacity) */
    uint32_t raster_blocks_per_cu[3]; /* persistent raster grid per pass, 256-thread workgroups per CU */
    uint32_t dom_tile;             /* hm_dom: 0 = direct kernels, 16 or 32 = binned, texels per tile side */
    uint32_t cam_tile;             /* hm_gbuffer: 0 = direct kernels, 16 or 32 = binned, pixels per tile side */
    uint32_t bin_box_limit;        /* tiles a segment's box may span before it is drawn with
                                      global atomics instead (1..64) */
    uint32_t bin_records;          /* tile record buffer length; 0 = the worst case (126 segments per
                                      strand x bin_box_limit tiles), which cannot overflow */
    uint32_t bin_blocks_per_cu;    /* binning grid: 256-thread workgroups per CU, each counting
                                      one contiguous run of segments in LDS (1..64) */
    uint32_t cam_chunk;            /* hm_gbuffer: records per tile workgroup; larger tiles are split
                                      over several workgroups (0 = one per tile, else 64..1048576) */
    uint32_t cam_bin_blocks_per_cu; /* hm_gbuffer's binning grid (0 = bin_blocks_per_cu, else 1..64):
                                      its per-workgroup tile histogram is small, so it takes more
                                      workgroups per CU than the DOM pass's */
    uint32_t light_cull;           /* shading: 0 = cull point and spot lights per wave when some
                                      light's range may end inside the groom (from 64 lights when
                                      shading is called on its own), 1 = never, 2 = always; the
                                      colours are the same either way */
    uint32_t dom_busy_blocks_per_cu; /* hm_dom with 16-texel tiles: 0 = one 256-thread workgroup per
                                      tile; 1 or 2 = that many persistent 1024-thread workgroups per
                                      CU, taking the tiles that hold records one at a time (each
                                      sample walked once, kept in LDS for the opacity pass) and
                                      then writing the empty tiles */
    uint32_t frame_serial;         /* hm_frame: 0 = the DOM pass on a second stream beside the camera
                                      pass (its generation after the camera generation, shading after
                                      both); 1 = the passes one after the other on one stream. The
                                      results are the same either way */
    uint32_t batch_lanes;          /* hm_batch_create (1..4): views are dealt in turn to this many
                                      lanes, each a chain on streams and work buffers of its own,
                                      joined at the end of the captured graph; the first lane is the
                                      context itself. A batch created while a scene depth is set
                                      uses one lane. The results are the same either way */
    uint32_t filter_tiles;         /* hm_frame's filter on the binned path: 0 = one thread per pixel of
                                      the camera tiles' lists, taps from memory; 1 = one workgroup per
                                      16 x 16 screen tile, taps from an LDS copy of its neighbourhood.
                                      The results are the same either way */
} hm_tuning;

hm_status hm_get_tuning(hm_ctx* ctx, hm_tuning* out);
hm_status hm_set_tuning(hm_ctx* ctx, const hm_tuning* tuning);
/* Times the raster variants on this groom and view, `reps` launches each
 * from cleared targets, keeps the fastest per pass and returns it: direct
 * kernels (flags x grid size) per pass; camera tiles 16 / 32 x 64 / 128 / 256 /
 * 512 / 1024 / 2048 records per tile workgroup (cam_chunk); DOM tiles 16 / 32 x
 * binning workgroups per CU 1 / 2 / 4, 16-texel tiles also with persistent tile
 * workgroups (dom_busy_blocks_per_cu 0 / 1 / 2); then whole frames with the DOM
 * pass beside or after the camera pass (frame_serial), and with the filter
 * over the pixel lists or by screen tiles (filter_tiles). Leaves the targets
 * dirty: run a frame afterwards. */
hm_status hm_autotune(hm_ctx* ctx, hm_groom* groom, const hm_params params[3], const hm_frame_opts* opts,
                      uint32_t reps, hm_tuning* best);

/* ---- roofline microbenchmarks -------------------------------------------- */

/* STREAM-style copy of `bytes` between two fresh buffers, `reps` times;
 * returns read + write GB/s. */
hm_status hm_bench_copy(hm_ctx* ctx, size_t bytes, uint32_t reps, float* gbps);
/* `count` random-address atomics over a `target_bytes` buffer: kind 0 u64
 * atomicMin, 1 u32 atomicMin, 2 u32 atomicAdd; returns atomics per second. */
hm_status hm_bench_atomics(hm_ctx* ctx, size_t target_bytes, uint32_t kind, uint64_t count, float* per_second);
/* About `count` random-address LDS atomics, 256-thread workgroups (8 per CU)
 * each over its own array of the largest power-of-two word count within
 * `lds_bytes` (8 KiB to 64 KiB): kind 0 u64 min, 1 u32 min, 2 u32 add;
 * returns atomics per second over the whole chip (the tile kernels' LDS
 * atomic peak). */
hm_status hm_bench_lds_atomics(hm_ctx* ctx, uint32_t lds_bytes, uint32_t kind, uint64_t count, float* per_second);

/* ---- allocation-free scan (no counterpart in the reference) --------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. hm_scan
 * (scan.wgsl) over caller-owned device buffers: `lod` holds bundle_count
 * hm_lod_records (as hm_lod writes them), `refs` capacity hm_strand_refs and
 * `indirect` 4 words; the results are those hm_scan gives for the same
 * records, bit for bit. Its scratch (the per-bundle prefix) is a static
 * device array of HM_SCAN_NOALLOC_MAX_BUNDLES + 1 words (256 KiB) that each
 * GPU holds until the process ends; more bundles return HM_ERR_CAPACITY.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from any context or stream) race. Allocates nothing; asynchronous on the
 * context's stream. */
#define HM_SCAN_NOALLOC_MAX_BUNDLES 65536u
hm_status hm_scan_noalloc(hm_ctx* ctx, const hm_lod_record* lod, uint32_t bundle_count,
                          hm_strand_ref* refs, uint32_t capacity, uint32_t* indirect);

/* ---- learned shading (no counterpart in the reference) -------------------
 * A 32 -> 64 -> 64 -> 64 -> 16 MLP (bf16 weights on the matrix cores, f32
 * accumulation) stands in for dual_scatter (shade.wesl:153-190): its outputs
 * are the dual-scattering terms from which every light's back lobe and the
 * key light's global multiple scattering and DOM forward-scattering spread
 * (shade.wesl:220-242) are evaluated; single scattering stays analytic.
 * Features per pixel (float[32]): tangent[3], view[3], key light
 * direction[3], the 16 DOM layer transmittances (bilinear over the pixel's
 * four DOM taps), sigma_a * 0.25 [3], roughness, the key light's theta_h / pi,
 * cos(theta_o), the key light's optical depth * 0.125. Outputs (float[16]):
 * log(v + 1e-6) of the forward R lobe sum, forward TT[3], forward TRT[3],
 * a_back[3], then delta_b[3] (linear), then log(sigma_b^2 + 1e-6)[3].
 * Weights: float[11264] (W0 64x32, W1 64x64, W2 64x64, W3 16x64, out x in
 * row-major), biases float[208]. */

/* He-uniform initialisation from `seed`; zero biases. */
hm_status hm_mlp_init(hm_ctx* ctx, uint32_t seed);
hm_status hm_mlp_set_weights(hm_ctx* ctx, const float* weights, const float* biases);
hm_status hm_mlp_get_weights(hm_ctx* ctx, float* weights, float* biases);
/* y[n][16] = network(x[n][32]); host arrays, synchronous. */
hm_status hm_mlp_forward(hm_ctx* ctx, const float* x, uint32_t n, float* y);
/* Loss sum_o ((y_o - t_o) * scale_o)^2 / (16 n) over the 16 outputs and its
 * gradients (float[11264], float[208]) for one batch; synchronous. */
hm_status hm_mlp_gradients(hm_ctx* ctx, const float* x, const float* t, uint32_t n, const float scale[16],
                           float* loss, float* grad_weights, float* grad_biases);
/* Sample buffer for training (features + plain-method targets); resets the count. */
hm_status hm_mlp_reserve_samples(hm_ctx* ctx, uint32_t capacity);
/* Appends samples from the current G-buffer/DOM (after the raster passes of
 * a frame): each hit pixel is kept with probability `keep`. */
hm_status hm_mlp_collect(hm_ctx* ctx, const hm_params* camera_params, float keep, uint32_t seed, uint32_t* total);
hm_status hm_mlp_read_samples(hm_ctx* ctx, float* x, float* t, uint32_t n);
/* Replaces the held samples with host ones (x[n][32], t[n][16]); n <= capacity. */
hm_status hm_mlp_set_samples(hm_ctx* ctx, const float* x, const float* t, uint32_t n);
/* Feature range the network was trained on (NULL clears it); tiles with a
 * pixel outside it are shaded by the plain method. */
hm_status hm_mlp_set_range(hm_ctx* ctx, const float lo[32], const float hi[32]);
/* Adam over the collected samples for `epochs` passes; returns the last
 * epoch's mean loss. */
hm_status hm_mlp_train(hm_ctx* ctx, uint32_t epochs, uint32_t batch, float learning_rate, const float scale[16],
                       uint32_t seed, float* final_loss);
/* Network shading of the frame's G-buffer into HM_BUF_SHADED. The covered
 * pixels are shaded in tiles of 64 consecutive covered pixels (one wave
 * each). A pixel whose features leave the trained range is shaded by the plain
 * method. With `tolerance` > 0, every 4th tile free of such pixels also shades
 * one probe pixel by the plain method and is re-shaded by it when the network
 * colour misses by more than `tolerance` (relative, L1 over RGB); the tiles
 * follow the list order, which varies between runs, so only tolerance 0 gives
 * run-to-run identical images. `fallback_tiles` (may be NULL, then the call
 * is asynchronous) receives the count of tiles with plain-shaded pixels. */
hm_status hm_shade_mlp(hm_ctx* ctx, const hm_params* params, float exposure, float tolerance,
                       uint32_t* fallback_tiles);

/* ---- batched multi-view frames (no counterpart in the reference) ----------
 * A batch renders a fixed list of views - groom, its three HairParams, target
 * and DOM size, frame options - back to back on the context's stream, each
 * view's final colour (filtered when the groom's filter flag is set, else
 * shaded) and depth copied into the batch's own output slot. hm_batch_create
 * runs the views once eagerly (which grows every buffer to what the views
 * need), then captures the whole sequence into one HIP graph; hm_batch_run
 * replays it with one launch. The graph holds the addresses of the context's
 * buffers and the light counts: hm_batch_run fails when a buffer was
 * reallocated since (a larger target, more lights, a scene depth...), the
 * light counts changed or hm_set_env_probe was called - create the batch
 * again. Light values, cube map texels of an unchanged probe layout and
 * network weights are read at each replay. The grooms must outlive the batch. Calls
 * that overlap a running batch on the same context race. */
typedef struct hm_batch hm_batch;
typedef struct hm_batch_view {
    hm_groom* groom;
    hm_params params[3];            /* camera, DOM depth, DOM opacity (hm_make_params) */
    uint32_t width, height, dom_size;
    hm_frame_opts opts;
} hm_batch_view;
hm_status hm_batch_create(hm_ctx* ctx, const hm_batch_view* views, uint32_t count, hm_batch** out);
/* Enqueues `reps` replays of the captured frame (asynchronous). */
hm_status hm_batch_run(hm_ctx* ctx, hm_batch* batch, uint32_t reps);
/* Wall time of `reps` replays in milliseconds (synchronous). */
hm_status hm_batch_time(hm_ctx* ctx, hm_batch* batch, uint32_t reps, float* total_ms);
/* One view's output: rgba float[width*height*4], depth float[width*height]
 * (either may be NULL); synchronizes the stream. */
hm_status hm_batch_read(hm_ctx* ctx, hm_batch* batch, uint32_t view, float* rgba, float* depth);
void hm_batch_free(hm_ctx* ctx, hm_batch* batch);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* HAIR_MI_H */
