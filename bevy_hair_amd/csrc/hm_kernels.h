// Launch interface between the host driver (hm_host.cpp) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hair_mi.h"

namespace hm {

constexpr uint32_t RASTER_PRECHECK = 1u;  // screen each atomic min with a plain load
constexpr uint32_t RASTER_XCD = 2u;       // contiguous strand ranges per XCD

// Everything one raster pass reads or writes (the raster.wgsl bind group,
// pipeline.rs:110-170, plus the work list bound and a counter block).
struct RasterArgs {
    hm_params view;
    const hm_bundle* bundles;
    const hm_layer* layers;
    const float* style;             // float4 per texel
    const hm_strand_ref* refs;
    const uint32_t* indirect;       // [0] = strands to draw
    uint32_t capacity;              // refs length
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    uint32_t* dom;
    const float* scene_depth;       // W*H or null
    float scene_clear;
    uint64_t* counters;             // counting variant only
    uint32_t opt;                   // RASTER_* flags
    const uint32_t* gate;           // when set and zero, the launch does nothing
};

// Covered-pixel list split into LIST_SUBS sub-lists with a counter each, so
// that writers spread their atomics: entry j of sub-list s is
// list[s * cap + j], j < counts[s].
constexpr uint32_t LIST_SUBS = 64u;
struct PixelList {
    uint32_t* list;
    uint32_t* counts;  // LIST_SUBS words
    uint32_t cap;      // entries per sub-list
};

// The environment probe as the shading kernels read it: RGBA float cube
// maps in device memory (faces +x, -x, +y, -y, +z, -z; specular levels one
// after another, level m starting at texel spec_base[m]).
constexpr uint32_t ENV_MAX_MIPS = 16u;
struct EnvArgs {
    const float4* specular;
    const float4* diffuse;
    uint32_t spec_size, spec_mips, diff_size, on;
    float rotation[4];
    float intensity;
    uint32_t spec_base[ENV_MAX_MIPS];
};

struct ShadeArgs {
    hm_params view;
    const uint64_t* center;
    const uint32_t* beta;
    const uint32_t* dom;
    const hm_point_light* points;   // point lights, then spot lights (each in the caller's order)
    uint32_t point_count;           // both kinds
    uint32_t spot_first;            // index of the first spot light
    const hm_directional_light* directionals;
    uint32_t directional_count;
    const hm_rect_light* rects;
    uint32_t rect_count;
    float exposure;
    float4* shaded;
    float* shaded_depth;
    uint32_t cull;  // cull point and spot lights per wave
    EnvArgs env;
};

// Variants of the shading kernels: SHADE_CULL culls point and spot lights
// per wave; SHADE_EXTRAS evaluates rect lights and the environment probe
// (compiled out when the view has neither).
constexpr uint32_t SHADE_CULL = 1u, SHADE_EXTRAS = 2u;
__host__ __device__ inline uint32_t shade_variant(const ShadeArgs& a) {
    return (a.cull ? SHADE_CULL : 0u) | ((a.rect_count || a.env.on) ? SHADE_EXTRAS : 0u);
}

struct FilterArgs {
    hm_params view;
    const uint64_t* center;
    const uint64_t* cons;
    const float4* shaded;
    const float* shaded_depth;
    float4* filtered;
    float* filtered_depth;
};

// Binned raster (hm_binned.hip). Control points of strand slot s sit at
// s * BIN_STRIDE + index; segment ids are (s << seg_shift) + index of the
// segment's first control point, with 2^seg_shift (16 .. 128) at least the
// pass's most segments per strand, so the binning passes walk no more empty
// segment slots than they must.
constexpr uint32_t BIN_STRIDE = 128u;
constexpr uint32_t BIN_NONE = 0xffffffffu;  // no box (behind the view, off the map, or long)
// Status words at the head of BinArgs::meta.
constexpr uint32_t BIN_RECORDS = 0u;     // records the tiles asked for
constexpr uint32_t BIN_OK = 1u;          // binned kernels run
constexpr uint32_t BIN_FALLBACK = 2u;    // direct kernels run
constexpr uint32_t BIN_LONG = 3u;        // binned and long segments exist
constexpr uint32_t BIN_LONG_COUNT = 4u;
constexpr uint32_t BIN_LIST_COUNT = 5u;  // (tile, rank) entries the counting workgroups wrote
constexpr uint32_t BIN_SPLIT = 8u;       // binned and some tile has more than one chunk
constexpr uint32_t BIN_LONG_APPEND = 9u; // k_bin_gen's long-segment cursor; k_bin_scan moves it to
                                         // BIN_LONG_COUNT and zeroes it, so it is zero whenever a
                                         // k_bin_gen starts (k_bin_gen clears every other status
                                         // word and the tile counts itself)
constexpr uint32_t BIN_BUSY = 10u;       // DOM: tiles holding records (k_bin_scan lists them after the
                                         // chunk offsets)
constexpr uint32_t BIN_BUSY_NEXT = 11u;  // DOM: next busy-tile list entry a tile workgroup takes
constexpr uint32_t BIN_EMPTY_NEXT = 12u; // DOM: next 16-tile run whose empty tiles a workgroup writes
constexpr uint32_t BIN_SHADE_SUBS = 16u;                        // camera: shade sub-list counters
constexpr uint32_t BIN_FILTER_SUBS = BIN_SHADE_SUBS + LIST_SUBS; // camera: filter sub-list counters
constexpr uint32_t BIN_META_WORDS = BIN_FILTER_SUBS + LIST_SUBS; // then count[tiles], offset[tiles + 1], cursor[tiles],
                                         // chunk offset[tiles + 1], busy tiles[tiles]
constexpr uint32_t BIN_MAX_TILES = 16384u;  // one LDS word per tile in the binning workgroups

struct BinArgs {
    hm_params view;                 // DOM view (the two DOM passes share it)
    const hm_bundle* bundles;
    const hm_layer* layers;
    const float* style;
    const hm_strand_ref* refs;
    const uint32_t* indirect;
    uint32_t capacity;              // refs length
    float4* cps;                    // per control point [slot * BIN_STRIDE + idx]: world xyz, light-space x
    float* cpy;                     // light-space y
    float* dcp;                     // DOM pass: 5 floats per control point (world xyz, light-space x, y)
    uint32_t* seg_box;              // per segment id: tile box or BIN_NONE
    uint32_t* meta;
    uint32_t* records;              // segment ids grouped by tile
    uint32_t record_capacity;
    uint32_t can_overflow;          // record_capacity is below the worst case: the fallback may run
    uint2* lists;                   // per counting workgroup: (tile, rank within the tile)
    uint32_t list_capacity;
    uint2* wg_list;                 // per counting workgroup: (first list entry, entries)
    uint32_t chunk;                 // records per tile workgroup (0: one workgroup per tile)
    uint32_t* chunk_tile;           // tile of each chunk
    uint32_t chunk_capacity;
    uint32_t* gb_slab;              // camera: centre | conservative | beta words, cleared when tiles split
    uint64_t gb_words;
    uint32_t fill_clear;            // k_bin_fill also clears: 1 the G-buffer if tiles split, 2 the DOM fronts if long
    const hm_strand_ref* reuse_refs;  // DOM: camera work list whose control points cps still holds, or null
    const uint32_t* reuse_indirect;
    uint32_t* long_list;            // segments whose box spans more than box_limit tiles
    uint32_t long_capacity;
    uint32_t tiles_x, tiles, tile_shift, box_limit;
    uint32_t seg_shift;             // log2 of the segment slots per strand (4 .. 7)
    uint32_t dom_size;
    uint32_t* dom;                  // fronts S*S, then counts S*S*16
    // camera pass
    float4* cpt;                    // per control point: world tangent xyz, ao
    float4* sinfo;                  // per strand slot: uv.x, uv.y, beta, control points (bits)
    uint32_t width, height;
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    const float* scene_depth;
    float scene_clear;
    // With emit_lists the camera tiles also write the shade and filter pixel
    // lists (sub-list of a tile: tile % LIST_SUBS; counters in meta) and the
    // empty results of uncovered pixels, the work of k_compact<0> and <1>.
    uint32_t emit_lists;
    uint32_t* shade_list;
    uint32_t* filter_list;
    uint32_t list_cap;              // entries per sub-list
    float4* shaded;
    float* shaded_depth;
    float4* filtered;
    float* filtered_depth;
};

// Strand slot, index and control point of segment `id`.
__device__ __forceinline__ uint32_t seg_strand(const BinArgs& a, uint32_t id) { return id >> a.seg_shift; }
__device__ __forceinline__ uint32_t seg_index(const BinArgs& a, uint32_t id) { return id & ((1u << a.seg_shift) - 1u); }
__device__ __forceinline__ uint32_t seg_cp(const BinArgs& a, uint32_t id) {
    return seg_strand(a, id) * BIN_STRIDE + seg_index(a, id);
}

// Generation, count, scan and fill of one pass; `after_gen`, when given, is
// recorded once the generation kernel is enqueued (the DOM pass of a frame
// waits on it before reading the camera's control points).
hipError_t launch_bin(hipStream_t s, const BinArgs& a, int mode, int gen_blocks, int bin_blocks,
                      hipEvent_t after_gen = nullptr);
hipError_t launch_dom_long(hipStream_t s, const BinArgs& a, int mode, int long_blocks);  // mode 1 before, 2 after the tiles
// 16 x 16 tiles with busy_blocks > 0: the persistent busy-tile kernel on that
// many workgroups; otherwise one workgroup per tile.
hipError_t launch_dom_tiles(hipStream_t s, const BinArgs& a, int busy_blocks);
hipError_t launch_cam_tiles(hipStream_t s, const BinArgs& a, int tile_blocks);
hipError_t launch_cam_finish(hipStream_t s, const BinArgs& a, int long_blocks);  // split tiles' lists, long segments

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes, int blocks);
// Copies `words` 32-bit words (a graph-safe device copy).
hipError_t launch_copy32(hipStream_t s, const void* src, void* dst, size_t words, int max_blocks);
hipError_t launch_atomics(hipStream_t s, void* buf, size_t words, uint32_t threads, uint32_t per_thread, uint32_t kind);
// `blocks` workgroups of random LDS atomics over `words` (a power of two) targets each.
hipError_t launch_lds_atomics(hipStream_t s, uint32_t* sink, uint32_t blocks, uint32_t words, uint32_t per_thread,
                              uint32_t kind);
hipError_t launch_fill32(hipStream_t s, uint32_t* dst, size_t words, uint32_t value, int max_blocks,
                         const uint32_t* gate = nullptr);
hipError_t launch_lod(hipStream_t s, const hm_params& prm, const hm_bundle* bundles, const hm_layer* layers,
                      hm_lod_record* out, uint32_t nb, uint32_t ovr_cps, float ovr_frac);
hipError_t launch_scan(hipStream_t s, const hm_lod_record* lod, uint32_t nb, uint32_t capacity, uint32_t* prefix,
                       uint32_t* indirect, hm_strand_ref* refs);
hipError_t launch_raster(hipStream_t s, const RasterArgs& a, int mode, bool count, int blocks);
// Sub-list capacity launch_compact needs for `pixels` (gated or not).
uint32_t compact_list_cap(uint32_t pixels, bool gated);
hipError_t launch_compact(hipStream_t s, int mode, const uint64_t* center, const uint64_t* cons, uint32_t pixels,
                          const PixelList& out, float4* color_out, float* depth_out, const uint32_t* gate = nullptr);
hipError_t launch_shade(hipStream_t s, const ShadeArgs& a, const PixelList& list, uint32_t pixels, int blocks);
hipError_t launch_filter(hipStream_t s, const FilterArgs& a, const PixelList& list, uint32_t pixels, int blocks);
// The filter over every 16 x 16 tile of the screen, for the pixels with a
// conservative word (the same pixels and results as launch_filter with the
// camera tiles' lists; it writes nothing for the other pixels).
hipError_t launch_filter_tiles(hipStream_t s, const FilterArgs& a, uint32_t width, uint32_t height);

}  // namespace hm
