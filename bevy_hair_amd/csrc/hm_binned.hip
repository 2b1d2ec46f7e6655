// Binned deep-opacity-map raster. The two DOM passes of raster.wgsl
// (draw_dom_segment / write_dom_pixel, raster.wgsl:118-190, run with
// pass_mode.x = 1 and then 2 by pass.rs:560-583) are evaluated per light-space
// tile in LDS, and each tile's finished texels are written to HBM once.
//
//   k_bin_gen        zero the status words and per-tile counts; two strands
//                    per wave: control points and their projection to HBM,
//                    each segment's tile box, long segments to a list
//   k_bin_count      per-tile counts: LDS histogram per workgroup, one global
//                    atomic per (workgroup, tile)
//   k_bin_scan       exclusive scan of the counts; gate words; the long-segment
//                    count to its status word, the append cursor back to zero
//   k_bin_fill       segment ids into per-tile record lists (same histogram,
//                    one global reservation per (workgroup, tile)); then the
//                    DOM fronts to empty             (only with long segments)
//   k_dom_long<1>    long segments, global atomic min  (same)
//   k_dom_tiles<T>   per tile: depth pass (LDS min), front to HBM, opacity
//                    pass (LDS add), counts to HBM: the whole DOM is written,
//                    so no separate clear is needed
//   k_dom_long<2>    long segments, global atomic add  (same)
// The camera pass bins the same way into screen tiles (k_bin_gen<0>); its
// tiles are split into chunks of records so that the few tiles holding most
// of the hair spread over many workgroups (k_cam_tiles, k_cam_finish).
// When the records outgrow their buffer the gates switch every binned kernel
// off and the direct kernels (clear + k_raster<1> + k_raster<2>) run instead.
// DOM words are built from min and add only, so the result does not depend on
// the order in which segments reach a texel and equals the direct path's bit
// for bit.
#include "hm_kernels.h"

#include "hm_device.h"
#include "hm_lists.h"
#include "hm_raster.h"

namespace hm {

namespace {

__device__ __forceinline__ const View& bin_view(const BinArgs& a) { return *reinterpret_cast<const View*>(&a.view); }

__device__ __forceinline__ int64_t mx64(int64_t a, int64_t b) { return a > b ? a : b; }
__device__ __forceinline__ int64_t mn64(int64_t a, int64_t b) { return a < b ? a : b; }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Texels the Bresenham walk of Drawer::dom_walk can write between projected
// end points a and b, clipped to the map, as a tile box. The walk moves each
// coordinate by at most one per step toward its end point and draws at most
// 1024 pixels; when both spans are below 2^20 the error term cannot overflow
// and the walk stays inside the end points' bounding box.
__device__ __forceinline__ uint32_t segment_box(const View& p, F2 a, F2 b, uint32_t shift, uint32_t box_limit,
                                                bool& is_long) {
    is_long = false;
    int64_t x0 = to_i32(floorf(a.x)), y0 = to_i32(floorf(a.y));
    int64_t x1 = to_i32(floorf(b.x)), y1 = to_i32(floorf(b.y));
    int64_t lx = x0 - 1023, hx = x0 + 1023, ly = y0 - 1023, hy = y0 + 1023;
    int64_t spanx = x1 > x0 ? x1 - x0 : x0 - x1, spany = y1 > y0 ? y1 - y0 : y0 - y1;
    if (spanx < (1 << 20) && spany < (1 << 20)) {
        lx = mx64(lx, x0 < x1 ? x0 : x1);
        hx = mn64(hx, x0 < x1 ? x1 : x0);
        ly = mx64(ly, y0 < y1 ? y0 : y1);
        hy = mn64(hy, y0 < y1 ? y1 : y0);
    }
    const int64_t w = to_i32(p.screen.x), h = to_i32(p.screen.y);
    lx = mx64(lx, 0);
    ly = mx64(ly, 0);
    hx = mn64(hx, w - 1);
    hy = mn64(hy, h - 1);
    if (lx > hx || ly > hy) return BIN_NONE;
    uint32_t tx0 = (uint32_t)lx >> shift, tx1 = (uint32_t)hx >> shift;
    uint32_t ty0 = (uint32_t)ly >> shift, ty1 = (uint32_t)hy >> shift;
    uint32_t bw = tx1 - tx0 + 1u, bh = ty1 - ty0 + 1u;
    if (bw * bh > box_limit) {
        is_long = true;
        return BIN_NONE;
    }
    return tx0 | (ty0 << 10) | ((bw - 1u) << 20) | ((bh - 1u) << 26);
}

// Pixels the camera DDA of Drawer::segment_walk can write between projected
// end points a and b, clipped to the screen, as a tile box. Samples are
// floor(mix(a, b, t)) for t in [0, 1] (plus the two 4-connected pixels
// formed from consecutive samples); below 2^21 in magnitude the rounding of
// mix moves a sample by less than half a pixel, so one pixel of margin covers
// every sample. Larger or non-finite coordinates are drawn by the long path.
__device__ __forceinline__ uint32_t camera_box(const View& p, F2 a, F2 b, uint32_t shift, uint32_t box_limit,
                                               bool& is_long) {
    is_long = false;
    const float m = mx(mx(fabsf(a.x), fabsf(a.y)), mx(fabsf(b.x), fabsf(b.y)));
    if (!(m < 2097152.0f)) {
        is_long = true;
        return BIN_NONE;
    }
    int64_t lx = (int64_t)floorf(mn(a.x, b.x)) - 1, hx = (int64_t)floorf(mx(a.x, b.x)) + 1;
    int64_t ly = (int64_t)floorf(mn(a.y, b.y)) - 1, hy = (int64_t)floorf(mx(a.y, b.y)) + 1;
    const int64_t w = to_i32(p.screen.x), h = to_i32(p.screen.y);
    lx = mx64(lx, 0);
    ly = mx64(ly, 0);
    hx = mn64(hx, w - 1);
    hy = mn64(hy, h - 1);
    if (lx > hx || ly > hy) return BIN_NONE;
    uint32_t tx0 = (uint32_t)lx >> shift, tx1 = (uint32_t)hx >> shift;
    uint32_t ty0 = (uint32_t)ly >> shift, ty1 = (uint32_t)hy >> shift;
    uint32_t bw = tx1 - tx0 + 1u, bh = ty1 - ty0 + 1u;
    if (bw * bh > box_limit) {
        is_long = true;
        return BIN_NONE;
    }
    return tx0 | (ty0 << 10) | ((bw - 1u) << 20) | ((bh - 1u) << 26);
}

// Calls f(tile) for every tile of a box, row by row. Most boxes are one
// tile: that case takes no loop.
template <class F>
__device__ __forceinline__ void for_box_tiles(uint32_t box, uint32_t tiles_x, F f) {
    if (box == BIN_NONE) return;
    uint32_t row = ((box >> 10) & 1023u) * tiles_x + (box & 1023u);
    if ((box >> 20) == 0u) {
        f(row);
        return;
    }
    const uint32_t bw = ((box >> 20) & 63u) + 1u, bh = ((box >> 26) & 63u) + 1u;
#pragma nounroll
    for (uint32_t y = 0; y < bh; ++y, row += tiles_x)
#pragma nounroll
        for (uint32_t x = 0; x < bw; ++x) f(row + x);
}

// Appends flagged lanes' segment ids to the long list with one atomic per wave.
__device__ __forceinline__ void append_long(bool flag, uint32_t id, uint32_t* count, uint32_t* list, uint32_t cap) {
    uint64_t m = __ballot(flag);
    if (m == 0u) return;
    const int lane = (int)(threadIdx.x & 63u);
    int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    uint32_t pos = base + lanes_below(m);
    if (flag && pos < cap) list[pos] = id;
}

// Tiles (or list entries) per thread whose global round trips the binning
// kernels keep in flight together.
constexpr uint32_t BIN_BATCH = 8u;
// Segment boxes per thread per round of k_bin_count / k_bin_fill.
constexpr uint32_t BIN_SEGS = 16u;
// Records per thread a DOM tile keeps in registers between its two passes.
constexpr uint32_t DOM_KEEP = 2u;

// Workgroup b takes the contiguous run [first, end) of the n segment-box
// words, in whole uint4 groups (n is a multiple of 4: 128 slots per strand);
// its threads interleave within the run.
struct Run {
    uint32_t first, end;  // in uint4 groups
};

__device__ __forceinline__ Run workgroup_run(uint32_t n) {
    const uint32_t groups = n / 4u;
    const uint32_t per = (groups + gridDim.x - 1u) / gridDim.x;
    uint32_t first = blockIdx.x * per;
    first = first < groups ? first : groups;
    const uint32_t end = first + per < groups ? first + per : groups;
    return {first, end};
}

}  // namespace

// Two strands per wave: the half-wave at lane gb (0 or 32) takes strand
// s = 2 * pair + gb / 32, and lane gb + r holds its control points
// idx = r + 32 * j for the S slots j < S = ceil(max ncp / 32) of the pair, so
// a pair of 65-point strands takes 3 slot passes where one strand per wave
// took 2 each. Neighbouring control points are in lane r +- 1 of the same
// slot, or in lane 31 of slot j - 1 / lane 0 of slot j + 1 of the same half.
// MODE 0: camera (also stores world tangents, ao and per-strand uv / beta /
// control-point count for the payload); MODE 1: deep opacity map.
template <int MODE, int S>
__device__ __forceinline__ void gen_pair(const BinArgs& a, const View& p, uint32_t s, bool live, uint32_t reuse_total,
                                         int r, int gb) {
    const int lane = gb + r;
    const hm_strand_ref ref = a.refs[s];
    const size_t base = (size_t)s * BIN_STRIDE;
    const uint32_t ncp_live = ref.control_points < 2u ? 2u : (ref.control_points > 127u ? 127u : ref.control_points);
    const uint32_t ncp = live ? ncp_live : 0u;  // a missing second strand stores nothing
    // DOM pass: when the camera pass of this frame generated the same strand
    // (same bundle, strand and control-point count) its world control points
    // are already in cps; otherwise generate them.
    bool reuse = false;
    if (MODE == 1 && s < reuse_total) {
        const hm_strand_ref c = a.reuse_refs[s];
        reuse = c.bundle == ref.bundle && c.strand == ref.strand && c.control_points == ref.control_points;
    }
    F3 cw[S], ct[S];
    float ao[S];
    StrandLayers L;
    if (reuse) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const uint32_t idx = (uint32_t)r + 32u * (uint32_t)j;
            const float4 q = idx < ncp ? a.cps[base + idx] : make_float4(0, 0, 0, 0);
            cw[j] = F3{q.x, q.y, q.z};
        }
    } else {
        L = strand_layers(ref, a.bundles, a.layers, a.style, r, gb);
        F3 cp[S];
        float w;
#pragma unroll
        for (int j = 0; j < S; ++j) cp[j] = strand_cp(L, (uint32_t)r + 32u * (uint32_t)j, gb, w, ao[j]);
        if (MODE == 0) {
            // Control-point tangents (raster.wgsl:286-317).
#pragma unroll
            for (int j = 0; j < S; ++j) {
                const uint32_t idx = (uint32_t)r + 32u * (uint32_t)j;
                const F3 up = shfl3(cp[j], r > 0 ? lane - 1 : lane);
                const F3 dn = shfl3(cp[j], r < 31 ? lane + 1 : lane);
                const F3 up_prev = j > 0 ? shfl3(cp[j > 0 ? j - 1 : 0], gb + 31) : cp[j];
                const F3 dn_next = j + 1 < S ? shfl3(cp[j + 1 < S ? j + 1 : j], gb) : cp[j];
                const F3 prev = r > 0 ? up : up_prev;
                const F3 next = idx + 1u < ncp_live ? (r < 31 ? dn : dn_next) : cp[j];
                ct[j] = safe_normalize(calc_tangent(idx, ncp_live, prev, cp[j], next));
            }
        }
#pragma unroll
        for (int j = 0; j < S; ++j) cw[j] = to_world(p, cp[j]);
    }
    Projected pr[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t idx = (uint32_t)r + 32u * (uint32_t)j;
        pr[j] = project_world(p, cw[j]);
        if (idx < ncp && MODE == 1) {
            // a segment's two end points are 40 contiguous bytes for the DOM tiles
            float* d = a.dcp + (base + idx) * 5u;
            d[0] = cw[j].x;
            d[1] = cw[j].y;
            d[2] = cw[j].z;
            d[3] = pr[j].screen.x;
            d[4] = pr[j].screen.y;
        }
        if (idx < ncp && MODE == 0) {
            a.cps[base + idx] = make_float4(cw[j].x, cw[j].y, cw[j].z, pr[j].screen.x);
            a.cpy[base + idx] = pr[j].screen.y;
            // raster.wgsl:319-337: world tangents, ao; uv and beta per strand.
            const F3 wt = dir_world(p, ct[j]);
            a.cpt[base + idx] = make_float4(wt.x, wt.y, wt.z, ao[j]);
        }
    }
    if (MODE == 0 && r == 0 && live) a.sinfo[s] = make_float4(L.uv.x, L.uv.y, L.beta, __uint_as_float(ncp));
    // Segment idx -> idx + 1: the end point's projection is in the next lane,
    // or in lane 0 of the next slot.
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const uint32_t idx = (uint32_t)r + 32u * (uint32_t)j;
        const int nl = r < 31 ? lane + 1 : lane;  // r = 31 of the last slot: idx + 1 >= ncp, unused
        F2 e = {__shfl(pr[j].screen.x, nl), __shfl(pr[j].screen.y, nl)};
        bool ef = __shfl((int)pr[j].in_front, nl) != 0;
        if (j + 1 < S) {
            const Projected& q = pr[j + 1 < S ? j + 1 : j];
            const F2 e1 = {__shfl(q.screen.x, gb), __shfl(q.screen.y, gb)};
            const bool ef1 = __shfl((int)q.in_front, gb) != 0;
            if (r == 31) {
                e = e1;
                ef = ef1;
            }
        }
        const uint32_t id = (s << a.seg_shift) + idx;
        bool is_long = false;
        uint32_t box = BIN_NONE;
        if (idx + 1u < ncp && (pr[j].in_front || ef))
            box = MODE == 0 ? camera_box(p, pr[j].screen, e, a.tile_shift, a.box_limit, is_long)
                            : segment_box(p, pr[j].screen, e, a.tile_shift, a.box_limit, is_long);
        // Every segment slot of the strand is written, so the binning passes
        // can walk segment ids without the strand's control-point count (the
        // host's slot count covers every segment: idx < 2^seg_shift here).
        if (live && idx < (1u << a.seg_shift)) a.seg_box[id] = box;
        append_long(is_long, id, a.meta + BIN_LONG_APPEND, a.long_list, a.long_capacity);
    }
#pragma unroll
    for (int j = S; j < (int)(BIN_STRIDE / 32u); ++j) {
        const uint32_t idx = (uint32_t)r + 32u * (uint32_t)j;
        if (live && idx < (1u << a.seg_shift)) a.seg_box[(s << a.seg_shift) + idx] = BIN_NONE;
    }
}

// MAXS: the most slots a pair of the pass can need, ceil(max ncp / 32), from
// the host's control-point bound. Passes of at most 96 control points (65 with
// LOD on) leave out the four-slot instance, whose register arrays set the
// kernel's register count, and ask for five (camera) or six (DOM) waves per
// SIMD.
template <int MODE, int MAXS>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MAXS <= 3 ? (MODE == 1 ? 6 : 5) : 1, 8))) void k_bin_gen(BinArgs a) {
    const View& p = bin_view(a);
    // Status words and tile counts to zero for this pass's k_bin_count and
    // k_bin_scan; the long-segment cursor, the only word written here, is
    // already zero.
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < BIN_META_WORDS + a.tiles; i += gridDim.x * blockDim.x)
        if (i != BIN_LONG_APPEND) a.meta[i] = 0u;
    const int lane = (int)(threadIdx.x & 63u), r = lane & 31, gb = lane & 32;
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (threadIdx.x >> 6));
    uint32_t reuse_total = 0u;
    if (MODE == 1 && a.reuse_refs) {
        reuse_total = a.reuse_indirect[0];
        reuse_total = reuse_total < a.capacity ? reuse_total : a.capacity;
    }
    for (uint32_t pair = w0; 2u * pair < total; pair += gridDim.x * wpb) {
        const uint32_t s_want = 2u * pair + (gb ? 1u : 0u);
        const bool live = s_want < total;
        const uint32_t s = live ? s_want : 2u * pair;  // the missing second strand of an odd total
        // slots for the longer strand of the pair (wave-uniform)
        const uint32_t c0 = a.refs[2u * pair].control_points;
        const uint32_t c1 = 2u * pair + 1u < total ? a.refs[2u * pair + 1u].control_points : 0u;
        uint32_t need = c0 > c1 ? c0 : c1;
        need = need < 2u ? 2u : (need > 127u ? 127u : need);
        const uint32_t slots = __builtin_amdgcn_readfirstlane((need + 31u) / 32u);
        if (slots <= 1u) gen_pair<MODE, 1>(a, p, s, live, reuse_total, r, gb);
        else if (slots == 2u) gen_pair<MODE, 2>(a, p, s, live, reuse_total, r, gb);
        else if (MAXS <= 3 || slots == 3u) gen_pair<MODE, 3>(a, p, s, live, reuse_total, r, gb);
        else gen_pair<MODE, MAXS <= 3 ? 3 : 4>(a, p, s, live, reuse_total, r, gb);
    }
}

// Per-workgroup tile histogram in LDS (dynamic size: one word per tile).
// Each workgroup adds up the (segment, tile) pairs of one contiguous run of
// segment ids, adds each nonzero entry to the global count once, and keeps
// the count's previous value, its rank within the tile, as a (tile, rank)
// list for k_bin_fill.
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_bin_count(BinArgs a) {
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t touched, list_base, list_pos;
    for (uint32_t t = threadIdx.x; t < a.tiles; t += blockDim.x) hist[t] = 0u;
    if (threadIdx.x == 0) {
        touched = 0u;
        list_pos = 0u;
    }
    __syncthreads();
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    const Run run = workgroup_run(total << a.seg_shift);
    // Thread i takes segment ids first + i + k * blockDim.x: a wave's lanes
    // hold consecutive segments, so the loads are coalesced dwords and, in
    // k_bin_fill, a tile's slots go to consecutive segments in lane order.
    const uint32_t wend = run.end * 4u;
    for (uint32_t r0 = run.first * 4u; r0 < wend; r0 += BIN_SEGS * blockDim.x) {
        uint32_t bx[BIN_SEGS];  // all loads in flight
#pragma unroll
        for (uint32_t k = 0; k < BIN_SEGS; ++k) {
            const uint32_t w = r0 + threadIdx.x + k * blockDim.x;
            bx[k] = w < wend ? a.seg_box[w] : BIN_NONE;
        }
#pragma unroll
        for (uint32_t k = 0; k < BIN_SEGS; ++k)
            for_box_tiles(bx[k], a.tiles_x, [&](uint32_t t) { atomicAdd(hist + t, 1u); });
    }
    __syncthreads();
    uint32_t mine = 0u;
    for (uint32_t t = threadIdx.x; t < a.tiles; t += blockDim.x) mine += hist[t] ? 1u : 0u;
    if (mine) atomicAdd(&touched, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        list_base = touched ? atomicAdd(a.meta + BIN_LIST_COUNT, touched) : 0u;
        a.wg_list[blockIdx.x] = make_uint2(list_base, touched);
    }
    __syncthreads();
    uint32_t* count = a.meta + BIN_META_WORDS;
    // BIN_BATCH tiles per thread at a time: their global atomics are all in
    // flight before the first result is used.
    for (uint32_t t0 = threadIdx.x; t0 < a.tiles; t0 += BIN_BATCH * blockDim.x) {
        uint32_t h[BIN_BATCH], rank[BIN_BATCH];
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k) {
            const uint32_t t = t0 + k * blockDim.x;
            h[k] = t < a.tiles ? hist[t] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k) rank[k] = h[k] ? atomicAdd(count + t0 + k * blockDim.x, h[k]) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k) {
            if (!h[k]) continue;
            const uint32_t pos = list_base + atomicAdd(&list_pos, 1u);
            if (pos < a.list_capacity) a.lists[pos] = make_uint2(t0 + k * blockDim.x, rank[k]);
        }
    }
}

// Exclusive scan of the per-tile counts (one block of 1024 threads, 16 tiles
// per thread), then the gate words. Counts and results pass through LDS so
// that every global load and store is coalesced.
// With chunk > 0 it also splits each tile's records into chunks of at most
// `chunk` (at least one chunk per tile, so empty tiles are still written):
// chunk offsets after the cursors, the tile of every chunk in chunk_tile,
// and BIN_SPLIT set when some tile has more than one chunk.
// With `busy` it also lists the tiles that hold records, in tile order, and
// their number in BIN_BUSY.
// NT threads over at most MAXT tiles. A grid of at most 2048 tiles (the
// camera pass at 1080p) takes the 256-thread, 9 KB variant: a workgroup that
// small starts as soon as one workgroup of the kernel running beside it on
// the other stream retires, where a 1024-thread one with 70 KB of LDS waits
// for a CU to drain (measured: the camera's scan took 61 us instead of 12
// beside the DOM generation).
template <uint32_t NT, uint32_t MAXT>
__global__ __launch_bounds__(NT) void k_bin_scan(uint32_t* meta, uint32_t tiles, uint32_t record_capacity,
                                                 uint32_t long_capacity, uint32_t list_capacity, uint32_t chunk,
                                                 uint32_t* chunk_tile, uint32_t chunk_capacity, uint32_t* busy) {
    constexpr uint32_t PER = MAXT / NT, NW = NT / 64u;
    // Tile i sits at word i + i / PER: a thread's run of PER tiles starts
    // PER + 1 words after its neighbour's, so the run-wise accesses below hit
    // distinct banks, as do the tile-wise ones.
    __shared__ uint32_t stage[MAXT + MAXT / PER];
    __shared__ uint32_t wave_total[NW], wave_chunks[NW], wave_busy[NW], split, ok_s;
    auto pad = [](uint32_t i) { return i + i / PER; };
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t* cnt = meta + BIN_META_WORDS;
    uint32_t* off = meta + BIN_META_WORDS + tiles;
    uint32_t* cur = off + tiles + 1u;
    uint32_t* choff = cur + tiles;
    if (tid == 0) split = 0u;
    // All loads in flight at once: every lane loads (a clamped index, so no
    // load sits behind a branch and its own wait), then masks.
    uint32_t ld[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = tid + NT * k;
        ld[k] = cnt[i < tiles ? i : 0u];
    }
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = tid + NT * k;
        stage[pad(i)] = i < tiles ? ld[k] : 0u;
    }
    __syncthreads();
    const uint32_t b = tid * PER;
    uint32_t v[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) v[k] = stage[pad(b + k)];
    uint32_t sum = 0u, csum = 0u, bsum = 0u, nch[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        sum += v[k];
        nch[k] = b + k < tiles ? (chunk && v[k] > chunk ? (v[k] + chunk - 1u) / chunk : 1u) : 0u;
        csum += nch[k];
        bsum += (b + k < tiles && v[k] != 0u) ? 1u : 0u;
    }
    uint32_t inc = sum, cinc = csum, binc = bsum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d), cy = __shfl_up(cinc, d), by = __shfl_up(binc, d);
        if ((int)lane >= d) {
            inc += y;
            cinc += cy;
            binc += by;
        }
    }
    if (lane == 63u) {
        wave_total[wave] = inc;
        wave_chunks[wave] = cinc;
        wave_busy[wave] = binc;
    }
    __syncthreads();
    uint32_t run = inc - sum, crun = cinc - csum, brun = binc - bsum;
    for (uint32_t w = 0; w < wave; ++w) {
        run += wave_total[w];
        crun += wave_chunks[w];
        brun += wave_busy[w];
    }
    if (busy) {
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k)
            if (b + k < tiles && v[k] != 0u) busy[brun++] = b + k;
        if (tid == NT - 1u) meta[BIN_BUSY] = brun;
    }
    // Record offsets through LDS, then out (offset and cursor arrays).
    uint32_t r = run, cr = crun;
    bool any_split = false;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        stage[pad(b + k)] = r;
        r += v[k];
        any_split |= nch[k] > 1u;
    }
    if (any_split) split = 1u;
    __syncthreads();
    for (uint32_t i = tid; i < tiles; i += NT) {
        const uint32_t o = stage[pad(i)];
        off[i] = o;
        cur[i] = o;
    }
    __syncthreads();
    // Chunk offsets the same way; then each chunk finds its tile by binary
    // search over them (coalesced chunk_tile stores).
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        stage[pad(b + k)] = cr;
        cr += nch[k];
    }
    __syncthreads();
    uint32_t chunks = 0u;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) chunks += wave_chunks[w];
    for (uint32_t i = tid; i < tiles; i += NT) choff[i] = stage[pad(i)];
    const uint32_t nmap = !chunk_tile ? 0u : (chunks < chunk_capacity ? chunks : chunk_capacity);
    for (uint32_t c = tid; c < nmap; c += NT) {
        uint32_t lo = 0u, hi = tiles;  // last tile whose first chunk <= c
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (stage[pad(mid)] <= c) lo = mid;
            else hi = mid;
        }
        chunk_tile[c] = lo;
    }
    if (tid == NT - 1u) {
        const uint32_t total = r;
        off[tiles] = total;
        choff[tiles] = chunks;
        const uint32_t longs = meta[BIN_LONG_APPEND];
        meta[BIN_LONG_COUNT] = longs;
        meta[BIN_LONG_APPEND] = 0u;
        const uint32_t ok = (total <= record_capacity && longs <= long_capacity &&
                             meta[BIN_LIST_COUNT] <= list_capacity && chunks <= chunk_capacity)
                                ? 1u
                                : 0u;
        meta[BIN_RECORDS] = total;
        meta[BIN_OK] = ok;
        meta[BIN_FALLBACK] = 1u - ok;
        meta[BIN_LONG] = (ok && longs) ? 1u : 0u;
        ok_s = ok;
    }
    __syncthreads();
    if (tid == 0) meta[BIN_SPLIT] = (split && ok_s) ? 1u : 0u;
}

// Same runs as k_bin_count: the workgroup's (tile, rank) list gives the first
// slot of its block in each touched tile; slots are handed out in LDS.
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_bin_fill(BinArgs a) {
    if (a.meta[BIN_OK] == 0u) return;
    extern __shared__ uint32_t hist[];
    const uint32_t* off = a.meta + BIN_META_WORDS + a.tiles;
    const uint2 wl = a.wg_list[blockIdx.x];
    for (uint32_t i0 = threadIdx.x; i0 < wl.y; i0 += BIN_BATCH * blockDim.x) {
        uint2 e[BIN_BATCH];
        uint32_t o[BIN_BATCH];
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k) {
            const uint32_t i = i0 + k * blockDim.x;
            e[k] = i < wl.y ? a.lists[wl.x + i] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k) o[k] = i0 + k * blockDim.x < wl.y ? off[e[k].x] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < BIN_BATCH; ++k)
            if (i0 + k * blockDim.x < wl.y) hist[e[k].x] = o[k] + e[k].y;
    }
    __syncthreads();
    uint32_t total = a.indirect[0];
    total = total < a.capacity ? total : a.capacity;
    const Run run = workgroup_run(total << a.seg_shift);
    uint32_t* records = a.records;
    const uint32_t wend = run.end * 4u;
    for (uint32_t r0 = run.first * 4u; r0 < wend; r0 += BIN_SEGS * blockDim.x) {
        uint32_t bx[BIN_SEGS];  // all loads in flight
#pragma unroll
        for (uint32_t k = 0; k < BIN_SEGS; ++k) {
            const uint32_t w = r0 + threadIdx.x + k * blockDim.x;
            bx[k] = w < wend ? a.seg_box[w] : BIN_NONE;
        }
#pragma unroll
        for (uint32_t k = 0; k < BIN_SEGS; ++k) {
            const uint32_t id = r0 + threadIdx.x + k * blockDim.x;
            for_box_tiles(bx[k], a.tiles_x, [&](uint32_t t) { records[atomicAdd(hist + t, 1u)] = id; });
        }
    }
    // The clear the next kernels need, gated as they are: split camera tiles
    // take atomic minima into a cleared G-buffer, long DOM segments into
    // fronts set to the empty depth (no launch of its own when not needed).
    uint32_t* dst = nullptr;
    size_t words = 0;
    uint32_t value = 0u;
    if (a.fill_clear == 1u && a.meta[BIN_SPLIT] != 0u) {
        dst = a.gb_slab;
        words = (size_t)a.gb_words;
        value = 0xffffffffu;
    } else if (a.fill_clear == 2u && a.meta[BIN_LONG] != 0u) {
        dst = a.dom;
        words = (size_t)a.dom_size * a.dom_size;
        value = HM_DOM_EMPTY;
    }
    if (dst) {
        const size_t quads = words / 4u;  // both buffers are 16-byte aligned
        const uint4 v4 = make_uint4(value, value, value, value);
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < quads; i += (size_t)gridDim.x * blockDim.x)
            reinterpret_cast<uint4*>(dst)[i] = v4;
        for (size_t i = quads * 4u + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
             i += (size_t)gridDim.x * blockDim.x)
            dst[i] = value;
    }
}

// LDS tile of the map: fronts and 16 counts per texel.
template <int T>
struct DomTileSink {
    uint32_t* front;
    uint32_t* counts;
    int32_t ox, oy;
    __device__ __forceinline__ bool accepts(int32_t x, int32_t y) const {
        return (uint32_t)(x - ox) < (uint32_t)T && (uint32_t)(y - oy) < (uint32_t)T;
    }
    // The walk is monotone in x and y: once past the tile in its direction of
    // travel it cannot come back.
    __device__ __forceinline__ bool walked_out(int32_t x, int32_t y, int32_t sx, int32_t sy) const {
        // bitwise, not short-circuit: selects rather than branches per step
        const bool out_x = (sx > 0) ? (x >= ox + T) : (x < ox);
        const bool out_y = (sy > 0) ? (y >= oy + T) : (y < oy);
        return out_x | out_y;
    }
    __device__ __forceinline__ uint32_t at(int32_t x, int32_t y) const { return (uint32_t)((y - oy) * T + (x - ox)); }
    __device__ __forceinline__ void dom_min(size_t, int32_t x, int32_t y, uint32_t q, float, Counters&) {
        atomicMin(front + at(x, y), q);
    }
    __device__ __forceinline__ uint32_t dom_front(size_t, int32_t x, int32_t y) const { return front[at(x, y)]; }
    __device__ __forceinline__ void dom_add(size_t, int32_t x, int32_t y, uint32_t slice, Counters&) {
        atomicAdd(counts + at(x, y) * HM_DOM_LAYERS + slice, 1u);
    }
};

// One segment of the map from the stored control points (world position and
// light-space projection of both ends, raster.wgsl:150-156).
struct RecordEnds {
    float4 c0, c1;  // world xyz, projected x
    float y0, y1;   // projected y
    bool first;     // segment 0 of its strand
};

__device__ __forceinline__ RecordEnds load_record(const BinArgs& a, uint32_t id) {
    const float* d = a.dcp + (size_t)seg_cp(a, id) * 5u;
    return {make_float4(d[0], d[1], d[2], d[3]), make_float4(d[5], d[6], d[7], d[8]), d[4], d[9],
            seg_index(a, id) == 0u};
}

template <class D>
__device__ __forceinline__ void walk_ends(D& d, const RecordEnds& e) {
    d.dom_walk(F3{e.c0.x, e.c0.y, e.c0.z}, F3{e.c1.x, e.c1.y, e.c1.z}, F2{e.c0.w, e.y0}, F2{e.c1.w, e.y1}, e.first);
}

template <class D>
__device__ __forceinline__ void walk_record(D& d, const BinArgs& a, uint32_t id) {
    walk_ends(d, load_record(a, id));
}

template <int T>
__global__ __launch_bounds__(256) void k_dom_tiles(BinArgs a) {
    if (a.meta[BIN_OK] == 0u) return;
    __shared__ uint32_t front[T * T];
    __shared__ __align__(16) uint32_t counts[T * T * HM_DOM_LAYERS];
    const View& p = bin_view(a);
    const uint32_t tile = blockIdx.x;
    const int32_t ox = (int32_t)(tile % a.tiles_x) * T, oy = (int32_t)(tile / a.tiles_x) * T;
    const int32_t S = (int32_t)a.dom_size;
    const bool has_long = a.meta[BIN_LONG] != 0u;
    uint32_t* dom_front = a.dom;
    uint4* dom_counts = reinterpret_cast<uint4*>(a.dom + (size_t)S * (size_t)S);
    for (int i = (int)threadIdx.x; i < T * T; i += 256) {
        int32_t x = ox + i % T, y = oy + i / T;
        front[i] = (has_long && x < S && y < S) ? dom_front[(size_t)y * S + x] : HM_DOM_EMPTY;
    }
    for (int i = (int)threadIdx.x; i < T * T * 4; i += 256) reinterpret_cast<uint4*>(counts)[i] = make_uint4(0, 0, 0, 0);
    __syncthreads();

    const uint32_t* off = a.meta + BIN_META_WORDS + a.tiles;
    const uint32_t begin = off[tile], end = off[tile + 1u];
    DomTileSink<T> sink{front, counts, ox, oy};
    // Each thread's first DOM_KEEP records stay in registers for the opacity
    // pass, which then reads only a crowded tile's further records again.
    RecordEnds keep[DOM_KEEP];
    {
        Drawer<1, DomTileSink<T>> d{p, nullptr, 0.0f, sink, {0, 0, 0, 0}};
        uint32_t r = begin + threadIdx.x;
#pragma unroll
        for (uint32_t j = 0; j < DOM_KEEP; ++j, r += 256u)
            if (r < end) {
                keep[j] = load_record(a, a.records[r]);
                walk_ends(d, keep[j]);
            }
        for (; r < end; r += 256u) walk_record(d, a, a.records[r]);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < T * T; i += 256) {
        int32_t x = ox + i % T, y = oy + i / T;
        if (x < S && y < S) dom_front[(size_t)y * S + x] = front[i];
    }
    {
        Drawer<2, DomTileSink<T>> d{p, nullptr, 0.0f, sink, {0, 0, 0, 0}};
        uint32_t r = begin + threadIdx.x;
#pragma unroll
        for (uint32_t j = 0; j < DOM_KEEP; ++j, r += 256u)
            if (r < end) walk_ends(d, keep[j]);
        for (; r < end; r += 256u) walk_record(d, a, a.records[r]);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < T * T * 4; i += 256) {
        int texel = i >> 2;
        int32_t x = ox + texel % T, y = oy + texel / T;
        if (x < S && y < S) dom_counts[((size_t)y * S + x) * 4u + (uint32_t)(i & 3)] = reinterpret_cast<uint4*>(counts)[i];
    }
}

// 16 x 16 DOM tiles, persistent: workgroups of 1024 threads take the tiles
// that hold records from k_bin_scan's busy list (one counter), then 16-tile
// runs of the grid whose empty tiles they write (another counter). A busy
// tile walks its records once, in the depth pass: each sample takes its LDS
// atomic minimum and is kept in an LDS stash (texel, view distance), from
// which the opacity pass bins it once every front is final. Only samples the
// stash cannot hold are walked a second time: a thread's records from the one
// where its first sample missed the stash on, skipping that record's samples
// the stash took. Both passes test the same pixels in the same order, so the
// skipped samples are exactly the stashed ones.
constexpr uint32_t DOM_STASH = 12288u;  // samples per busy tile (60 KB of LDS)
constexpr uint32_t DOM_WG = 1024u;
constexpr uint32_t DOM_NO_FAIL = 0xffffffffu;

struct DomStashSink {
    static constexpr int T = 16;
    uint32_t* front;
    uint32_t* counts;
    int32_t ox, oy;
    float* s_dist;
    uint8_t* s_pix;
    uint32_t* s_n;
    // (record ordinal of this thread) << 11 | samples of that record so far
    // (a walk draws at most 1024), now and at the first sample the stash
    // could not take
    uint32_t pos = 0u;
    uint32_t fail = DOM_NO_FAIL;
    uint32_t skip = 0u;       // opacity re-walk: samples still to pass over
    __device__ __forceinline__ bool accepts(int32_t x, int32_t y) const {
        return (uint32_t)(x - ox) < (uint32_t)T && (uint32_t)(y - oy) < (uint32_t)T;
    }
    __device__ __forceinline__ bool walked_out(int32_t x, int32_t y, int32_t sx, int32_t sy) const {
        const bool out_x = (sx > 0) ? (x >= ox + T) : (x < ox);
        const bool out_y = (sy > 0) ? (y >= oy + T) : (y < oy);
        return out_x | out_y;
    }
    __device__ __forceinline__ uint32_t at(int32_t x, int32_t y) const { return (uint32_t)((y - oy) * T + (x - ox)); }
    __device__ __forceinline__ void dom_min(size_t, int32_t x, int32_t y, uint32_t q, float dist, Counters&) {
        const uint32_t i = at(x, y);
        atomicMin(front + i, q);
        // one stash reservation per wave for the lanes writing now
        const uint64_t m = __ballot(1);
        const uint32_t rank = lanes_below(m);
        uint32_t base = 0u;
        if (rank == 0u) base = atomicAdd(s_n, (uint32_t)__popcll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        const uint32_t slot = base + rank;
        if (fail == DOM_NO_FAIL) {
            if (slot < DOM_STASH) {
                s_dist[slot] = dist;
                s_pix[slot] = (uint8_t)i;
            } else {
                fail = pos;
            }
        }
        ++pos;
    }
    __device__ __forceinline__ uint32_t dom_front(size_t, int32_t x, int32_t y) {
        if (skip != 0u) {
            --skip;
            return HM_DOM_EMPTY;  // binned from the stash already
        }
        return front[at(x, y)];
    }
    __device__ __forceinline__ void dom_add(size_t, int32_t x, int32_t y, uint32_t slice, Counters&) {
        atomicAdd(counts + at(x, y) * HM_DOM_LAYERS + slice, 1u);
    }
};

__global__ __launch_bounds__(DOM_WG, 8) void k_dom_tiles_busy(BinArgs a) {
    if (a.meta[BIN_OK] == 0u) return;
    constexpr int T = 16;
    __shared__ uint32_t front[T * T];
    __shared__ __align__(16) uint32_t counts[T * T * HM_DOM_LAYERS];
    __shared__ float s_dist[DOM_STASH];
    __shared__ uint8_t s_pix[DOM_STASH];
    __shared__ uint32_t s_n, s_item, s_fail;
    const View& p = bin_view(a);
    const uint32_t tid = threadIdx.x;
    const int32_t S = (int32_t)a.dom_size;
    const bool has_long = a.meta[BIN_LONG] != 0u;
    const uint32_t* cnt = a.meta + BIN_META_WORDS;
    const uint32_t* off = cnt + a.tiles;
    const uint32_t* busy = cnt + 4u * a.tiles + 2u;
    const uint32_t n_busy = a.meta[BIN_BUSY];
    uint32_t* dom_front = a.dom;
    uint4* dom_counts = reinterpret_cast<uint4*>(a.dom + (size_t)S * (size_t)S);
    const DomSlab slab = dom_slab(p.screen.z, p.screen.w);
    // thread tid: front word tid (tid < 256) and count quad tid of a tile
    const int32_t fx = (int32_t)(tid % T), fy = (int32_t)((tid / T) % T);
    const int32_t qx = (int32_t)((tid >> 2) % T), qy = (int32_t)((tid >> 2) / T);
    for (;;) {
        if (tid == 0u) s_item = atomicAdd(a.meta + BIN_BUSY_NEXT, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= n_busy) break;
        const uint32_t tile = busy[item];
        const int32_t ox = (int32_t)(tile % a.tiles_x) * T, oy = (int32_t)(tile / a.tiles_x) * T;
        if (tid < (uint32_t)(T * T)) {
            const int32_t x = ox + fx, y = oy + fy;
            front[tid] = (has_long && x < S && y < S) ? dom_front[(size_t)y * S + x] : HM_DOM_EMPTY;
        }
        reinterpret_cast<uint4*>(counts)[tid] = make_uint4(0u, 0u, 0u, 0u);
        if (tid == 0u) {
            s_n = 0u;
            s_fail = 0u;
        }
        __syncthreads();
        const uint32_t begin = off[tile], end = off[tile + 1u];
        DomStashSink sink{front, counts, ox, oy, s_dist, s_pix, &s_n};
        {
            Drawer<1, DomStashSink> d{p, nullptr, 0.0f, sink, {0, 0, 0, 0}};
            uint32_t j = 0u;
            for (uint32_t r = begin + tid; r < end; r += DOM_WG, ++j) {
                sink.pos = j << 11;
                walk_record(d, a, a.records[r]);
            }
        }
        if (sink.fail != DOM_NO_FAIL) s_fail = 1u;
        __syncthreads();
        if (tid < (uint32_t)(T * T)) {
            const int32_t x = ox + fx, y = oy + fy;
            if (x < S && y < S) dom_front[(size_t)y * S + x] = front[tid];
        }
        const uint32_t n = s_n < DOM_STASH ? s_n : DOM_STASH;
        for (uint32_t i = tid; i < n; i += DOM_WG) {
            const uint32_t px = s_pix[i];
            const uint32_t fq = front[px];
            if (fq != HM_DOM_EMPTY) atomicAdd(counts + px * HM_DOM_LAYERS + dom_slice(slab, fq, s_dist[i]), 1u);
        }
        if (s_fail != 0u && sink.fail != DOM_NO_FAIL) {
            Drawer<2, DomStashSink> d{p, nullptr, 0.0f, sink, {0, 0, 0, 0}};
            const uint32_t j0 = sink.fail >> 11;
            uint32_t j = j0;
            for (uint32_t r = begin + tid + j * DOM_WG; r < end; r += DOM_WG, ++j) {
                sink.skip = j == j0 ? (sink.fail & 2047u) : 0u;
                walk_record(d, a, a.records[r]);
            }
        }
        __syncthreads();
        {
            const int32_t x = ox + qx, y = oy + qy;
            if (x < S && y < S) dom_counts[((size_t)y * S + x) * 4u + (tid & 3u)] = reinterpret_cast<uint4*>(counts)[tid];
        }
    }
    // Tiles without records: empty fronts (unless long segments wrote them)
    // and zero counts.
    const uint32_t runs = (a.tiles + 15u) / 16u;
    for (;;) {
        __syncthreads();  // s_item's last value has been read
        if (tid == 0u) s_item = atomicAdd(a.meta + BIN_EMPTY_NEXT, 1u);
        __syncthreads();
        const uint32_t run = s_item;
        if (run >= runs) break;
        const uint32_t t_end = run * 16u + 16u < a.tiles ? run * 16u + 16u : a.tiles;
        for (uint32_t t = run * 16u; t < t_end; ++t) {
            if (cnt[t] != 0u) continue;
            const int32_t ox = (int32_t)(t % a.tiles_x) * T, oy = (int32_t)(t / a.tiles_x) * T;
            const int32_t x = ox + qx, y = oy + qy;
            if (x < S && y < S) dom_counts[((size_t)y * S + x) * 4u + (tid & 3u)] = make_uint4(0u, 0u, 0u, 0u);
            if (!has_long && tid < (uint32_t)(T * T)) {
                const int32_t fx2 = ox + fx, fy2 = oy + fy;
                if (fx2 < S && fy2 < S) dom_front[(size_t)fy2 * S + fx2] = HM_DOM_EMPTY;
            }
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_dom_long(BinArgs a) {
    if (a.meta[BIN_LONG] == 0u) return;
    const View& p = bin_view(a);
    const uint32_t n = a.meta[BIN_LONG_COUNT];
    const size_t S = a.dom_size;
    GlobalSink sink{nullptr, nullptr, nullptr, a.dom, S * S, true};
    Drawer<MODE, GlobalSink> d{p, nullptr, 0.0f, sink, {0, 0, 0, 0}};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        walk_record(d, a, a.long_list[i]);
}

hipError_t launch_bin(hipStream_t s, const BinArgs& a, int mode, int gen_blocks, int bin_blocks, hipEvent_t after_gen) {
    if (a.tiles > BIN_MAX_TILES) return hipErrorInvalidValue;
    hipError_t e;
    // a strand has at most 2^seg_shift + 1 control points
    const bool four = a.seg_shift > 6u;
    if (mode == 0 && four) hipLaunchKernelGGL((k_bin_gen<0, 4>), dim3(gen_blocks), dim3(256), 0, s, a);
    else if (mode == 0) hipLaunchKernelGGL((k_bin_gen<0, 3>), dim3(gen_blocks), dim3(256), 0, s, a);
    else if (four) hipLaunchKernelGGL((k_bin_gen<1, 4>), dim3(gen_blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_bin_gen<1, 3>), dim3(gen_blocks), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (after_gen && (e = hipEventRecord(after_gen, s)) != hipSuccess) return e;
    const size_t lds = (size_t)a.tiles * sizeof(uint32_t);
    // A DOM pass's histogram (one word per tile: 64 KB at S = 2048 with 16-texel
    // tiles) leaves room for two workgroups per CU; with 1024 threads each they
    // keep eight waves per SIMD busy rather than two. The camera pass keeps
    // 256 threads (with 1024 the 2160p frame was slower).
    const bool wide = mode == 1 && a.tiles > 2048u;
    if (wide) hipLaunchKernelGGL(k_bin_count<1024u>, dim3(bin_blocks), dim3(1024), lds, s, a);
    else hipLaunchKernelGGL(k_bin_count<256u>, dim3(bin_blocks), dim3(256), lds, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t* busy = mode == 1 ? a.meta + BIN_META_WORDS + 4u * a.tiles + 2u : nullptr;
    if (a.tiles <= 2048u)
        hipLaunchKernelGGL((k_bin_scan<256u, 2048u>), dim3(1), dim3(256), 0, s, a.meta, a.tiles, a.record_capacity,
                           a.long_capacity, a.list_capacity, a.chunk, a.chunk_tile, a.chunk_capacity, busy);
    else
        hipLaunchKernelGGL((k_bin_scan<1024u, BIN_MAX_TILES>), dim3(1), dim3(1024), 0, s, a.meta, a.tiles,
                           a.record_capacity, a.long_capacity, a.list_capacity, a.chunk, a.chunk_tile,
                           a.chunk_capacity, busy);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (wide) hipLaunchKernelGGL(k_bin_fill<1024u>, dim3(bin_blocks), dim3(1024), lds, s, a);
    else hipLaunchKernelGGL(k_bin_fill<256u>, dim3(bin_blocks), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_dom_long(hipStream_t s, const BinArgs& a, int mode, int long_blocks) {
    // (before the tiles, k_bin_fill has set the fronts to empty when long segments exist)
    if (mode == 1) hipLaunchKernelGGL(k_dom_long<1>, dim3(long_blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_dom_long<2>, dim3(long_blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dom_tiles(hipStream_t s, const BinArgs& a, int busy_blocks) {
    if (a.tile_shift == 4u && busy_blocks > 0) {
        hipLaunchKernelGGL(k_dom_tiles_busy, dim3(busy_blocks), dim3(DOM_WG), 0, s, a);
    } else if (a.tile_shift == 4u) {
        hipLaunchKernelGGL(k_dom_tiles<16>, dim3(a.tiles), dim3(256), 0, s, a);
    } else if (a.tile_shift == 5u) {
        hipLaunchKernelGGL(k_dom_tiles<32>, dim3(a.tiles), dim3(256), 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- camera G-buffer ----------------------------------------------------------

// LDS tile of the G-buffer: centre and conservative words, beta words.
template <int T>
struct CamTileSink {
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    int32_t ox, oy;
    __device__ __forceinline__ bool accepts(int32_t x, int32_t y) const {
        return (uint32_t)(x - ox) < (uint32_t)T && (uint32_t)(y - oy) < (uint32_t)T;
    }
    __device__ __forceinline__ bool walked_out(int32_t, int32_t, int32_t, int32_t) const { return false; }
    __device__ __forceinline__ void cam(size_t, int32_t x, int32_t y, uint64_t payload, bool center_hit,
                                        uint32_t beta_word, Counters&) {
        const uint32_t i = (uint32_t)((y - oy) * T + (x - ox));
        atomicMin((unsigned long long*)(cons + i), (unsigned long long)payload);
        if (center_hit) {
            atomicMin((unsigned long long*)(center + i), (unsigned long long)payload);
            atomicMin(beta + i, beta_word);
        }
    }
};

// One camera segment from the stored control points (raster.wgsl:319-337 with
// the end points' projection, tangents, ao and the strand's uv and beta).
template <class D>
__device__ __forceinline__ void walk_cam_record(D& d, const BinArgs& a, uint32_t id) {
    const uint32_t cp = seg_cp(a, id);
    const float4 c0 = a.cps[cp], c1 = a.cps[cp + 1u];
    const float y0 = a.cpy[cp], y1 = a.cpy[cp + 1u];
    const float4 t0 = a.cpt[cp], t1 = a.cpt[cp + 1u];
    const float4 si = a.sinfo[seg_strand(a, id)];
    const uint32_t ncp = __float_as_uint(si.w);
    const uint32_t idx = seg_index(a, id);
    const float den = u2f(ncp - 1u > 1u ? ncp - 1u : 1u), inv_den = rdiv(1.0f, den);
    const float w0 = div_by_recip(u2f(idx), den, inv_den), w1 = div_by_recip(u2f(idx + 1u), den, inv_den);
    d.segment_walk(F3{c0.x, c0.y, c0.z}, F3{c1.x, c1.y, c1.z}, F2{c0.w, y0}, F2{c1.w, y1}, F3{t0.x, t0.y, t0.z},
                   F3{t1.x, t1.y, t1.z}, F3{si.x, si.y, w0}, F3{si.x, si.y, w1}, t0.w, t1.w, si.z);
}

// Final words of one pixel of a tile (every thread of the workgroup calls it):
// the G-buffer, and with emit_lists the pixel's entries in the shade / filter
// sub-lists of its tile and the empty results of uncovered pixels.
__device__ __forceinline__ void finish_pixel(const BinArgs& a, uint32_t tile, bool in, size_t g, uint64_t c,
                                             uint64_t k, uint32_t b, bool write_gbuffer, uint32_t* scratch) {
    if (in && write_gbuffer) {
        a.center[g] = c;
        a.cons[g] = k;
        a.beta[g] = b;
    }
    if (!a.emit_lists) return;
    const bool shade = in && !gb_empty(c);
    const bool filter = in && (shade || !gb_empty(k));
    if (in && !shade) {
        a.shaded[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.shaded_depth[g] = 0.0f;
    }
    if (in && !filter) {
        a.filtered[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.filtered_depth[g] = 0.0f;
    }
    const uint32_t sub = tile % LIST_SUBS;
    block_append(PixelList{a.shade_list, a.meta + BIN_SHADE_SUBS, a.list_cap}, sub, shade, (uint32_t)g, scratch);
    block_append(PixelList{a.filter_list, a.meta + BIN_FILTER_SUBS, a.list_cap}, sub, filter, (uint32_t)g, scratch);
}

// One workgroup per chunk of a screen tile's records. The tile starts empty
// and takes its segments with LDS atomics. A tile with one chunk is written as
// it is (this is its clear); the chunks of a larger tile take the minimum into
// the G-buffer, cleared beforehand, with global atomics, and k_cam_finish
// then emits that tile's pixel lists. Six waves per SIMD fit in 79 VGPRs
// without spills (the compiler's own choice, 87, allowed five); seven spill.
template <int T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_cam_tiles(BinArgs a) {
    static_assert((T * T) % 256 == 0, "the write-back loop runs whole waves");
    if (a.meta[BIN_OK] == 0u) return;
    const uint32_t* off = a.meta + BIN_META_WORDS + a.tiles;
    const uint32_t* choff = off + 2u * a.tiles + 1u;
    const uint32_t nchunks = choff[a.tiles];
    __shared__ uint64_t center[T * T], cons[T * T];
    __shared__ uint32_t beta[T * T];
    __shared__ uint32_t scratch[5];
    const View& p = bin_view(a);
    const int32_t W = (int32_t)a.width, H = (int32_t)a.height;
    // The grid covers the usual number of chunks with one workgroup each (the
    // hardware dispatcher balances them); a grid-stride loop takes any rest.
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t tile = a.chunk_tile[chunk];
        const uint32_t first = choff[tile], nch = choff[tile + 1u] - first;
        const int32_t ox = (int32_t)(tile % a.tiles_x) * T, oy = (int32_t)(tile / a.tiles_x) * T;
        for (int i = (int)threadIdx.x; i < T * T; i += 256) {
            center[i] = ~0ull;
            cons[i] = ~0ull;
            beta[i] = ~0u;
        }
        __syncthreads();
        const uint32_t tile_end = off[tile + 1u];
        const uint32_t begin = off[tile] + (chunk - first) * (a.chunk ? a.chunk : 0u);
        const uint32_t end = nch > 1u && begin + a.chunk < tile_end ? begin + a.chunk : tile_end;
        CamTileSink<T> sink{center, cons, beta, ox, oy};
        Drawer<0, CamTileSink<T>> d{p, a.scene_depth, a.scene_clear, sink, {0, 0, 0, 0}};
        for (uint32_t r = begin + threadIdx.x; r < end; r += 256u) walk_cam_record(d, a, a.records[r]);
        __syncthreads();
        for (int i = (int)threadIdx.x; i < T * T; i += 256) {  // T * T is a multiple of 256: whole waves
            const int32_t x = ox + i % T, y = oy + i / T;
            const bool in = x < W && y < H;
            const size_t g = in ? (size_t)y * (size_t)W + (size_t)x : 0u;
            if (nch > 1u) {
                if (in) {
                    if (cons[i] != ~0ull) atomicMin((unsigned long long*)a.cons + g, (unsigned long long)cons[i]);
                    if (center[i] != ~0ull) atomicMin((unsigned long long*)a.center + g, (unsigned long long)center[i]);
                    if (beta[i] != ~0u) atomicMin(a.beta + g, beta[i]);
                }
            } else {
                finish_pixel(a, tile, in, g, center[i], cons[i], beta[i], true, scratch);
            }
        }
        __syncthreads();  // the next chunk reuses the LDS tile
    }
}

// Tiles that were split into chunks: their finished words are in the
// G-buffer; emit the pixel lists (one workgroup per tile; the grid-stride
// loop only guards a smaller grid).
template <int T>
__global__ __launch_bounds__(256) void k_cam_finish(BinArgs a) {
    if (a.meta[BIN_SPLIT] == 0u || !a.emit_lists) return;
    const uint32_t* choff = a.meta + BIN_META_WORDS + 3u * a.tiles + 1u;
    __shared__ uint32_t scratch[5];
    const int32_t W = (int32_t)a.width, H = (int32_t)a.height;
    for (uint32_t tile = blockIdx.x; tile < a.tiles; tile += gridDim.x) {
        if (choff[tile + 1u] - choff[tile] <= 1u) continue;
        const int32_t ox = (int32_t)(tile % a.tiles_x) * T, oy = (int32_t)(tile / a.tiles_x) * T;
        for (int i = (int)threadIdx.x; i < T * T; i += 256) {
            const int32_t x = ox + i % T, y = oy + i / T;
            const bool in = x < W && y < H;
            const size_t g = in ? (size_t)y * (size_t)W + (size_t)x : 0u;
            const uint64_t c = in ? a.center[g] : ~0ull, k = in ? a.cons[g] : ~0ull;
            finish_pixel(a, tile, in, g, c, k, 0u, false, scratch);
        }
    }
}

// Long camera segments go to the G-buffer after the tiles have written it
// and emitted the pixel lists. The write that first covers a pixel (its
// atomic returns the empty word) appends the pixel to its tile's sub-list,
// so pixels that only long segments reach are shaded and filtered as well.
// Every centre write comes with a conservative write of the same payload,
// so "conservative word set" is the filter list's condition.
struct LongCamSink {
    uint64_t* center;
    uint64_t* cons;
    uint32_t* beta;
    PixelList shade, filter;
    bool emit;
    uint32_t tile_shift, tiles_x;
    __device__ __forceinline__ bool accepts(int32_t, int32_t) const { return true; }
    __device__ __forceinline__ bool walked_out(int32_t, int32_t, int32_t, int32_t) const { return false; }
    __device__ __forceinline__ void append(const PixelList& L, int32_t x, int32_t y, size_t pix) const {
        const uint32_t sub = ((((uint32_t)y >> tile_shift) * tiles_x) + ((uint32_t)x >> tile_shift)) % LIST_SUBS;
        const uint32_t pos = atomicAdd(L.counts + sub, 1u);
        if (pos < L.cap) L.list[(size_t)sub * L.cap + pos] = (uint32_t)pix;
    }
    __device__ __forceinline__ void cam(size_t pix, int32_t x, int32_t y, uint64_t payload, bool center_hit,
                                        uint32_t beta_word, Counters&) {
        // a loaded word at or below the payload is set already (the words only decrease)
        if (payload < cons[pix]) {
            const uint64_t old = atomicMin((unsigned long long*)(cons + pix), (unsigned long long)payload);
            if (emit && gb_empty(old)) append(filter, x, y, pix);
        }
        if (center_hit) {
            if (payload < center[pix]) {
                const uint64_t old = atomicMin((unsigned long long*)(center + pix), (unsigned long long)payload);
                if (emit && gb_empty(old)) append(shade, x, y, pix);
            }
            if (beta_word < beta[pix]) atomicMin(beta + pix, beta_word);
        }
    }
};

__global__ __launch_bounds__(256) void k_cam_long(BinArgs a) {
    if (a.meta[BIN_LONG] == 0u) return;
    const View& p = bin_view(a);
    const uint32_t n = a.meta[BIN_LONG_COUNT];
    LongCamSink sink{a.center,
                     a.cons,
                     a.beta,
                     PixelList{a.shade_list, a.meta + BIN_SHADE_SUBS, a.list_cap},
                     PixelList{a.filter_list, a.meta + BIN_FILTER_SUBS, a.list_cap},
                     a.emit_lists != 0u,
                     a.tile_shift,
                     a.tiles_x};
    Drawer<0, LongCamSink> d{p, a.scene_depth, a.scene_clear, sink, {0, 0, 0, 0}};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        walk_cam_record(d, a, a.long_list[i]);
}

hipError_t launch_cam_tiles(hipStream_t s, const BinArgs& a, int tile_blocks) {
    if (a.tile_shift != 4u && a.tile_shift != 5u) return hipErrorInvalidValue;
    // Split tiles take atomic minima into the G-buffer, which k_bin_fill
    // has cleared when a split exists.
    const uint32_t grid = a.chunk_capacity < (uint32_t)tile_blocks ? a.chunk_capacity : (uint32_t)tile_blocks;
    if (a.tile_shift == 4u) hipLaunchKernelGGL(k_cam_tiles<16>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_cam_tiles<32>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_cam_finish(hipStream_t s, const BinArgs& a, int long_blocks) {
    if (a.tile_shift == 4u) hipLaunchKernelGGL(k_cam_finish<16>, dim3(a.tiles), dim3(256), 0, s, a);
    else if (a.tile_shift == 5u) hipLaunchKernelGGL(k_cam_finish<32>, dim3(a.tiles), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cam_long, dim3(long_blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace hm
