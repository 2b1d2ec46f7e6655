// The offline baker: explicit strands -> the hair-mesh cages the rasterizer
// generates strands from (src/bake.rs:89-213). Host code; the result is the
// bundle / cage-corner / style arrays hm_groom_upload takes.
//
// Steps, each restating the reference function named beside it:
//   cluster roots in a 3D grid of ~180 strands per cell, big cells split
//   along their principal axis, sparse groups merged into a neighbour
//   (bake.rs:340-448);
//   split every group until its cage is coherent: every strand inside every
//   layer's quad, no styled quad sample more than 5 units from a strand, no
//   style texel whose strands disagree by more than 2, no member strand more
//   than 4 from its reconstruction (bake.rs:977-1088);
//   per group: 16 arc-length samples per strand, a PCA root frame and
//   parallel-transported layer frames with tight bounds (bake.rs:683-826),
//   corner tangents by differences of the quads, the mean residual from the
//   bilinear cage splatted into the 8 x 8 x 16 style volume (bake.rs:1090-1123);
//   ambient occlusion per corner from 8 hemisphere rays against the strands
//   (bake.rs:1139-1342), the corners spread over the host's threads.
// Floating point is f32 throughout, in the reference's operation order.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/hair_mi.h"

namespace {

constexpr size_t kLayers = HM_LAYER_COUNT;
constexpr size_t kStyleU = HM_STYLE_U, kStyleV = HM_STYLE_V, kTexels = HM_STYLE_TEXELS;
constexpr size_t kMinStrandsPerBundle = 8;
constexpr float kTargetStrandsPerCell = 180.0f;
constexpr float kMaxQuadGap = 5.0f;
constexpr float kMaxStyleStddev = 2.0f;
constexpr float kMaxMemberError = 4.0f;
constexpr size_t kAoRays = 8;
constexpr float kAoReach = 8.0f;
constexpr float kAoRadius = 0.45f;
constexpr float kAoMinT = 0.75f;

struct V3 {
    float x = 0, y = 0, z = 0;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3& operator+=(V3& a, V3 b) { return a = a + b; }
inline V3& operator-=(V3& a, V3 b) { return a = a - b; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float length_sq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(length_sq(a)); }
inline float distance_sq(V3 a, V3 b) { return length_sq(a - b); }
inline float distance(V3 a, V3 b) { return length(a - b); }
inline V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline V3 splat(float s) { return {s, s, s}; }
inline bool finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
// glam's normalize / normalize_or: v * (1 / |v|), the fallback when that
// reciprocal is not a finite positive number.
inline V3 normalize(V3 v) { return v * (1.0f / length(v)); }
inline V3 normalize_or(V3 v, V3 fallback) {
    const float r = 1.0f / length(v);
    return (std::isfinite(r) && r > 0.0f) ? v * r : fallback;
}
inline V3 lerp(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
const V3 kX = {1, 0, 0}, kY = {0, 1, 0}, kZ = {0, 0, 1};

struct V2 {
    float x = 0, y = 0;
};

struct M3 {  // columns
    V3 c[3];
    V3 operator*(V3 v) const { return c[0] * v.x + c[1] * v.y + c[2] * v.z; }
};

using Samples = std::array<V3, kLayers>;

struct Hair {
    const V3* points;
    size_t point_count;
    const uint32_t* strands;  // (start, count) pairs
    size_t strand_count;
    std::pair<uint32_t, uint32_t> strand(size_t i) const { return {strands[2 * i], strands[2 * i + 1]}; }
};

V3 average(const std::vector<V3>& points) {  // bake.rs:1344-1350
    V3 acc;
    for (const V3& p : points) acc += p;
    return acc / (float)std::max<size_t>(points.size(), 1);
}

V3 dominant_axis(const M3& cov, V3 seed) {  // bake.rs:1368-1378: power iteration
    V3 v = normalize_or(seed, kX);
    for (int i = 0; i < 24; ++i) {
        V3 next = cov * v;
        if (length_sq(next) < 1e-12f) break;
        v = normalize(next);
    }
    return v;
}

M3 deflate(const M3& cov, V3 axis) {  // bake.rs:1380-1383
    const float lambda = dot(axis, cov * axis);
    M3 out = cov;
    const V3 cols[3] = {axis * axis.x, axis * axis.y, axis * axis.z};
    for (int k = 0; k < 3; ++k) out.c[k] = cov.c[k] - cols[k] * lambda;
    return out;
}

// Mean and the two principal axes of a point set (bake.rs:1352-1366).
void plane_axes(const std::vector<V3>& points, V3& mean, V3& axis_u, V3& axis_v) {
    mean = average(points);
    M3 cov = {{splat(0), splat(0), splat(0)}};
    for (const V3& p : points) {
        const V3 d = p - mean;
        cov.c[0] += d * d.x;
        cov.c[1] += d * d.y;
        cov.c[2] += d * d.z;
    }
    axis_u = dominant_axis(cov, kX);
    axis_v = dominant_axis(deflate(cov, axis_u), kY);
    if (std::fabs(dot(axis_u, axis_v)) > 0.99f) axis_v = normalize_or(cross(axis_u, kY), kX);
    axis_v = normalize(axis_v);
}

struct Frame {  // bake.rs:625-681
    V3 mean, axis_u, axis_v;
    float min_u = 0, max_u = 0, min_v = 0, max_v = 0;
    float width_u() const { return max_u - min_u; }
    float width_v() const { return max_v - min_v; }
    std::array<V3, 4> quad() const {
        const V3 u0 = axis_u * min_u, u1 = axis_u * max_u, v0 = axis_v * min_v, v1 = axis_v * max_v;
        return {mean + u0 + v0, mean + u1 + v0, mean + u0 + v1, mean + u1 + v1};
    }
    V2 uv(V3 point) const {
        const V3 d = point - mean;
        const float lu = dot(d, axis_u), lv = dot(d, axis_v);
        const float u = max_u > min_u + 1e-8f ? (lu - min_u) / (max_u - min_u) : 0.5f;
        const float v = max_v > min_v + 1e-8f ? (lv - min_v) / (max_v - min_v) : 0.5f;
        return {u, v};
    }
    V3 longer_axis() const { return width_u() >= width_v() ? axis_u : axis_v; }
};

// Tight (u, v) bounds with a small pad and a 1e-2 minimum width (bake.rs:794-826).
void tight_bounds(const std::vector<V3>& points, V3 mean, V3 au, V3 av, Frame& f) {
    if (points.empty()) {
        f.min_u = -1e-3f, f.max_u = 1e-3f, f.min_v = -1e-3f, f.max_v = 1e-3f;
        return;
    }
    float min_u = 3.402823466e38f, max_u = -3.402823466e38f, min_v = 3.402823466e38f, max_v = -3.402823466e38f;
    for (const V3& p : points) {
        const V3 d = p - mean;
        min_u = std::min(min_u, dot(d, au));
        max_u = std::max(max_u, dot(d, au));
        min_v = std::min(min_v, dot(d, av));
        max_v = std::max(max_v, dot(d, av));
    }
    const float pad_u = std::max(max_u - min_u, 1e-3f) * 1e-3f + 1e-4f;
    const float pad_v = std::max(max_v - min_v, 1e-3f) * 1e-3f + 1e-4f;
    min_u -= pad_u, max_u += pad_u, min_v -= pad_v, max_v += pad_v;
    if (max_u - min_u < 1e-2f) {
        const float mid = (max_u + min_u) * 0.5f;
        min_u = mid - 5e-3f, max_u = mid + 5e-3f;
    }
    if (max_v - min_v < 1e-2f) {
        const float mid = (max_v + min_v) * 0.5f;
        min_v = mid - 5e-3f, max_v = mid + 5e-3f;
    }
    f.min_u = min_u, f.max_u = max_u, f.min_v = min_v, f.max_v = max_v;
}

std::vector<V3> layer_points(const std::vector<Samples>& strands, size_t layer) {
    std::vector<V3> pts;
    pts.reserve(strands.size());
    for (const Samples& s : strands) pts.push_back(s[layer]);
    return pts;
}

Frame root_frame(const std::vector<Samples>& strands) {  // bake.rs:716-732
    const std::vector<V3> pts = layer_points(strands, 0);
    Frame f;
    V3 au, av;
    plane_axes(pts, f.mean, au, av);
    au = normalize_or(au, kX);
    av = normalize_or(av - au * dot(av, au), normalize_or(cross(au, kY), kZ));
    f.axis_u = au, f.axis_v = av;
    tight_bounds(pts, f.mean, au, av, f);
    return f;
}

// The previous layer's axes carried into the plane normal to `normal` (bake.rs:768-792).
void transported_axes(V3 prev_u, V3 prev_v, V3 normal, V3& au, V3& av) {
    const V3 u_plane = prev_u - normal * dot(prev_u, normal);
    const V3 v_plane = prev_v - normal * dot(prev_v, normal);
    const float u2 = length_sq(u_plane), v2 = length_sq(v_plane);
    if (u2 >= v2 && u2 > 1e-8f) {
        au = normalize(u_plane);
        V3 v = cross(normal, au);
        if (dot(v, prev_v) < 0.0f) v = -v;
        av = normalize_or(v, kX);
    } else if (v2 > 1e-8f) {
        av = normalize(v_plane);
        V3 u = cross(av, normal);
        if (dot(u, prev_u) < 0.0f) u = -u;
        au = normalize_or(u, kX);
    } else {
        au = normalize_or(cross(normal, kY), kX);
        av = normalize_or(cross(normal, au), kZ);
    }
}

Frame oriented_frame(const std::vector<Samples>& strands, size_t layer, V3 ref_u, V3 ref_v) {  // bake.rs:734-755
    const std::vector<V3> pts = layer_points(strands, layer);
    Frame f;
    f.mean = average(pts);
    V3 tangent;
    for (const Samples& s : strands) {
        const V3 next = s[std::min(layer + 1, kLayers - 1)];
        const V3 prev = s[layer == 0 ? 0 : layer - 1];
        tangent += next - prev;
    }
    transported_axes(ref_u, ref_v, normalize_or(tangent, kY), f.axis_u, f.axis_v);
    tight_bounds(pts, f.mean, f.axis_u, f.axis_v, f);
    return f;
}

std::vector<Frame> frames_for(const std::vector<Samples>& strands) {  // bake.rs:757-766
    std::vector<Frame> frames;
    frames.reserve(kLayers);
    frames.push_back(root_frame(strands));
    for (size_t layer = 1; layer < kLayers; ++layer) {
        const Frame prev = frames[layer - 1];
        frames.push_back(oriented_frame(strands, layer, prev.axis_u, prev.axis_v));
    }
    return frames;
}

bool uv_in_unit(V2 uv) { return uv.x >= -1e-3f && uv.y >= -1e-3f && uv.x <= 1.0f + 1e-3f && uv.y <= 1.0f + 1e-3f; }
V2 clamp01(V2 uv) { return {std::min(std::max(uv.x, 0.0f), 1.0f), std::min(std::max(uv.y, 0.0f), 1.0f)}; }

bool samples_inside(const std::vector<Frame>& frames, const Samples& s) {  // bake.rs:828-833
    for (size_t l = 0; l < kLayers; ++l)
        if (!uv_in_unit(frames[l].uv(s[l]))) return false;
    return true;
}

// 16 arc-length samples per strand of at least two points and length >= 1e-5 (bake.rs:683-714).
std::vector<std::pair<size_t, Samples>> resample_indexed(const Hair& hair, const std::vector<size_t>& group) {
    std::vector<std::pair<size_t, Samples>> out;
    out.reserve(group.size());
    for (size_t strand : group) {
        const auto [start, count] = hair.strand(strand);
        if (count < 2) continue;
        const V3* pts = hair.points + start;
        std::vector<float> cum(count, 0.0f);
        for (uint32_t i = 1; i < count; ++i) cum[i] = cum[i - 1] + distance(pts[i], pts[i - 1]);
        const float len = cum[count - 1];
        if (len < 1e-5f) continue;
        Samples s;
        for (size_t layer = 0; layer < kLayers; ++layer) {
            const float target = len * (float)layer / (float)(kLayers - 1);
            uint32_t index = 1;
            while (index + 1 < count && cum[index] < target) ++index;
            const float span = std::max(cum[index] - cum[index - 1], 1e-6f);
            const float t = std::min(std::max((target - cum[index - 1]) / span, 0.0f), 1.0f);
            s[layer] = lerp(pts[index - 1], pts[index], t);
        }
        out.emplace_back(strand, s);
    }
    return out;
}

V3 bilinear(const std::array<V3, 4>& q, V2 uv) {  // bake.rs:1090-1097
    const float u = uv.x, v = uv.y;
    return q[0] * (1.0f - u) * (1.0f - v) + q[1] * u * (1.0f - v) + q[2] * (1.0f - u) * v + q[3] * u * v;
}

// The four style texels of (uv, layer) and their bilinear weights (bake.rs:1099-1123).
template <class F>
void for_splat(V2 uv, size_t layer, F f) {
    const float u = std::min(std::max(uv.x, 0.0f), 1.0f) * (float)(kStyleU - 1);
    const float v = std::min(std::max(uv.y, 0.0f), 1.0f) * (float)(kStyleV - 1);
    const uint32_t u0 = (uint32_t)std::floor(u), v0 = (uint32_t)std::floor(v);
    const uint32_t u1 = std::min<uint32_t>(u0 + 1, kStyleU - 1), v1 = std::min<uint32_t>(v0 + 1, kStyleV - 1);
    const float tu = u - (float)u0, tv = v - (float)v0;
    const uint32_t w = (uint32_t)layer;
    const std::tuple<uint32_t, uint32_t, float> corners[4] = {
        {u0, v0, (1.0f - tu) * (1.0f - tv)}, {u1, v0, tu * (1.0f - tv)}, {u0, v1, (1.0f - tu) * tv}, {u1, v1, tu * tv}};
    for (const auto& [iu, iv, alpha] : corners) {
        if (alpha <= 0.0f) continue;
        f((size_t)((w * kStyleV + iv) * kStyleU + iu), alpha);
    }
}

struct StyleFit {  // bake.rs:839-889
    std::vector<V3> sum;
    std::vector<float> sum_sq, weight;
};

StyleFit accumulate_style(const std::vector<Samples>& strands, const std::vector<Frame>& frames) {
    StyleFit fit{std::vector<V3>(kTexels), std::vector<float>(kTexels, 0.0f), std::vector<float>(kTexels, 0.0f)};
    std::vector<std::array<V3, 4>> quads(kLayers);
    for (size_t l = 0; l < kLayers; ++l) quads[l] = frames[l].quad();
    for (const Samples& s : strands) {
        V2 uv = frames[0].uv(s[0]);
        if (!uv_in_unit(uv)) continue;
        uv = clamp01(uv);
        for (size_t layer = 0; layer < kLayers; ++layer) {
            const V3 residual = s[layer] - bilinear(quads[layer], uv);
            const float len_sq = length_sq(residual);
            for_splat(uv, layer, [&](size_t i, float alpha) {
                fit.sum[i] += residual * alpha;
                fit.weight[i] += alpha;
                fit.sum_sq[i] += len_sq * alpha;
            });
        }
    }
    return fit;
}

std::vector<std::array<float, 4>> style_means(const StyleFit& fit) {  // bake.rs:891-904
    std::vector<std::array<float, 4>> style(kTexels, {0.0f, 0.0f, 0.0f, 0.0f});
    for (size_t i = 0; i < kTexels; ++i)
        if (fit.weight[i] > 0.0f) {
            const V3 r = fit.sum[i] / fit.weight[i];
            style[i] = {r.x, r.y, r.z, 0.0f};
        }
    return style;
}

float worst_stddev(const StyleFit& fit, size_t& layer) {  // bake.rs:906-923
    float worst = 0.0f;
    layer = 0;
    for (size_t i = 0; i < kTexels; ++i) {
        const float w = fit.weight[i];
        if (w <= 1.0f) continue;
        const V3 mean = fit.sum[i] / w;
        const float second = fit.sum_sq[i] / w;
        const float sd = std::sqrt(std::max(second - length_sq(mean), 0.0f));
        if (sd > worst) {
            worst = sd;
            layer = i / (kStyleU * kStyleV);
        }
    }
    return worst;
}

V3 style_slice(const std::vector<std::array<float, 4>>& style, V2 uv, size_t layer) {  // bake.rs:925-943
    const float u = std::min(std::max(uv.x, 0.0f), 1.0f) * (float)(kStyleU - 1);
    const float v = std::min(std::max(uv.y, 0.0f), 1.0f) * (float)(kStyleV - 1);
    const size_t u0 = (size_t)std::floor(u), v0 = (size_t)std::floor(v);
    const size_t u1 = std::min(u0 + 1, kStyleU - 1), v1 = std::min(v0 + 1, kStyleV - 1);
    const float tu = u - (float)u0, tv = v - (float)v0;
    auto at = [&](size_t iu, size_t iv) {
        const auto& t = style[(layer * kStyleV + iv) * kStyleU + iu];
        return V3{t[0], t[1], t[2]};
    };
    return lerp(lerp(at(u0, v0), at(u1, v0), tu), lerp(at(u0, v1), at(u1, v1), tu), tv);
}

// Largest distance from a styled 5 x 5 quad sample to the nearest strand sample of the layer (bake.rs:945-975).
float worst_quad_gap(const std::vector<Samples>& strands, const std::vector<Frame>& frames,
                     const std::vector<std::array<float, 4>>& style, size_t& worst_layer, V3& worst_axis) {
    worst_layer = 0;
    worst_axis = frames[0].longer_axis();
    float worst = 0.0f;
    const float steps[5] = {0.0f, 0.25f, 0.5f, 0.75f, 1.0f};
    for (size_t layer = 0; layer < frames.size(); ++layer) {
        const auto quad = frames[layer].quad();
        float gap = 0.0f;
        for (float v : steps)
            for (float u : steps) {
                const V2 uv = {u, v};
                const V3 pos = bilinear(quad, uv) + style_slice(style, uv, layer);
                float nearest = 3.402823466e38f;
                for (const Samples& s : strands) nearest = std::min(nearest, distance_sq(pos, s[layer]));
                gap = std::max(gap, nearest);
            }
        gap = std::sqrt(gap);
        if (gap > worst) {
            worst = gap;
            worst_layer = layer;
            worst_axis = frames[layer].longer_axis();
        }
    }
    return worst;
}

float max_member_error(const std::vector<Samples>& strands, const std::vector<Frame>& frames,
                       const std::vector<std::array<float, 4>>& style, size_t& worst_layer, V3& axis) {
    // bake.rs:1063-1088
    float worst = 0.0f;
    worst_layer = 0;
    axis = frames[0].longer_axis();
    for (const Samples& s : strands) {
        V2 uv = frames[0].uv(s[0]);
        if (!uv_in_unit(uv)) continue;
        uv = clamp01(uv);
        for (size_t i = 0; i < kLayers; ++i) {
            const V3 predicted = bilinear(frames[i].quad(), uv) + style_slice(style, uv, i);
            const float err = distance(predicted, s[i]);
            if (err > worst) {
                worst = err;
                worst_layer = i;
                axis = frames[i].longer_axis();
            }
        }
    }
    return worst;
}

// Sort (key, strand) by key, stably, as f32::total_cmp orders finite keys.
void sort_by_key(std::vector<std::pair<float, size_t>>& order) {
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<float, size_t>& a, const std::pair<float, size_t>& b) { return a.first < b.first; });
}

void split_rec(const Hair& hair, const std::vector<size_t>& group, uint32_t depth, std::vector<std::vector<size_t>>& out) {
    // bake.rs:977-1061
    if (group.size() < 2 || depth >= 16) {
        out.push_back(group);
        return;
    }
    const auto indexed = resample_indexed(hair, group);
    if (indexed.size() < 2) {
        std::vector<size_t> ids;
        for (const auto& e : indexed) ids.push_back(e.first);
        out.push_back(ids);
        return;
    }
    std::vector<Samples> resampled;
    for (const auto& e : indexed) resampled.push_back(e.second);
    const std::vector<Frame> frames = frames_for(resampled);
    std::vector<size_t> inside, outside;
    for (const auto& e : indexed) (samples_inside(frames, e.second) ? inside : outside).push_back(e.first);
    if (!outside.empty() && !inside.empty()) {
        split_rec(hair, inside, depth + 1, out);
        split_rec(hair, outside, depth + 1, out);
        return;
    }
    const StyleFit fit = accumulate_style(resampled, frames);
    const auto style = style_means(fit);
    size_t gap_layer, var_layer, member_layer;
    V3 gap_axis, member_axis;
    const float gap = worst_quad_gap(resampled, frames, style, gap_layer, gap_axis);
    const float stddev = worst_stddev(fit, var_layer);
    const float member_error = max_member_error(resampled, frames, style, member_layer, member_axis);
    const bool gap_bad = gap > kMaxQuadGap, var_bad = stddev > kMaxStyleStddev, member_bad = member_error > kMaxMemberError;
    if (!gap_bad && !var_bad && !member_bad) {
        std::vector<size_t> ids;
        for (const auto& e : indexed) ids.push_back(e.first);
        out.push_back(ids);
        return;
    }
    size_t layer;
    V3 axis;
    if (member_bad && member_error / kMaxMemberError >= gap / kMaxQuadGap &&
        member_error / kMaxMemberError >= stddev / std::max(kMaxStyleStddev, 1e-3f)) {
        layer = member_layer, axis = member_axis;
    } else if (var_bad && (!gap_bad || stddev / kMaxStyleStddev > gap / kMaxQuadGap)) {
        layer = std::min(var_layer, frames.size() - 1);
        axis = frames[layer].longer_axis();
    } else {
        layer = gap_layer, axis = gap_axis;
    }
    const V3 mean = average(layer_points(resampled, layer));
    std::vector<std::pair<float, size_t>> order;
    for (const auto& e : indexed) order.emplace_back(dot(e.second[layer] - mean, axis), e.first);
    sort_by_key(order);
    const float span = order.back().first - order.front().first;
    const size_t mid = order.size() / 2;
    if (mid == 0) {
        out.push_back(group);
        return;
    }
    std::vector<size_t> left, right;
    if (span < 1e-3f) {  // strands that share a uv but disagree: each half its own bundle
        left.assign(group.begin(), group.begin() + (ptrdiff_t)mid);
        right.assign(group.begin() + (ptrdiff_t)mid, group.end());
    } else {
        for (size_t i = 0; i < order.size(); ++i) (i < mid ? left : right).push_back(order[i].second);
    }
    split_rec(hair, left, depth + 1, out);
    split_rec(hair, right, depth + 1, out);
}

std::vector<std::vector<size_t>> split_along_axis(const std::vector<V3>& roots, const std::vector<size_t>& group) {
    // bake.rs:381-397
    std::vector<V3> pts;
    for (size_t i : group) pts.push_back(roots[i]);
    V3 mean, axis, unused;
    plane_axes(pts, mean, axis, unused);
    std::vector<std::pair<float, size_t>> order;
    for (size_t i : group) order.emplace_back(dot(roots[i] - mean, axis), i);
    sort_by_key(order);
    if (order.back().first - order.front().first < 1e-3f) return {group};
    const size_t chunk = (size_t)kTargetStrandsPerCell;
    std::vector<std::vector<size_t>> out;
    for (size_t i = 0; i < order.size(); i += chunk) {
        std::vector<size_t> piece;
        for (size_t k = i; k < std::min(order.size(), i + chunk); ++k) piece.push_back(order[k].second);
        out.push_back(piece);
    }
    return out;
}

int32_t f2i(float x) {  // Rust's saturating f32 -> i32 cast
    if (std::isnan(x)) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

std::vector<std::vector<size_t>> cluster_roots_3d(const std::vector<V3>& roots, const std::vector<size_t>& which) {
    // bake.rs:340-379
    if (which.empty()) return {};
    if (which.size() <= (size_t)kTargetStrandsPerCell) return {which};
    V3 lo = splat(3.402823466e38f), hi = splat(-3.402823466e38f);
    for (size_t i : which) lo = vmin(lo, roots[i]), hi = vmax(hi, roots[i]);
    const V3 extent = vmax(hi - lo, splat(1e-3f));
    const int32_t cells = (int32_t)std::min(std::max(std::ceil(std::cbrt((float)which.size() / kTargetStrandsPerCell)),
                                                     2.0f), 32.0f);
    const V3 cell = extent / (float)cells;
    std::map<std::tuple<int32_t, int32_t, int32_t>, std::vector<size_t>> buckets;  // ordered by key
    for (size_t i : which) {
        const V3 p = (roots[i] - lo) / cell;
        const auto key = std::make_tuple(std::min(std::max(f2i(p.x), 0), cells - 1),
                                         std::min(std::max(f2i(p.y), 0), cells - 1),
                                         std::min(std::max(f2i(p.z), 0), cells - 1));
        buckets[key].push_back(i);
    }
    std::vector<std::vector<size_t>> groups;
    for (auto& [key, group] : buckets) {
        if (group.size() > (size_t)kTargetStrandsPerCell * 2) {
            for (auto& g : split_along_axis(roots, group)) groups.push_back(g);
        } else {
            groups.push_back(group);
        }
    }
    return groups;
}

V3 centroid(const std::vector<size_t>& group, const std::vector<V3>& roots) {  // bake.rs:442-448
    V3 acc;
    for (size_t i : group) acc += roots[i];
    return acc / (float)std::max<size_t>(group.size(), 1);
}

std::vector<std::vector<size_t>> absorb_sparse(std::vector<std::vector<size_t>> groups, const std::vector<V3>& roots) {
    // bake.rs:399-440: groups below 8 strands join the nearest group within 5 units
    groups.erase(std::remove_if(groups.begin(), groups.end(), [](const std::vector<size_t>& g) { return g.empty(); }),
                 groups.end());
    if (groups.size() <= 1) return groups;
    std::vector<V3> cents;
    for (const auto& g : groups) cents.push_back(centroid(g, roots));
    std::vector<bool> alive(groups.size(), true);
    const float merge_sq = kMaxQuadGap * kMaxQuadGap;
    for (size_t i = 0; i < groups.size(); ++i) {
        if (!alive[i] || groups[i].size() >= kMinStrandsPerBundle) continue;
        size_t best = SIZE_MAX;
        float best_d = 3.402823466e38f;
        for (size_t j = 0; j < groups.size(); ++j) {
            if (i == j || !alive[j]) continue;
            const float d = distance_sq(cents[i], cents[j]);
            if (d < best_d) best_d = d, best = j;
        }
        if (best == SIZE_MAX || best_d > merge_sq) continue;
        groups[best].insert(groups[best].end(), groups[i].begin(), groups[i].end());
        groups[i].clear();
        cents[best] = centroid(groups[best], roots);
        alive[i] = false;
    }
    std::vector<std::vector<size_t>> out;
    for (size_t i = 0; i < groups.size(); ++i)
        if (alive[i] && !groups[i].empty()) out.push_back(std::move(groups[i]));
    return out;
}

// ---- ambient occlusion (bake.rs:1125-1342) ------------------------------------

struct OccluderGrid {
    V3 origin;
    float cell = 4.0f;
    std::vector<std::pair<V3, V3>> segments;
    std::map<std::tuple<int32_t, int32_t, int32_t>, std::vector<uint32_t>> buckets;

    std::tuple<int32_t, int32_t, int32_t> key_of(V3 p) const {
        const V3 q = (p - origin) / cell;
        return {f2i(std::floor(q.x)), f2i(std::floor(q.y)), f2i(std::floor(q.z))};
    }

    explicit OccluderGrid(const Hair& hair) {
        V3 lo = splat(3.402823466e38f), hi = splat(-3.402823466e38f);
        for (size_t s = 0; s < hair.strand_count; ++s) {
            const auto [start, count] = hair.strand(s);
            if (count < 2) continue;
            const V3* pts = hair.points + start;
            // about 16 coarse segments per strand
            const size_t step = std::max<size_t>(1, (size_t)std::max(1.0f, std::ceil((float)(count - 1) / 16.0f)));
            size_t index = 0;
            while (index + 1 < count) {
                const size_t next = std::min<size_t>(index + step, count - 1);
                if (next > index) {
                    const V3 a = pts[index], b = pts[next];
                    lo = vmin(vmin(lo, a), b);
                    hi = vmax(vmax(hi, a), b);
                    segments.emplace_back(a, b);
                }
                if (next + 1 >= count) break;
                index = next;
            }
        }
        if (!finite(lo)) lo = splat(0.0f);
        origin = lo;
        for (size_t i = 0; i < segments.size(); ++i) {
            const auto& [a, b] = segments[i];
            for (V3 p : {a, b, (a + b) * 0.5f}) buckets[key_of(p)].push_back((uint32_t)i);
        }
    }

    bool occluded(V3 o, V3 dir) const;
};

bool ray_segment(V3 o, V3 dir, V3 a, V3 b) {  // bake.rs:1294-1319
    const V3 ab = b - a, v = o - a;
    const float d2 = dot(dir, ab), d3 = std::max(length_sq(ab), 1e-12f), d4 = dot(v, dir), d5 = dot(v, ab);
    const float denom = d3 - d2 * d2;
    float s = std::fabs(denom) < 1e-8f ? 0.0f : std::min(std::max((d5 - d4 * d2) / denom, 0.0f), 1.0f);
    float t = dot(a + ab * s - o, dir);
    if (!(t >= kAoMinT && t <= kAoReach)) {
        t = std::min(std::max(t, kAoMinT), kAoReach);
        s = std::min(std::max(dot(o + dir * t - a, ab) / d3, 0.0f), 1.0f);
        t = dot(a + ab * s - o, dir);
        if (!(t >= kAoMinT && t <= kAoReach)) return false;
    }
    return distance_sq(o + dir * t, a + ab * s) <= kAoRadius * kAoRadius;
}

bool OccluderGrid::occluded(V3 o, V3 dir) const {  // bake.rs:1257-1281
    const int32_t steps = std::min(std::max(f2i(std::ceil(kAoReach / (cell * 0.5f))), 1), 8);
    const float dt = kAoReach / (float)steps;
    for (int32_t step = 0; step <= steps; ++step) {
        const auto it = buckets.find(key_of(o + dir * ((float)step * dt)));
        if (it == buckets.end()) continue;
        for (uint32_t i : it->second)
            if (ray_segment(o, dir, segments[i].first, segments[i].second)) return true;
    }
    return false;
}

V3 hemisphere_dir(size_t index, size_t count, V3 normal) {  // bake.rs:1284-1292
    const float z = std::sqrt(((float)index + 0.5f) / (float)count);
    const float phi = (float)index * 2.3999631f;
    const float radius = std::sqrt(std::max(1.0f - z * z, 0.0f));
    const V3 local = {radius * std::cos(phi), radius * std::sin(phi), z};
    const V3 tangent = normalize_or(cross(normal, kY), kX);
    const V3 bitangent = normalize_or(cross(normal, tangent), kZ);
    return normalize_or(tangent * local.x + bitangent * local.y + normal * local.z, normal);
}

struct Corner {
    V3 position, tangent;
    float ao = 1.0f;
};

void bake_ambient_occlusion(std::vector<Corner>& corners, const std::vector<hm_bundle>& bundles, const Hair& hair) {
    if (corners.empty()) return;
    const OccluderGrid grid(hair);
    std::vector<float> ao(corners.size(), 1.0f);
    auto work = [&](size_t b0, size_t b1) {
        for (size_t b = b0; b < b1; ++b) {
            const size_t base = bundles[b].layer_offset;
            for (size_t layer = 0; layer < kLayers; ++layer) {
                V3 center;
                for (size_t c = 0; c < 4; ++c) center += corners[base + layer * 4 + c].position;
                center = center * 0.25f;
                for (size_t c = 0; c < 4; ++c) {
                    const size_t idx = base + layer * 4 + c;
                    const V3 pos = corners[idx].position;
                    const V3 tan = normalize_or(corners[idx].tangent, kY);
                    V3 outward = pos - center;
                    outward -= tan * dot(outward, tan);
                    outward = normalize_or(outward, normalize_or(cross(tan, kY), kX));
                    float occluded = 0.0f;
                    for (size_t ray = 0; ray < kAoRays; ++ray)
                        if (grid.occluded(pos, hemisphere_dir(ray, kAoRays, outward))) occluded += 1.0f;
                    ao[idx] = std::min(std::max(1.0f - occluded / (float)kAoRays, 0.0f), 1.0f);
                }
            }
        }
    };
    const size_t nthreads = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 16));
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nthreads; ++t)
        pool.emplace_back(work, bundles.size() * t / nthreads, bundles.size() * (t + 1) / nthreads);
    for (auto& th : pool) th.join();
    for (size_t i = 0; i < corners.size(); ++i) corners[i].ao = ao[i];
}

}  // namespace

struct hm_bake_result {
    std::vector<hm_bundle> bundles;
    std::vector<hm_layer> layers;
    std::vector<float> style;  // 4 per texel
    float bounds_min[3] = {0, 0, 0}, bounds_max[3] = {0, 0, 0};
};

extern "C" {

hm_status hm_bake(const float* points, uint32_t point_count, const uint32_t* strands, uint32_t strand_count,
                  hm_bake_result** out) {
    if (!out || (point_count && !points) || (strand_count && !strands)) return HM_ERR_INVALID;
    *out = nullptr;
    for (uint32_t s = 0; s < strand_count; ++s)
        if ((uint64_t)strands[2 * s] + strands[2 * s + 1] > point_count) return HM_ERR_INVALID;
    hm_bake_result* r = new hm_bake_result();
    const Hair hair{reinterpret_cast<const V3*>(points), point_count, strands, strand_count};
    if (strand_count == 0) {  // bake.rs:95-97: the empty mesh
        *out = r;
        return HM_OK;
    }
    // bake.rs:99-113
    std::vector<V3> roots(strand_count);
    for (uint32_t s = 0; s < strand_count; ++s) roots[s] = hair.points[strands[2 * s]];
    std::vector<size_t> all(strand_count);
    for (size_t i = 0; i < all.size(); ++i) all[i] = i;
    std::vector<std::vector<size_t>> coherent;
    for (const auto& group : absorb_sparse(cluster_roots_3d(roots, all), roots)) {
        if (group.empty()) continue;
        split_rec(hair, group, 0, coherent);
    }
    // bake.rs:115-204
    std::vector<Corner> corners;
    std::vector<std::array<float, 4>> style;
    V3 bmin = splat(3.402823466e38f), bmax = splat(-3.402823466e38f);
    for (const auto& group : coherent) {
        const auto indexed = resample_indexed(hair, group);
        if (indexed.empty()) continue;
        std::vector<Samples> resampled;
        for (const auto& e : indexed) resampled.push_back(e.second);
        const std::vector<Frame> frames = frames_for(resampled);
        const Frame root = frames[0];
        std::vector<const Samples*> kept;
        for (const auto& e : indexed)
            if (samples_inside(frames, e.second)) kept.push_back(&e.second);
        if (kept.empty()) continue;
        std::vector<std::array<V3, 4>> quads;
        for (const Frame& f : frames) {
            const auto q = f.quad();
            for (const V3& c : q) bmin = vmin(bmin, c), bmax = vmax(bmax, c);
            quads.push_back(q);
        }
        const uint32_t layer_offset = (uint32_t)corners.size(), style_offset = (uint32_t)style.size();
        std::vector<V3> sum(kTexels);
        std::vector<float> weight(kTexels, 0.0f);
        for (const Samples* s : kept) {
            V2 uv = root.uv((*s)[0]);
            if (!uv_in_unit(uv)) continue;
            uv = clamp01(uv);
            for (size_t layer = 0; layer < kLayers; ++layer) {
                const V3 residual = (*s)[layer] - bilinear(quads[layer], uv);
                for_splat(uv, layer, [&](size_t i, float alpha) {
                    sum[i] += residual * alpha;
                    weight[i] += alpha;
                });
            }
        }
        for (size_t layer = 0; layer < kLayers; ++layer) {
            const auto& q = quads[layer];
            const auto& prev = quads[layer == 0 ? 0 : layer - 1];
            const auto& next = quads[std::min(layer + 1, kLayers - 1)];
            for (size_t c = 0; c < 4; ++c) {
                V3 t;
                if (layer == 0) t = next[c] - q[c];
                else if (layer + 1 == kLayers) t = q[c] - prev[c];
                else t = (next[c] - prev[c]) * 0.5f;
                corners.push_back({q[c], t, 1.0f});
            }
        }
        for (size_t i = 0; i < kTexels; ++i) {
            const V3 res = weight[i] > 0.0f ? sum[i] / weight[i] : V3{};
            style.push_back({res.x, res.y, res.z, 0.0f});
        }
        r->bundles.push_back({layer_offset, style_offset, (uint32_t)kept.size(), 0u});
    }
    if (r->bundles.empty()) {
        *out = r;
        return HM_OK;
    }
    bake_ambient_occlusion(corners, r->bundles, hair);
    r->layers.resize(corners.size());
    for (size_t i = 0; i < corners.size(); ++i) {
        const Corner& c = corners[i];
        r->layers[i] = {{c.position.x, c.position.y, c.position.z, c.ao}, {c.tangent.x, c.tangent.y, c.tangent.z, 0.0f}};
    }
    r->style.resize(style.size() * 4);
    for (size_t i = 0; i < style.size(); ++i) std::memcpy(&r->style[i * 4], style[i].data(), 16);
    const float lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
    std::memcpy(r->bounds_min, lo, sizeof(lo));
    std::memcpy(r->bounds_max, hi, sizeof(hi));
    *out = r;
    return HM_OK;
}

hm_status hm_bake_get(const hm_bake_result* r, hm_groom_desc* desc) {
    if (!r || !desc) return HM_ERR_INVALID;
    desc->bundles = r->bundles.data();
    desc->bundle_count = (uint32_t)r->bundles.size();
    desc->layers = r->layers.data();
    desc->layer_count = (uint32_t)r->layers.size();
    desc->style = r->style.data();
    desc->style_count = (uint32_t)(r->style.size() / 4);
    std::memcpy(desc->bounds_min, r->bounds_min, sizeof(r->bounds_min));
    std::memcpy(desc->bounds_max, r->bounds_max, sizeof(r->bounds_max));
    return HM_OK;
}

void hm_bake_free(hm_bake_result* r) { delete r; }

}  // extern "C"
