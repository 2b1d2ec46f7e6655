// Shading functions of shade.wesl (paths relative to the reference repository):
//   gaussian3 shade.wesl:51-54, fresnel :56-60, absorption_from_color :70-77,
//   azimuthal :94-116, lobe_shape :124-128, dual_scatter :153-190,
//   hair_bsdf :193-244, dom_optical_depth/texel/visibility :246-310,
//   spot_mask :312-327, quat_rotate/cubemap_dir :329-339, probe_radiance
//   :396-427 (sample_environment :365-382), rect lights :503-518; Bevy's
//   getRangeFalloff / getDistanceAttenuation restated.
// Shaded RGB is held to a relative tolerance, so the continuous terms below
// use the hardware approximations (v_exp_f32, v_log_f32, v_sqrt_f32,
// v_rcp_f32, v_cos_f32; about 1 ulp) rather than the correctly rounded
// sequences the library is compiled with. The G-buffer decode and every
// discrete choice (texel, slice) keep the exact arithmetic of hm_device.h.
#pragma once

#include "hm_device.h"
#include "hm_kernels.h"

namespace hm {

__device__ __forceinline__ float fexp(float x) { return __expf(x); }
__device__ __forceinline__ float flog(float x) { return __logf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fcos(float x) { return __cosf(x); }
__device__ __forceinline__ float fpow(float b, float e) { return __expf(e * __logf(b)); }
// The bare base-2 instructions (v_exp_f32, v_log_f32) for the BSDF and DOM
// lookup: no range scaling for results below 2^-126, which cannot change a
// colour held to a relative tolerance (the library forms above add a
// compare and two selects per call).
constexpr float kLog2e = 1.44269504088896341f;
constexpr float kGaussLog2 = 0.849321800288019f;    // sqrt(log2(e) / 2): exp(-z^2 / 2) = exp2(-(z kGaussLog2)^2)
constexpr float kInvSqrt2Pi = 0.398942280401432678f;
__device__ __forceinline__ float hexp(float x) { return __builtin_amdgcn_exp2f(x * kLog2e); }
__device__ __forceinline__ float hlog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ float hpow(float b, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(b)); }
// Abramowitz & Stegun 4.4.46 (Hastings): asin x = pi/2 - sqrt(1 - x) P7(x)
// on [0, 1], |error| <= 2e-8; odd extension for x < 0. The polynomial's
// multiply-adds fuse: an error of this size in the result is not amplified
// (the conditioning near |x| = 1 concerns x itself, which is exact here).
#pragma clang fp contract(fast)
__device__ __forceinline__ float fasin(float x) {
    float a = fabsf(x);
    float p = -0.0012624911f;
    p = p * a + 0.0066700901f;
    p = p * a - 0.0170881256f;
    p = p * a + 0.0308918810f;
    p = p * a - 0.0501743046f;
    p = p * a + 0.0889789874f;
    p = p * a - 0.2145988016f;
    p = p * a + 1.5707963050f;
    float r = 1.5707963268f - fsqrt(1.0f - a) * p;
    return x < 0.0f ? -r : r;
}
#pragma clang fp contract(off)
__device__ __forceinline__ F3 fnormalize(F3 v) {  // safe_normalize with the hardware reciprocal square root
    float l2 = dot3(v, v);
    if (l2 < 1e-10f) return {0.0f, 1.0f, 0.0f};
    float r = __builtin_amdgcn_rsqf(l2);
    return {v.x * r, v.y * r, v.z * r};
}
__device__ __forceinline__ F3 fdiv3(F3 a, F3 b) { return {fdiv(a.x, b.x), fdiv(a.y, b.y), fdiv(a.z, b.z)}; }
// clamp / max in one instruction each (v_med3_f32, v_max_f32): the same
// values as clampf / mx for every non-NaN input (a zero keeps its sign).
__device__ __forceinline__ float sclamp(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
__device__ __forceinline__ float smax(float a, float b) { return __builtin_fmaxf(a, b); }

struct Lobes { F3 r, tt, trt; };
// inv_beta: the reciprocal gaussian3 takes of the widths, formed once per pixel.
struct Shape { F3 alpha, beta, inv_beta; };
struct DualScatter { Lobes fwd; F3 a_f, beta_f2, a_back, delta_b, sigma_b2; };

// The BSDF, light terms and dual scattering below are continuous in every
// input: multiply-adds may fuse there (one rounding instead of two).
#pragma clang fp contract(fast)

__device__ __forceinline__ F3 vexp(F3 a) { return {hexp(a.x), hexp(a.y), hexp(a.z)}; }
__device__ __forceinline__ F3 vsqrt(F3 a) { return {fsqrt(a.x), fsqrt(a.y), fsqrt(a.z)}; }

// 1 / max(b, 1e-3) per channel: the reciprocal width of gaussian3.
__device__ __forceinline__ F3 gaussian_rcp(F3 b) {
    F3 bb = vmax(b, splat3(1e-3f));
    return {frcp(bb.x), frcp(bb.y), frcp(bb.z)};
}

// gaussian3 from the reciprocal width r = gaussian_rcp(b).
__device__ __forceinline__ F3 gaussian3_r(F3 r, F3 x) {
    const float s2pi = 2.5066282746f;  // sqrt(2 pi)
    F3 z = x * r;
    return vexp(-0.5f * z * z) * r * (1.0f / s2pi);
}

__device__ __forceinline__ F3 gaussian3(F3 b, F3 x) { return gaussian3_r(gaussian_rcp(b), x); }

__device__ __forceinline__ float fresnel(float cos_t) {
    const float ior = 1.55f;
    float r0 = (1.0f - ior) / (1.0f + ior);
    float f0 = r0 * r0;
    float x = 1.0f - sclamp(cos_t, 0.0f, 1.0f);
    float x2 = x * x;
    return f0 + (1.0f - f0) * (x2 * x2 * x);
}

__device__ __forceinline__ F3 absorption_from_color(F3 color, float roughness) {
    float b = clampf(roughness, 0.0f, 1.0f);
    float b2 = b * b, b3 = b2 * b;
    float d = 5.969f - 0.215f * b + 2.532f * b2 - 10.73f * b3 + 5.574f * b2 * b2 + 0.245f * b3 * b2;
    F3 c = vclamp(color, 1e-4f, 1.0f);
    float rd = frcp(d);
    F3 l = {hlog(c.x) * rd, hlog(c.y) * rd, hlog(c.z) * rd};
    return l * l;
}

// The parts of azimuthal that depend on theta_d only (one light or one
// pixel's quadrature): 1 / cos theta_d, Karis's a = 1 / eta, and the TRT
// lobe's Fresnel factor (1 - f)^2 f and transmittance exp(-3.2 sigma_a / cos theta_d).
struct AzimuthalD {
    float inv_cd, a, trt_f;
    F3 t_trt;
};
__device__ __forceinline__ AzimuthalD azimuthal_d(float cos_theta_d, F3 sigma_a) {
    AzimuthalD d;
    d.inv_cd = frcp(cos_theta_d);
    d.a = frcp(1.19f * d.inv_cd + 0.36f * cos_theta_d);
    const float f_trt = fresnel(cos_theta_d * 0.5f);
    d.trt_f = (1.0f - f_trt) * (1.0f - f_trt) * f_trt;
    d.t_trt = vexp(sigma_a * (-3.2f * d.inv_cd));
    return d;
}

// N_R and N_TT of azimuthal (shade.wesl:94-116) at one angle: cos(phi),
// cos(phi / 2) = sqrt(0.5 + 0.5 cos phi), Karis's offset term 0.6 - 0.8 cos phi
// and the TT weight exp(-3.65 cos phi - 3.98).
__device__ __forceinline__ void azimuthal_r_tt(float cos_theta_d, const AzimuthalD& d, F3 sigma_a, float cos_half_phi,
                                               float h_k, float w_tt, float& n_r, F3& n_tt) {
    n_r = 0.25f * cos_half_phi * fresnel(cos_theta_d * cos_half_phi);
    const float h = cos_half_phi * (1.0f + d.a * h_k);
    const float f_tt = fresnel(cos_theta_d * fsqrt(sclamp(1.0f - h * h, 0.0f, 1.0f)));
    const float s_tt = fsqrt(sclamp(1.0f - h * h * d.a * d.a, 0.0f, 1.0f));
    n_tt = vexp(sigma_a * (-2.0f * s_tt * d.inv_cd)) * (w_tt * (1.0f - f_tt) * (1.0f - f_tt));
}

__device__ __forceinline__ F3 lobe_mean(const Lobes& l, F3 total, F3 q) {
    return fdiv3(l.r * q.x + l.tt * q.y + l.trt * q.z, vmax(total, splat3(1e-5f)));
}

// The 16 azimuthal angles of dual_scatter's quadrature (shade.wesl:158-162):
// sample i < 8 backward, phi = (i + 0.5) pi / 16; i >= 8 forward,
// phi = pi / 2 + (i - 7.5) pi / 16. Per sample, rounded to f32 from double
// precision: cos phi, cos(phi / 2) = sqrt(0.5 + 0.5 cos phi), 0.6 - 0.8 cos phi,
// exp(-3.65 cos phi - 3.98) (TT) and exp(17 cos phi - 16.78) (TRT).
struct DualAngle { float cos_phi, cos_half_phi, h_k, w_tt, w_trt; };
constexpr DualAngle kDualAngles[16] = {
    {0.9951847195625305f, 0.9987954497337341f, -0.19614776968955994f, 0.0004942721570841968f, 1.1481364965438843f},
    {0.9569403529167175f, 0.9891765117645264f, -0.1655522882938385f, 0.0005683163763023913f, 0.5992873907089233f},
    {0.8819212317466736f, 0.9700312614440918f, -0.10553698241710663f, 0.000747323443647474f, 0.16740503907203674f},
    {0.7730104327201843f, 0.9415440559387207f, -0.01840834692120552f, 0.0011121189454570413f, 0.026283269748091698f},
    {0.6343932747840881f, 0.903989315032959f, 0.09248538315296173f, 0.0018445213790982962f, 0.002490394050255418f},
    {0.47139665484428406f, 0.8577286005020142f, 0.22288267314434052f, 0.003343965858221054f, 0.0001559060619911179f},
    {0.2902846336364746f, 0.803207516670227f, 0.36777228116989136f, 0.006476733833551407f, 7.17318198439898e-06f},
    {0.0980171337723732f, 0.7409511208534241f, 0.5215862989425659f, 0.013065729290246964f, 2.7302328931000375e-07f},
    {-0.09801722317934036f, 0.6715589165687561f, 0.6784137487411499f, 0.026722826063632965f, 9.747131990422986e-09f},
    {-0.2902847230434418f, 0.5956993103027344f, 0.8322277665138245f, 0.05390884354710579f, 3.709921259087423e-10f},
    {-0.4713968336582184f, 0.5141026973724365f, 0.9771174788475037f, 0.10441295057535172f, 1.706918807686808e-11f},
    {-0.6343932747840881f, 0.4275550842285156f, 1.1075146198272705f, 0.18929198384284973f, 1.0685851075525887e-12f},
    {-0.7730104923248291f, 0.3368898332118988f, 1.2184083461761475f, 0.31395310163497925f, 1.0125053735970252e-13f},
    {-0.8819213509559631f, 0.24298009276390076f, 1.3055371046066284f, 0.4672050476074219f, 1.5896729801110655e-14f},
    {-0.9569403529167175f, 0.146730437874794f, 1.365552306175232f, 0.614363968372345f, 4.44060410729856e-15f},
    {-0.9951847195625305f, 0.04906770959496498f, 1.3961477279663086f, 0.7063984870910645f, 2.3178410738991647e-15f},
};
// Sums of the TRT weights over the backward and forward samples: N_TRT is
// that weight times factors that do not depend on phi, so each half of the
// quadrature sums it in one product.
constexpr float kDualTrtBack = 1.9437659917028196f, kDualTrtFwd = 1.0136385360493365e-08f;

// Azimuthal sample i of the quadrature (i < 8 backward, i >= 8 forward).
__device__ __forceinline__ Lobes dual_sample(uint32_t i, float cos_theta_d, F3 sigma_a) {
    const DualAngle s = kDualAngles[i];
    const AzimuthalD d = azimuthal_d(cos_theta_d, sigma_a);
    float n_r;
    F3 n_tt;
    azimuthal_r_tt(cos_theta_d, d, sigma_a, s.cos_half_phi, s.h_k, s.w_tt, n_r, n_tt);
    return {splat3(n_r), n_tt, d.t_trt * (s.w_trt * d.trt_f)};
}

__device__ __forceinline__ Lobes lobes_add(const Lobes& a, const Lobes& b) {
    return {a.r + b.r, a.tt + b.tt, a.trt + b.trt};
}

// dual_scatter's closing part from the forward and backward quadrature sums.
__device__ DualScatter dual_finish(Lobes fwd, Lobes back, const Shape& shape);

// dual_scatter (shade.wesl:153-190): the 16 samples unrolled over the angle
// table; the theta_d parts are formed once and the TRT sums in closed form.
__device__ DualScatter dual_scatter(float cos_theta_d, F3 sigma_a, const Shape& shape) {
    const AzimuthalD d = azimuthal_d(cos_theta_d, sigma_a);
    float r_back = 0.0f, r_fwd = 0.0f;
    F3 tt_back = splat3(0.0f), tt_fwd = splat3(0.0f);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        float n_r;
        F3 n_tt;
        azimuthal_r_tt(cos_theta_d, d, sigma_a, kDualAngles[i].cos_half_phi, kDualAngles[i].h_k, kDualAngles[i].w_tt,
                       n_r, n_tt);
        if (i < 8) {
            r_back += n_r;
            tt_back = tt_back + n_tt;
        } else {
            r_fwd += n_r;
            tt_fwd = tt_fwd + n_tt;
        }
    }
    const Lobes back = {splat3(r_back), tt_back, d.t_trt * (kDualTrtBack * d.trt_f)};
    const Lobes fwd = {splat3(r_fwd), tt_fwd, d.t_trt * (kDualTrtFwd * d.trt_f)};
    return dual_finish(fwd, back, shape);
}

__device__ DualScatter dual_finish(Lobes fwd, Lobes back, const Shape& shape) {
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
    F3 a_back = fdiv3(a_b * af2, rest) + fdiv3(ab3 * af2, rest3);
    F3 delta_b = alpha_b * (splat3(1.0f) - fdiv3(2.0f * ab2, rest * rest)) +
                 fdiv3(alpha_b * (2.0f * rest * rest + 4.0f * af2 * ab2), rest3);
    F3 beta_f = vsqrt(beta_f2), beta_b = vsqrt(beta_b2);
    F3 sigma_b = fdiv3((splat3(1.0f) + 0.7f * af2) *
                           (a_b * vsqrt(2.0f * beta_f2 + beta_b2) + ab3 * vsqrt(2.0f * beta_f2 + 3.0f * beta_b2)),
                       vmax(a_b + ab3 * (2.0f * beta_f + 3.0f * beta_b), splat3(1e-5f)));
    return {fwd, a_f, beta_f2, a_back, delta_b, sigma_b * sigma_b};
}

// The geometry below follows the oracle's unfused, correctly rounded
// operations: theta_i = asin(t . l) and the components across the fiber are
// ill-conditioned where the light or the view runs along the fiber, and an ulp
// there moves a narrow lobe by far more than the tolerance.
#pragma clang fp contract(off)

// Angles of one light/view pair around the fiber (shade.wesl:203-215).
struct BsdfGeom {
    F3 t, l;
    float theta_h, cos_theta_d, cos_phi;
};

// The view's side of bsdf_geom, once per pixel: t and v normalized
// (safe_normalize of the tangent and view vectors), sin and asin of theta_o,
// the view's component across the fiber and its squared length.
struct ViewGeom {
    F3 t, v, t_o;
    float sin_o, theta_o, to2;
};
__device__ __forceinline__ ViewGeom view_geom(F3 t, F3 v) {
    ViewGeom g;
    g.t = t;
    g.v = v;
    g.sin_o = clampf(dot3(t, v), -1.0f, 1.0f);
    g.theta_o = fasin(g.sin_o);
    g.t_o = v - t * g.sin_o;
    g.to2 = dot3(g.t_o, g.t_o);
    return g;
}

// bsdf_geom for a light direction `l` already normalized. cos_phi is formed
// with the exact operations: near cos_phi = -1 the R lobe takes
// sqrt(0.5 + 0.5 cos_phi), which turns an ulp of cos_phi into a large
// relative change.
__device__ __forceinline__ BsdfGeom bsdf_geom_u(const ViewGeom& vg, F3 l) {
    BsdfGeom g;
    float sin_i = sclamp(dot3(vg.t, l), -1.0f, 1.0f);
    float theta_i = fasin(sin_i);
    g.theta_h = theta_i + vg.theta_o;
    // cos(|theta_i - theta_o| / 2): v_cos takes revolutions, and halving the
    // reciprocal of 2 pi is exact, so this is fcos(0.5 |...|) bit for bit
    g.cos_theta_d = smax(__builtin_amdgcn_cosf(fabsf(theta_i - vg.theta_o) * (0.5f * 0.15915494309189535f)), 1e-3f);
    F3 t_i = l - vg.t * sin_i;
    // proj^2 lies in [1e-8, 16] and proj in [1e-4, 4] (t, v and l are unit
    // or zero vectors): the mid-range root and, for |dot| >= 2^-60, divide
    float proj = sqrt_rn_mid(smax(dot3(t_i, t_i) * vg.to2, 1e-8f));
    const float num = dot3(t_i, vg.t_o);
    float c = 0.0f;
    if (div_mid_ok(num)) c = div_rn_mid(num, proj);
    else c = rdiv(num, proj);
    g.cos_phi = sclamp(c, -1.0f, 1.0f);
    g.t = vg.t;
    g.l = l;
    return g;
}

// t and v already normalized; `light` is normalized here (safe_normalize).
__device__ __forceinline__ BsdfGeom bsdf_geom_n(F3 t, F3 v, F3 light) {
    return bsdf_geom_u(view_geom(t, v), safe_normalize(light));
}

#pragma clang fp contract(fast)

// Local backscattering lobe (shade.wesl:220); inv_sb = gaussian_rcp(sqrt(sigma_b2)).
__device__ __forceinline__ F3 local_back_r(float theta_h, F3 a_back, F3 delta_b, F3 inv_sb) {
    return 2.0f * a_back * gaussian3_r(inv_sb, splat3(theta_h) - delta_b) * (1.0f / kPi);
}

__device__ __forceinline__ F3 local_back(float theta_h, F3 a_back, F3 delta_b, F3 sigma_b2) {
    return local_back_r(theta_h, a_back, delta_b, gaussian_rcp(vsqrt(sigma_b2)));
}

// Global forward scattering through `depth` of hair (shade.wesl:225-242);
// zero at depth <= 1e-3.
__device__ __forceinline__ F3 global_spread(float theta_h, float depth, float direct, const Shape& shape,
                                            const DualScatter& ds) {
    if (!(depth > 1e-3f)) return splat3(0.0f);
    F3 sigma_f2 = depth * ds.beta_f2;
    F3 transmit = {hpow(ds.a_f.x, depth), hpow(ds.a_f.y, depth), hpow(ds.a_f.z, depth)};
    F3 beta2 = shape.beta * shape.beta;
    F3 x = splat3(theta_h);
    F3 spread = gaussian3(vsqrt(splat3(beta2.x) + sigma_f2), x - splat3(shape.alpha.x)) * ds.fwd.r +
                gaussian3(vsqrt(splat3(beta2.y) + sigma_f2), x - splat3(shape.alpha.y)) * ds.fwd.tt +
                gaussian3(vsqrt(splat3(beta2.z) + sigma_f2), x - splat3(shape.alpha.z)) * ds.fwd.trt;
    F3 back_spread = 2.0f * ds.a_back * gaussian3(vsqrt(ds.sigma_b2 + sigma_f2), x - ds.delta_b) * (1.0f / kPi);
    return (1.0f - direct) * transmit * 0.7f * (spread * (1.0f / kPi) + kPi * 0.7f * back_spread);
}


// The DOM lookup's texel and slice choices keep unfused arithmetic.
#pragma clang fp contract(off)

struct DomTaps {
    uint32_t front[4];
    uint4 counts[4][4];
    float tx, ty, dist;
};

// Where one DOM lookup reads: the four taps' texels (a tap outside the map
// reads texel 0 and is marked outside), the bilinear fractions and the
// light-space distance shifted by -ln(beta) * dom_scale (shade.wesl:279-299).
// False when the pixel projects outside the light's clip volume or the DOM
// is off (visibility 1).
struct DomPos {
    uint32_t pix[4];
    bool inside[4];
    float tx, ty, dist;
};
__device__ __forceinline__ bool dom_pos(const View& p, F3 world, float beta, DomPos& P) {
    if (p.flags[3] == 0u) return false;
    float cx = mat_row(p.lclip, 0, world.x, world.y, world.z, 1.0f);
    float cy = mat_row(p.lclip, 1, world.x, world.y, world.z, 1.0f);
    float cw = mat_row(p.lclip, 3, world.x, world.y, world.z, 1.0f);
    if (cw <= 0.0f) return false;
    float ndc_x = rdiv(cx, cw), ndc_y = rdiv(cy, cw);
    float ux = ndc_x * 0.5f + 0.5f;
    float uy = 1.0f - (ndc_y * 0.5f + 0.5f);
    if (ux < 0.0f || uy < 0.0f || ux > 1.0f || uy > 1.0f) return false;
    uint32_t dom_w = to_u32(p.dom_info.x), dom_h = to_u32(p.dom_info.y);
    P.dist = dot3(world - xyz(p.light_eye), xyz(p.light_forward)) - (-hlog(mx(beta, 1e-4f)) * p.filter_params.z);
    float fx = ux * u2f(dom_w) - 0.5f;
    float fy = uy * u2f(dom_h) - 0.5f;
    float flx = floorf(fx), fly = floorf(fy);
    int32_t x0 = to_i32(flx), y0 = to_i32(fly);
    P.tx = fx - flx;
    P.ty = fy - fly;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int32_t x = x0 + (c & 1), y = y0 + (c >> 1);
        P.inside[c] = x >= 0 && y >= 0 && x < (int32_t)dom_w && y < (int32_t)dom_h;
        P.pix[c] = P.inside[c] ? (uint32_t)y * dom_w + (uint32_t)x : 0u;
    }
    return true;
}

// The four bilinear taps of one DOM lookup: every load is issued before any
// is used, so the lookup costs one memory round trip.
__device__ __forceinline__ bool dom_taps(const View& p, const uint32_t* __restrict__ dom, F3 world, float beta,
                                         DomTaps& T) {
    DomPos P;
    if (!dom_pos(p, world, beta, P)) return false;
    T.tx = P.tx;
    T.ty = P.ty;
    T.dist = P.dist;
    const uint4* counts = reinterpret_cast<const uint4*>(dom + (size_t)to_u32(p.dom_info.x) * to_u32(p.dom_info.y));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        uint32_t f = dom[P.pix[c]];
        T.front[c] = P.inside[c] ? f : HM_DOM_EMPTY;
#pragma unroll
        for (int q = 0; q < 4; ++q) T.counts[c][q] = counts[(size_t)P.pix[c] * (HM_DOM_LAYERS / 4) + q];
    }
    return true;
}

// The slice position t of light-space distance `dist` behind a tap's front
// (shade.wesl:265-272): (dist - z_front - 0.05 thickness) / thickness held
// at >= 0; a NaN t is taken as 0 (it adds no slice).
__device__ __forceinline__ float dom_tap_t(const View& p, uint32_t front_q, float dist) {
    float near_ = p.dom_info.z, far_ = p.dom_info.w;
    float z_front = dequantize_depth(front_q, near_, far_);
    float thickness = rdiv(mx(far_ - near_, 1e-3f), 16.0f);
    float rel = mx(dist - z_front - thickness * 0.05f, 0.0f);
    float t = rdiv(rel, thickness);
    return t > 0.0f ? t : 0.0f;
}

// dom_optical_depth's sum over four slices s0 .. s0 + 3, continued from
// `occ` in slice order: slice i adds count * min(max(t - i, 0), 1), which is
// the count itself where i + 1 <= t (t - i >= 1 rounds to >= 1), the count
// times t - i where i < t < i + 1, and +0 elsewhere: the same additions and
// roundings as the reference's branches, without them.
__device__ __forceinline__ float dom_occ4(float occ, uint4 c, float t, int s0) {
    const uint32_t cs[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = __builtin_amdgcn_fmed3f(t - (float)(s0 + k), 0.0f, 1.0f);
        occ = occ + u2f(cs[k]) * w;
    }
    return occ;
}

// Visibility through one tap at light-space distance `dist`.
__device__ __forceinline__ float dom_tap_visibility(const View& p, uint32_t front_q, const uint4 counts[4], float dist) {
    if (front_q == HM_DOM_EMPTY) return 1.0f;
    const float t = dom_tap_t(p, front_q, dist);
    float occ = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) occ = dom_occ4(occ, counts[q], t, 4 * q);
    return hexp(-occ * p.filter_params.w);
}

__device__ __forceinline__ float dom_taps_visibility(const View& p, const DomTaps& T) {
    float v00 = dom_tap_visibility(p, T.front[0], T.counts[0], T.dist);
    float v10 = dom_tap_visibility(p, T.front[1], T.counts[1], T.dist);
    float v01 = dom_tap_visibility(p, T.front[2], T.counts[2], T.dist);
    float v11 = dom_tap_visibility(p, T.front[3], T.counts[3], T.dist);
    return mixs(mixs(v00, v10, T.tx), mixs(v01, v11, T.tx), T.ty);
}

// The shading kernels' lookup: the taps in two pairs, each pair's loads issued
// together (two memory round trips), so that 34 words rather than 68 are in
// flight at the kernels' register peak. The same taps and blend as
// dom_taps_visibility.
__device__ float dom_visibility(const View& p, const uint32_t* __restrict__ dom, F3 world, float beta) {
    DomPos P;
    if (!dom_pos(p, world, beta, P)) return 1.0f;
    const uint4* counts = reinterpret_cast<const uint4*>(dom + (size_t)to_u32(p.dom_info.x) * to_u32(p.dom_info.y));
    float v[4];
#pragma unroll
    for (int h = 0; h < 4; h += 2) {
        uint32_t front[2];
        uint4 cnt[2][4];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint32_t f = dom[P.pix[h + c]];
            front[c] = P.inside[h + c] ? f : HM_DOM_EMPTY;
#pragma unroll
            for (int q = 0; q < 4; ++q) cnt[c][q] = counts[(size_t)P.pix[h + c] * (HM_DOM_LAYERS / 4) + q];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) v[h + c] = dom_tap_visibility(p, front[c], cnt[c], P.dist);
    }
    return mixs(mixs(v[0], v[1], P.tx), mixs(v[2], v[3], P.tx), P.ty);
}

#pragma clang fp contract(fast)

// Bevy's range window saturate(1 - (d^2 / r^2)^2)^2: zero from the range on.
__device__ __forceinline__ float range_falloff(float dist2, float inv_range2) {
    float factor = dist2 * inv_range2;
    float smooth = sclamp(1.0f - factor * factor, 0.0f, 1.0f);
    return smooth * smooth;
}

__device__ __forceinline__ float distance_attenuation(float dist2, float inv_range2) {
    return range_falloff(dist2, inv_range2) * frcp(smax(dist2, 1e-4f));
}

__device__ __forceinline__ float spot_mask(const hm_point_light& L, F3 world) {
    F3 sd = {L.light_custom_data[0], 0.0f, L.light_custom_data[1]};
    sd.y = fsqrt(mx(0.0f, 1.0f - sd.x * sd.x - sd.z * sd.z));
    if (L.flags & HM_LIGHT_FLAG_SPOT_Y_NEGATIVE) sd.y = -sd.y;
    F3 ltf = F3{L.position_radius[0], L.position_radius[1], L.position_radius[2]} - world;
    float cd = dot3(sd * -1.0f, fnormalize(ltf));
    float att = sclamp(cd * L.light_custom_data[2] + L.light_custom_data[3], 0.0f, 1.0f);
    return att * att;
}


#pragma clang fp contract(off)

// Per-pixel inputs of shade.wesl:fs (shade.wesl:539-571, 442-446).
struct PixelShade {
    F3 world, tangent, view, albedo, sigma_a;
    F3 t_n, v_n;  // safe_normalize(tangent), safe_normalize(view)
    ViewGeom vg;  // the view's side of every light's bsdf_geom
    float view_dist, beta, roughness, cos_theta_o, ao;
    Shape shape;
    // The light terms' per-pixel factors: -log2(e) sigma_a (absorption as a
    // base-2 exponent), the Gaussian widths' reciprocals times sqrt(log2(e) / 2)
    // (M_p = exp2(-((theta_h - alpha_p) gs_p)^2) ms_p) and over sqrt(2 pi).
    F3 sig2, gs, ms;
};

__device__ __forceinline__ PixelShade pixel_setup(const View& p, uint64_t word, uint32_t beta_word, uint32_t x,
                                                  uint32_t y) {
    PixelShade s;
    uint32_t low = (uint32_t)word;
    s.view_dist = dequantize_depth(gb_depth(word), p.screen.z, p.screen.w);
    s.world = reconstruct_world(p, (int32_t)x, (int32_t)y, s.view_dist);
    s.tangent = oct_decode(gb_tx(word), gb_ty(word));
    float uvw_z = u2f((low >> 6u) & 0x3fu) / 63.0f;
    s.beta = u2f(beta_word & 0xffu) / 255.0f;
    s.ao = p.flags[2] == 0u ? 1.0f : u2f(low & 0x3fu) / 63.0f;  // shade.wesl:544-547
    s.roughness = mx(p.appearance.x * (0.65f + 0.7f * uvw_z), 0.04f);
    float tilt = p.appearance.y;
    s.albedo = xyz(p.albedo) * (0.65f + 0.35f * uvw_z);
    // Orthographic when the projection's last row is (0, 0, 0, 1).
    bool ortho = p.clip[3] == 0.0f && p.clip[7] == 0.0f && p.clip[11] == 0.0f && p.clip[15] == 1.0f;
    s.view = ortho ? safe_normalize(F3{p.clip[2], p.clip[6], p.clip[10]}) : safe_normalize(xyz(p.camera_pos) - s.world);
    float bb = clampf(s.roughness * s.roughness, 0.02f, 1.0f);
    s.shape.alpha = {tilt, -0.5f * tilt, -1.5f * tilt};
    s.shape.beta = {bb, 0.5f * bb, 2.0f * bb};
    s.shape.inv_beta = gaussian_rcp(s.shape.beta);
    s.t_n = safe_normalize_unit(s.tangent);  // the decoded tangent is normalized
    s.v_n = safe_normalize(s.view);
    s.vg = view_geom(s.t_n, s.v_n);
    float tv = dot3(s.tangent, s.view);
    s.cos_theta_o = sqrtf(mx(1.0f - tv * tv, 1e-6f));
    s.sigma_a = absorption_from_color(s.albedo, s.roughness);
    s.sig2 = s.sigma_a * -kLog2e;
    s.gs = s.shape.inv_beta * kGaussLog2;
    s.ms = s.shape.inv_beta * kInvSqrt2Pi;
    return s;
}

__device__ __forceinline__ F3 point_color(const hm_point_light& L) {
    return {L.color_inverse_square_range[0], L.color_inverse_square_range[1], L.color_inverse_square_range[2]};
}
__device__ __forceinline__ F3 point_pos(const hm_point_light& L) {
    return {L.position_radius[0], L.position_radius[1], L.position_radius[2]};
}

__device__ __forceinline__ int key_light(const ShadeArgs& a) {
    for (uint32_t i = 0; i < a.directional_count; ++i)
        if (a.directionals[i].flags & HM_LIGHT_FLAG_SHADOWS) return (int)i;
    return -1;
}

__device__ __forceinline__ F3 dir_of(const hm_directional_light& L) {
    return {L.direction_to_light[0], L.direction_to_light[1], L.direction_to_light[2]};
}

#pragma clang fp contract(fast)

// The dual-scattering terms of the local back lobe that every light uses,
// once per pixel: 0.7 times the lobe (shade.wesl:220, 232) is
// kb exp2(-((theta_h - delta_b) gb)^2) with r = gaussian_rcp(sqrt(sigma_b^2)),
// kb = 0.7 * 2 a_back r / (sqrt(2 pi) pi) and gb = r sqrt(log2(e) / 2).
struct LocalBack {
    F3 kb, gb, delta_b;
};

__device__ __forceinline__ LocalBack local_terms(const DualScatter& ds) {
    const F3 r = gaussian_rcp(vsqrt(ds.sigma_b2));
    return {ds.a_back * r * (1.4f * kInvSqrt2Pi / kPi), r * kGaussLog2, ds.delta_b};
}

// 0.7 times the local back lobe at theta_h.
__device__ __forceinline__ F3 back_lobe(float theta_h, const LocalBack& lb) {
    const F3 z = (splat3(theta_h) - lb.delta_b) * lb.gb;
    return lb.kb * F3{__builtin_amdgcn_exp2f(-z.x * z.x), __builtin_amdgcn_exp2f(-z.y * z.y),
                      __builtin_amdgcn_exp2f(-z.z * z.z)};
}

// Single scattering M_p N_p summed over R, TT and TRT (shade.wesl:217-219,
// azimuthal :94-116) for one light's angles. The exponentials of each of
// M_TT N_TT and M_TRT N_TRT are taken as one exp2 per channel: the
// Gaussian's exponent, the absorption exp(-sigma_a k) and the azimuthal
// weight exp(c0 cos phi + c1) summed in base 2.
__device__ __forceinline__ F3 single_lobes(const BsdfGeom& g, const PixelShade& px) {
    const float ct = g.cos_theta_d, cp = g.cos_phi;
    const float inv_cd = frcp(ct);
    const float a = frcp(1.19f * inv_cd + 0.36f * ct);
    const float f_trt = fresnel(ct * 0.5f);
    const float chp = fsqrt(sclamp(0.5f + 0.5f * cp, 0.0f, 1.0f));  // cos(phi / 2)
    const float n_r = 0.25f * chp * fresnel(ct * chp);
    const float h = chp * (1.0f + a * (0.6f - 0.8f * cp));
    const float f_tt = fresnel(ct * fsqrt(sclamp(1.0f - h * h, 0.0f, 1.0f)));
    const float s_tt = fsqrt(sclamp(1.0f - h * h * a * a, 0.0f, 1.0f));
    const float th = g.theta_h;
    const float dr = (th - px.shape.alpha.x) * px.gs.x;
    const float dt = (th - px.shape.alpha.y) * px.gs.y;
    const float dp = (th - px.shape.alpha.z) * px.gs.z;
    const float r = __builtin_amdgcn_exp2f(-dr * dr) * px.ms.x * n_r;
    const float e_tt = cp * (-3.65f * kLog2e) + (-3.98f * kLog2e) - dt * dt;
    const float e_trt = cp * (17.0f * kLog2e) + (-16.78f * kLog2e) - dp * dp;
    const float k_tt = 2.0f * s_tt * inv_cd, k_trt = 3.2f * inv_cd;
    const float w_tt = px.ms.y * (1.0f - f_tt) * (1.0f - f_tt);
    const float w_trt = px.ms.z * ((1.0f - f_trt) * (1.0f - f_trt) * f_trt);
    auto ch = [&](float sig2) {
        return r + w_tt * __builtin_amdgcn_exp2f(sig2 * k_tt + e_tt) + w_trt * __builtin_amdgcn_exp2f(sig2 * k_trt + e_trt);
    };
    return {ch(px.sig2.x), ch(px.sig2.y), ch(px.sig2.z)};
}

// single_lobes + back_lobe, the back lobe's weight taken in a fused multiply-add.
__device__ __forceinline__ F3 lobes(const BsdfGeom& g, const PixelShade& px, const LocalBack& lb) {
    const F3 single = single_lobes(g, px);
    const float th = g.theta_h;
    auto ch = [&](float sc, float kb, float gb, float db) {
        const float z = (th - db) * gb;
        return __builtin_fmaf(kb, __builtin_amdgcn_exp2f(-z * z), sc);
    };
    return {ch(single.x, lb.kb.x, lb.gb.x, lb.delta_b.x), ch(single.y, lb.kb.y, lb.gb.y, lb.delta_b.y),
            ch(single.z, lb.kb.z, lb.gb.z, lb.delta_b.z)};
}

// The key light's multiple scattering per unit light colour: hair_bsdf
// (shade.wesl:193-244) minus its direct single scattering,
// direct * 0.7 * back + [depth > 1e-3] * global spread.
__device__ __forceinline__ F3 key_multiple_scattering(float theta_h, float depth, const Shape& shape,
                                                      const DualScatter& ds) {
    const float direct = hexp(-depth);
    F3 ms = direct * (0.7f * local_back(theta_h, ds.a_back, ds.delta_b, ds.sigma_b2));
    if (depth > 1e-3f) ms = ms + global_spread(theta_h, depth, direct, shape, ds);
    return ms;
}

// Light i of a light array the kernel only reads, loaded as constant data at
// a wave-uniform index (scalar loads).
template <class T>
__device__ __forceinline__ T uniform_light(const T* lights, uint32_t i) {
    static_assert(sizeof(T) % 16 == 0, "lights are whole 16-byte vectors");
    constexpr uint32_t V = sizeof(T) / 16;
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(4))) u32x4* const_vec;
    const const_vec c = (const_vec)lights + (size_t)__builtin_amdgcn_readfirstlane(i) * V;
    struct Words {
        u32x4 v[V];
    } w;
#pragma unroll
    for (uint32_t k = 0; k < V; ++k) w.v[k] = c[k];
    return __builtin_bit_cast(T, w);
}

// Adds point light L's contribution (shade.wesl:449-505) with back-lobe
// terms `lb` to `color`; nothing where the light's range does not reach the
// pixel, as the cluster would not list the light.
// KIND: 0 point light, 1 spot light, -1 by the light's own kind.
template <int KIND = -1>
__device__ __forceinline__ void point_add(const hm_point_light& L, const PixelShade& px, const LocalBack& lb, F3& color) {
    F3 ltf = point_pos(L) - px.world;
    float att = distance_attenuation(dot3(ltf, ltf), L.color_inverse_square_range[3]);
    if (att == 0.0f) return;
    // safe_normalize twice, as the light loop (shade.wesl:457) and hair_bsdf
    // (shade.wesl:204) each normalize; the second one's vector is within a few
    // ulps of unit length, where safe_normalize_unit gives the same bits cheaper
    BsdfGeom g = bsdf_geom_u(px.vg, safe_normalize_unit(safe_normalize_mid(ltf)));
    const F3 lobe = lobes(g, px, lb);
    if (KIND == 1 || (KIND < 0 && L.kind != HM_LIGHT_POINT)) att = att * spot_mask(L, px.world);
    const F3 w = point_color(L) * att;
    color = {__builtin_fmaf(lobe.x, w.x, color.x), __builtin_fmaf(lobe.y, w.y, color.y),
             __builtin_fmaf(lobe.z, w.z, color.z)};
}

// Directional light i's contribution: for the key light (optical depth
// `depth`) exp(-depth) * single + key_ms, for the others single + 0.7 * back.
__device__ __forceinline__ F3 directional_term(const hm_directional_light& L, bool is_key, const PixelShade& px,
                                               const LocalBack& lb, float depth, F3 key_ms) {
    BsdfGeom g = bsdf_geom_u(px.vg, safe_normalize_unit(dir_of(L)));
    F3 lobe = is_key ? hexp(-depth) * single_lobes(g, px) + key_ms : lobes(g, px, lb);
    return lobe * F3{L.color[0], L.color[1], L.color[2]};
}

__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Rect light L's contribution (shade.wesl:503-518): hair_bsdf with shadow 1
// and no optical depth, times the rectangle's solid angle and range window.
__device__ __forceinline__ F3 rect_term(const hm_rect_light& L, const PixelShade& px, const LocalBack& lb) {
    F3 to_center = F3{L.position[0], L.position[1], L.position[2]} - px.world;
    float dist2 = mx(dot3(to_center, to_center), 1e-4f);
    F3 incident = to_center * __builtin_amdgcn_rsqf(dist2);
    F3 normal = fnormalize(cross3(F3{L.up[0], L.up[1], L.up[2]}, F3{L.right[0], L.right[1], L.right[2]}));
    float facing = mx(dot3(normal, incident), 0.0f);
    float falloff = range_falloff(dist2, frcp(mx(L.range * L.range, 1e-4f)));
    if (facing == 0.0f || falloff == 0.0f) return splat3(0.0f);
    float solid = L.width * L.height * facing * frcp(dist2);
    BsdfGeom g = bsdf_geom_u(px.vg, safe_normalize_unit(incident));
    return lobes(g, px, lb) * F3{L.color[0], L.color[1], L.color[2]} * (solid * falloff);
}

#pragma clang fp contract(off)

// ---- the environment probe --------------------------------------------------

__device__ __forceinline__ F3 cubemap_dir(const EnvArgs& e, F3 d) {  // quat_rotate, then z negated
    F3 u = {e.rotation[0], e.rotation[1], e.rotation[2]};
    F3 t = 2.0f * cross3(u, d);
    F3 r = d + e.rotation[3] * t + cross3(u, t);
    return {r.x, r.y, -r.z};
}

// Vulkan's cube map face selection (ties go to x, then y; a zero or
// non-finite direction reads the middle of face +x): face and its
// coordinates s, t in [0, 1].
struct CubeUV { uint32_t face; float s, t; };
__device__ __forceinline__ CubeUV cube_face(F3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t face;
    float ma, sc, tc;
    if (ax >= ay && ax >= az) {
        ma = ax;
        face = d.x >= 0.0f ? 0u : 1u;
        sc = d.x >= 0.0f ? -d.z : d.z;
        tc = -d.y;
    } else if (ay >= az) {
        ma = ay;
        face = d.y >= 0.0f ? 2u : 3u;
        sc = d.x;
        tc = d.y >= 0.0f ? d.z : -d.z;
    } else {
        ma = az;
        face = d.z >= 0.0f ? 4u : 5u;
        sc = d.z >= 0.0f ? d.x : -d.x;
        tc = -d.y;
    }
    if (!(ma > 0.0f)) return {0u, 0.5f, 0.5f};
    float r = frcp(ma);
    return {face, 0.5f * (sc * r + 1.0f), 0.5f * (tc * r + 1.0f)};
}

// Bilinear inside one face, clamp to edge, RGB of the RGBA texels. The
// texel coordinate is first held to [-1, size] (which changes no finite
// result) so that every index stays inside the face.
__device__ __forceinline__ F3 cube_bilinear(const float4* __restrict__ level, uint32_t size, const CubeUV& c) {
    const float4* face = level + (size_t)c.face * size * size;
    const float fs = u2f(size);
    float u = fminf(fmaxf(c.s * fs - 0.5f, -1.0f), fs);
    float v = fminf(fmaxf(c.t * fs - 0.5f, -1.0f), fs);
    float fu = floorf(u), fv = floorf(v);
    float wx = u - fu, wy = v - fv;
    int32_t x0 = (int32_t)fu, y0 = (int32_t)fv, hi = (int32_t)size - 1;
    uint32_t xa = (uint32_t)min(max(x0, 0), hi), xb = (uint32_t)min(max(x0 + 1, 0), hi);
    uint32_t ya = (uint32_t)min(max(y0, 0), hi), yb = (uint32_t)min(max(y0 + 1, 0), hi);
    float4 t00 = face[ya * size + xa], t10 = face[ya * size + xb];
    float4 t01 = face[yb * size + xa], t11 = face[yb * size + xb];
    F3 top = {mixs(t00.x, t10.x, wx), mixs(t00.y, t10.y, wx), mixs(t00.z, t10.z, wx)};
    F3 bot = {mixs(t01.x, t11.x, wx), mixs(t01.y, t11.y, wx), mixs(t01.z, t11.z, wx)};
    return {mixs(top.x, bot.x, wy), mixs(top.y, bot.y, wy), mixs(top.z, bot.z, wy)};
}

// The specular map at `level` (held to [0, mips - 1]), linear between levels.
__device__ __forceinline__ F3 cube_sample_spec(const EnvArgs& e, const CubeUV& c, float level) {
    const float top = u2f(e.spec_mips - 1u);
    level = fminf(fmaxf(level, 0.0f), top);
    const float fl = floorf(level);
    const uint32_t l0 = (uint32_t)fl, l1 = min(l0 + 1u, e.spec_mips - 1u);
    const F3 a = cube_bilinear(e.specular + e.spec_base[l0], max(e.spec_size >> l0, 1u), c);
    if (l1 == l0) return a;
    const F3 b = cube_bilinear(e.specular + e.spec_base[l1], max(e.spec_size >> l1, 1u), c);
    const float f = level - fl;
    return {mixs(a.x, b.x, f), mixs(a.y, b.y, f), mixs(a.z, b.z, f)};
}

// probe_radiance (shade.wesl:396-427) of one hit pixel.
__device__ F3 probe_radiance(const EnvArgs& e, const PixelShade& px) {
    const F3 t = px.tangent, v = px.view;
    const F3 fiber_n = safe_normalize(v - t * dot3(v, t));
    const float theta_o = fasin(clampf(dot3(t, v), -1.0f, 1.0f));
    const float theta_i = px.shape.alpha.x - theta_o;  // tilt - theta_o
    float si, ci;
    __sincosf(theta_i, &si, &ci);
    const F3 spec = cube_sample_spec(e, cube_face(cubemap_dir(e, t * si + fiber_n * ci)),
                                     px.roughness * u2f(e.spec_mips - 1u));
    const F3 diff = cube_bilinear(e.diffuse, e.diff_size, cube_face(cubemap_dir(e, fiber_n)));
    const float cos_o = mx(fcos(theta_o), 1e-3f);
    const float f_r = fresnel(cos_o), f_trt = fresnel(cos_o * 0.5f);
    const F3 trt = vexp(px.sigma_a * (-3.2f * frcp(cos_o))) * (0.76f * (1.0f - f_trt) * (1.0f - f_trt) * f_trt);
    // scatter_tint(albedo, 0.5) = (albedo / luminance(albedo))^0.25 (shade.wesl:62-85)
    const F3 al = px.albedo;
    const float inv_lum = frcp(mx(al.x * 0.2126f + al.y * 0.7152f + al.z * 0.0722f, 1e-4f));
    const F3 scatter = al * F3{hpow(al.x * inv_lum, 0.25f), hpow(al.y * inv_lum, 0.25f), hpow(al.z * inv_lum, 0.25f)};
    return (spec * (splat3(f_r) + trt) + diff * scatter) * e.intensity;
}

// ---- light culling per wave -------------------------------------------------

// World-space box of the wave's valid pixels (every lane of the wave calls).
struct WaveBox { F3 lo, hi; };
__device__ __forceinline__ WaveBox wave_box(F3 w, bool valid) {
    const float inf = __builtin_huge_valf();
    F3 lo = valid ? w : splat3(inf), hi = valid ? w : splat3(-inf);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = vmin(lo, F3{__shfl_xor(lo.x, o, 64), __shfl_xor(lo.y, o, 64), __shfl_xor(lo.z, o, 64)});
        hi = vmax(hi, F3{__shfl_xor(hi.x, o, 64), __shfl_xor(hi.y, o, 64), __shfl_xor(hi.z, o, 64)});
    }
    // the same in every lane: held in scalar registers
    auto u = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
    return {{u(lo.x), u(lo.y), u(lo.z)}, {u(hi.x), u(hi.y), u(hi.z)}};
}

// Lights base .. min(base + 63, end - 1) of the point/spot list whose range
// reaches the box, one bit per light (every lane of
// the wave calls; lane i tests light base + i). A light is dropped only when
// the box lies beyond its range by a relative margin of 1e-5, far above the
// rounding of the per-pixel distance: distance_attenuation is then zero at
// every pixel of the wave, so dropping it changes no colour.
__device__ __forceinline__ uint64_t reaching_lights(const ShadeArgs& a, const WaveBox& box, uint32_t base,
                                                    uint32_t end) {
    const uint32_t j = base + (uint32_t)(threadIdx.x & 63u);
    bool keep = false;
    if (j < end) {
        const hm_point_light& L = a.points[j];
        const F3 c = point_pos(L);
        const F3 d = vmax(vmax(box.lo - c, c - box.hi), splat3(0.0f));
        keep = !(dot3(d, d) * L.color_inverse_square_range[3] > 1.00001f);
    }
    return __ballot(keep);
}

#pragma clang fp contract(fast)

// Lights begin .. end - 1 of the point/spot list (all point lights, or all
// spot lights) added to `color` in order. The list is read-only and the index
// wave-uniform: each light's words come through the scalar cache, and the
// next light's are requested before this light's terms are formed, so that
// the load's latency hides behind them.
template <bool SPOT>
__device__ __forceinline__ F3 light_range(const ShadeArgs& a, uint32_t begin, uint32_t end, const PixelShade& px,
                                          const LocalBack& lb, F3 color) {
    if (begin >= end) return color;
    hm_point_light L = uniform_light(a.points, begin);
    for (uint32_t i = begin; i < end; ++i) {
        const hm_point_light next = uniform_light(a.points, i + 1u < end ? i + 1u : i);
        point_add<SPOT ? 1 : 0>(L, px, lb, color);
        L = next;
    }
    return color;
}

// Colour of one hit pixel (shade.wesl:442-521, 585-588), Bevy shadow factors
// 1, from the back-lobe terms `lb` and the key light's multiple scattering
// `key_ms`. Point lights, then spot lights (the order of Bevy's cluster
// ranges), then directional, then rect lights; then the probe times the AO.
// CULL: every lane of the wave calls, `valid` false for lanes without a
// pixel, and the point and spot lists are culled against the wave's pixels,
// which is what Bevy's clusters do for a froxel. Otherwise only lanes with
// a pixel call, and each light's range window is tested per pixel. Rect
// lights and the probe only in the SHADE_EXTRAS variant.
template <uint32_t F>
__device__ F3 pixel_color(const ShadeArgs& a, const PixelShade& px, bool valid, const LocalBack& lb, int key,
                          float depth, F3 key_ms) {
    constexpr bool CULL = (F & SHADE_CULL) != 0u;
    F3 color = splat3(0.0f);
    if (!CULL) {
        color = light_range<false>(a, 0u, a.spot_first, px, lb, color);
        color = light_range<true>(a, a.spot_first, a.point_count, px, lb, color);
    } else if (a.point_count) {
        const WaveBox box = wave_box(px.world, valid);
        for (int pass = 0; pass < 2; ++pass) {
            const uint32_t lo = pass == 0 ? 0u : a.spot_first, hi = pass == 0 ? a.spot_first : a.point_count;
            for (uint32_t base = lo; base < hi; base += 64u) {
                uint64_t reach = reaching_lights(a, box, base, hi);
                while (reach) {
                    const uint32_t i = base + (uint32_t)__builtin_ctzll(reach);
                    reach &= reach - 1u;
                    // read-only, wave-uniform: through the scalar cache in one
                    // load rather than as dependent vector loads
                    const hm_point_light L = uniform_light(a.points, i);
                    if (valid) point_add(L, px, lb, color);
                }
            }
        }
    }
    if (!valid) return color;
    for (uint32_t i = 0; i < a.directional_count; ++i)
        color = color + directional_term(uniform_light(a.directionals, i), (int)i == key, px, lb, depth, key_ms);
    if (F & SHADE_EXTRAS) {
        for (uint32_t i = 0; i < a.rect_count; ++i) color = color + rect_term(uniform_light(a.rects, i), px, lb);
        if (a.env.on) color = color + probe_radiance(a.env, px) * px.ao;
    }
    return color * a.exposure;
}

// theta_h of the key light at this pixel (bsdf_geom_n's).
__device__ __forceinline__ float key_theta_h(const PixelShade& px, F3 key_dir) {
    return bsdf_geom_u(px.vg, safe_normalize_unit(key_dir)).theta_h;
}

// The plain method's colour of one hit pixel: dual scattering evaluated
// (shade.wesl:153-190), the DOM visibility for the first shadow-casting
// directional light (the key light). SHADE_CULL: every lane of the wave
// calls, lanes with `valid` false have no pixel and return zero.
template <uint32_t F>
__device__ F3 plain_pixel_color(const ShadeArgs& a, const View& p, const PixelShade& px, bool valid) {
    const int key = key_light(a);
    float depth = 0.0f;
    F3 key_ms = splat3(0.0f);
    LocalBack lb = {splat3(0.0f), splat3(0.0f), splat3(0.0f)};
    if (valid) {
        const DualScatter ds = dual_scatter(px.cos_theta_o, px.sigma_a, px.shape);
        if (key >= 0) {
            depth = -hlog(mx(dom_visibility(p, a.dom, px.world, px.beta), 1e-6f));
            key_ms = key_multiple_scattering(key_theta_h(px, dir_of(a.directionals[key])), depth, px.shape, ds);
        }
        lb = local_terms(ds);
    }
    return pixel_color<F>(a, px, valid, lb, key, depth, key_ms);
}

#pragma clang fp contract(off)

}  // namespace hm
