"""Independent pins of the oracle's shading, DOM lookup and filter functions
(shade.wesl:153-327, 523-593; common.wgsl:212-229; filter.wgsl:9-135).

None of these restates the oracle's expression: each expected value comes
from another route, chosen so that a one-symbol slip (a dropped term, a
wrong sign, a swapped lobe, a wrong exponent) fails a check:
  - series sums and Gaussian-mixture moments for the dual-scattering terms
    (Zinke et al. 2008, "Dual Scattering Approximation for Fast Multiple
    Scattering in Hair", Eqs. 13-17) and the global spread;
  - the Karis 2016 single-scattering product M_p(theta_h) N_p from the pinned
    azimuthal terms;
  - hand-built one-texel and four-texel deep opacity maps;
  - closed forms of Bevy's spot cone and range window;
  - projection round trips and finite differences on the screen;
  - hand-computed two-tap filter weights.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from bevy_hair_amd import _native as N
from bevy_hair_amd import plugin, scene
from bevy_hair_amd.groom import synthetic_groom
from tests.oracle_scenes import screen_params

L = oracle.lib()
EMPTY64 = 0xFFFFFFFFFFFFFFFF


def d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def shape(roughness, tilt):
    """Karis 2016 lobe shifts and widths as the reference parameterizes them
    (shade.wesl:124-128): alpha = (t, -t/2, -3t/2), beta = (b, b/2, 2b)."""
    b = min(max(roughness * roughness, 0.02), 1.0)
    return np.array([tilt, -0.5 * tilt, -1.5 * tilt]), np.array([b, 0.5 * b, 2.0 * b])


def gauss(sd, x):
    sd = np.maximum(sd, 1e-3)
    return np.exp(-0.5 * x * x / (sd * sd)) / (math.sqrt(2.0 * math.pi) * sd)


def sigma_of(albedo, roughness):
    out = (C.c_double * 3)()
    L.or_absorption_from_color(d3(albedo), roughness, out)
    return np.array(list(out))


def dual_scatter(ctd, sigma, roughness, tilt):
    out = (C.c_double * 24)()
    L.or_dual_scatter(ctd, d3(sigma), roughness, tilt, out)
    o = np.array(list(out)).reshape(8, 3)
    return dict(fwd_r=o[0], fwd_tt=o[1], fwd_trt=o[2], a_f=o[3], beta_f2=o[4], a_back=o[5], delta_b=o[6],
                sigma_b2=o[7])


def dual_terms(fwd, back, roughness, tilt):
    out = (C.c_double * 18)()
    L.or_dual_terms((C.c_double * 9)(*np.ravel(fwd)), (C.c_double * 9)(*np.ravel(back)), roughness, tilt, out)
    o = np.array(list(out)).reshape(6, 3)
    return dict(a_f=o[0], alpha_f=o[1], beta_f2=o[2], a_back=o[3], delta_b=o[4], sigma_b2=o[5])


# ---- dual scattering (shade.wesl:153-190) ----------------------------------------------


def test_back_sums_match_quadrature():
    """The back lobe sums: twice the integral of N_p over the backward
    quarter turn phi in [0, pi/2] (the midpoint rule's target)."""
    from scipy.integrate import quad

    sigma, ctd = (0.2, 0.5, 1.5), 0.8
    out = (C.c_double * 18)()
    L.or_scatter_sums(ctd, d3(sigma), out)
    s = np.array(list(out)).reshape(6, 3)
    az = (C.c_double * 7)()

    def n(phi, k):
        L.or_azimuthal(ctd, math.cos(phi), d3(sigma), az)
        return az[k]

    nodes = [(i + 0.5) / 8.0 * 0.5 * math.pi for i in range(8)]
    for lobe, col in ((3, 0), (4, 1), (5, 4)):
        for ch in range(3):
            k = col + (ch if lobe > 3 else 0)
            # the midpoint rule applied by hand: 8 nodes over [0, pi/2], weight pi / 8 (both halves)
            mid = math.pi / 8.0 * sum(n(phi, k) for phi in nodes)
            # (the oracle forms the node angles with the shader's f32 pi)
            assert abs(s[lobe, ch] - mid) <= 1e-6 * abs(mid), (lobe, ch)
            val, _ = quad(lambda p: n(p, k), 0.0, math.pi / 2, epsabs=1e-12)
            assert abs(s[lobe, ch] - 2 * val) <= 5e-2 * abs(2 * val) + 1e-9, (lobe, ch)
    # forward and backward halves are different integrals (the phi ranges are not swapped)
    assert not np.allclose(s[0:3], s[3:6])


@pytest.fixture(scope="module")
def lobes():
    """Random lobe sums (forward total below 0.95, so a_f is not capped)."""
    rng = np.random.default_rng(5)
    cases = []
    for _ in range(12):
        fwd = rng.uniform(0.02, 0.3, (3, 3))
        fwd[0, :] = fwd[0, 0]  # R keeps the light's colour (one value for all channels)
        back = rng.uniform(0.01, 0.25, (3, 3))
        back[0, :] = back[0, 0]
        cases.append((fwd, back, rng.uniform(0.1, 0.8), rng.uniform(-0.2, 0.2)))
    return cases


def test_forward_energy_and_lobe_means(lobes):
    for fwd, back, r, t in lobes:
        d = dual_terms(fwd, back, r, t)
        a_f = np.minimum(fwd.sum(0), 0.95)
        al, be = shape(r, t)
        assert np.allclose(d["a_f"], a_f, rtol=1e-12)
        # energy-weighted means of the lobes' shifts and variances
        assert np.allclose(d["alpha_f"], (fwd * al[:, None]).sum(0) / a_f, rtol=1e-12, atol=1e-15)
        assert np.allclose(d["beta_f2"], (fwd * (be**2)[:, None]).sum(0) / a_f, rtol=1e-12)
    # a_f is capped at 0.95
    fwd = np.full((3, 3), 0.5)
    assert np.allclose(dual_terms(fwd, np.full((3, 3), 0.1), 0.3, 0.0)["a_f"], 0.95)


def test_back_attenuation_is_the_one_and_three_bounce_series(lobes):
    """Zinke Eq. 13 as the sum over paths: one backward bounce between i >= 1
    forward passes each way, a_b sum a_f^(2i), and three backward bounces,
    a_b^3 sum C(i+1, 2) a_f^(2i) (i = 1.. 4000)."""
    i = np.arange(1, 4001, dtype=np.float64)
    for fwd, back, r, t in lobes:
        d = dual_terms(fwd, back, r, t)
        a_f, a_b = np.minimum(fwd.sum(0), 0.95), back.sum(0)
        x = a_f[:, None] ** (2 * i[None, :])
        one = a_b * x.sum(1)
        three = a_b**3 * (x * ((i + 1) * i / 2)[None, :]).sum(1)
        assert np.allclose(d["a_back"], one + three, rtol=1e-10)


def test_back_shift_structure():
    """Zinke Eq. 14, as the reference writes it (shade.wesl:181-182): the
    backward shift is the back lobes' mean shift alpha_b times
    F = 1 - 2 a_b^2 / (1 - a_f^2)^2 + (2 (1 - a_f^2)^2 + 4 a_f^2 a_b^2) / (1 - a_f^2)^3.
    (The reference uses alpha_b in the second term where the paper's
    derivation carries the forward passes' shift; the oracle keeps the
    reference.) Checked through its limits and its linearity in the tilt."""
    r = 0.4
    fwd = np.array([[0.1] * 3, [0.2, 0.15, 0.1], [0.05, 0.04, 0.03]])
    back = np.array([[0.08] * 3, [0.03, 0.02, 0.01], [0.1, 0.08, 0.06]])
    assert np.allclose(dual_terms(fwd, back, r, 0.0)["delta_b"], 0.0, atol=1e-15)
    d1, d2 = dual_terms(fwd, back, r, 0.1), dual_terms(fwd, back, r, 0.2)
    assert np.allclose(d2["delta_b"], 2.0 * d1["delta_b"], rtol=1e-12)
    al, _ = shape(r, 0.1)
    alpha_b = (back * al[:, None]).sum(0) / back.sum(0)
    a_f, a_b = fwd.sum(0), back.sum(0)
    q = 1.0 - a_f**2
    F = 1.0 - 2.0 * a_b**2 / q**2 + (2.0 * q**2 + 4.0 * a_f**2 * a_b**2) / q**3
    assert np.allclose(d1["delta_b"], alpha_b * F, rtol=1e-12)
    # limits that a wrong exponent of (1 - a_f^2) would miss: a_b -> 0 gives
    # 1 + 2 / (1 - a_f^2); a_f -> 0 gives 3 - 2 a_b^2
    tiny = back * 1e-4  # (the lobe means divide by max(total, 1e-5): stay above it)
    al_b = (tiny * al[:, None]).sum(0) / tiny.sum(0)
    assert np.allclose(dual_terms(fwd, tiny, r, 0.1)["delta_b"] / al_b, 1.0 + 2.0 / q, rtol=1e-6)
    assert not np.allclose(1.0 + 2.0 / q, 1.0 + 2.0 / q**2, rtol=1e-2)
    zf = fwd * 1e-8
    assert np.allclose(dual_terms(zf, back, r, 0.1)["delta_b"] / alpha_b, 3.0 - 2.0 * a_b**2, rtol=1e-6)


def test_back_variance_single_lobe_case():
    """Zinke Eqs. 15-16: with all energy in one lobe p, beta_f = beta_b =
    beta_p and sigma_b = (1 + d_b a_f^2) (a_b sqrt(3) + a_b^3 sqrt(5)) beta_p
    / (a_b + 5 a_b^3 beta_p), d_b = 0.7 (shade.wesl:43)."""
    for lobe in range(3):
        for r in (0.2, 0.5):
            fwd = np.zeros((3, 3))
            back = np.zeros((3, 3))
            fwd[lobe] = [0.3, 0.2, 0.1]
            back[lobe] = [0.2, 0.1, 0.05]
            if lobe == 0:
                fwd[0], back[0] = 0.3, 0.2
            d = dual_terms(fwd, back, r, 0.05)
            _, be = shape(r, 0.05)
            bp = be[lobe]
            a_f, a_b = fwd.sum(0), back.sum(0)
            sb = (1.0 + 0.7 * a_f**2) * (a_b * math.sqrt(3.0) + a_b**3 * math.sqrt(5.0)) * bp \
                / (a_b + 5.0 * a_b**3 * bp)
            assert np.allclose(d["sigma_b2"], sb**2, rtol=1e-12), (lobe, r)
            assert np.allclose(d["beta_f2"], bp * bp, rtol=1e-12)


# ---- hair_bsdf (shade.wesl:193-244) over theta_h ----------------------------------------

T_AXIS, VIEW = (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)  # view normal to the fiber: theta_o = 0, cos theta_o = 1
ALBEDO, ROUGH, TILT = (0.35, 0.18, 0.09), 0.25, 0.07


def bsdf_parts(theta_i, depth, phi0=0.6):
    l = (math.sin(phi0) * math.cos(theta_i), math.sin(theta_i), math.cos(phi0) * math.cos(theta_i))
    out = (C.c_double * 10)()
    L.or_bsdf_parts(d3(T_AXIS), d3(VIEW), d3(l), d3(ALBEDO), ROUGH, TILT, depth, out)
    o = np.array(list(out))
    return o[0:3], o[3:6], o[6:9], o[9]


@pytest.fixture(scope="module")
def theta_grid():
    th = np.linspace(-math.pi / 2 + 1e-6, math.pi / 2 - 1e-6, 8001)
    return th, th[1] - th[0]


def moments(th, dx, f):
    m0 = f.sum(0) * dx
    m1 = (f * th[:, None]).sum(0) * dx / m0
    m2 = (f * th[:, None] ** 2).sum(0) * dx / m0 - m1**2
    return m0, m1, m2


def test_local_back_lobe_moments(theta_grid):
    """The local back term is 2 A_b N(theta_h; Delta_b, sigma_b^2) / pi: its
    integral over theta_h, mean and variance are those dual terms."""
    th, dx = theta_grid
    ds = dual_scatter(1.0, sigma_of(ALBEDO, ROUGH), ROUGH, TILT)
    back = np.array([bsdf_parts(t, 0.0)[1] for t in th])
    m0, m1, m2 = moments(th, dx, back)
    assert np.allclose(m0, 2.0 * ds["a_back"] / math.pi, rtol=1e-6)
    assert np.allclose(m1, ds["delta_b"], atol=1e-7)
    assert np.allclose(m2, ds["sigma_b2"], rtol=1e-5)


@pytest.mark.parametrize("depth", [0.6, 2.5])
def test_global_spread_moments(theta_grid, depth):
    """The global term (Tafuri 2019 weighting of Zinke's forward scattering):
    (1 - e^-d) a_f^d d_f (S / pi + pi d_b B) with S the forward lobes widened
    by d beta_f^2 and B the back lobe widened the same way. Its integral,
    mean and variance over theta_h follow from the Gaussian mixture."""
    th, dx = theta_grid
    ds = dual_scatter(1.0, sigma_of(ALBEDO, ROUGH), ROUGH, TILT)
    al, be = shape(ROUGH, TILT)
    parts = [bsdf_parts(t, depth) for t in th]
    glob = np.array([p[2] for p in parts])
    assert all(abs(p[3] - math.exp(-depth)) < 1e-15 for p in parts[:3])
    k = (1.0 - math.exp(-depth)) * ds["a_f"] ** depth * 0.7
    fwd = np.stack([ds["fwd_r"], ds["fwd_tt"], ds["fwd_trt"]])
    w = np.concatenate([fwd / math.pi, (math.pi * 0.7 * 2.0 * ds["a_back"] / math.pi)[None, :]])  # 4 x rgb
    mu = np.concatenate([np.repeat(al[:, None], 3, 1), ds["delta_b"][None, :]])
    var = np.concatenate([np.repeat((be**2)[:, None], 3, 1), ds["sigma_b2"][None, :]]) + depth * ds["beta_f2"]
    m0, m1, m2 = moments(th, dx, glob)
    assert np.allclose(m0, k * w.sum(0), rtol=1e-6)
    mean = (w * mu).sum(0) / w.sum(0)
    assert np.allclose(m1, mean, atol=1e-7)
    assert np.allclose(m2, (w * (var + mu**2)).sum(0) / w.sum(0) - mean**2, rtol=1e-5)
    # below depth 1e-3 there is no global term
    assert np.all(bsdf_parts(0.1, 5e-4)[2] == 0.0)


def test_single_scattering_is_karis_product():
    """single = sum_p M_p(theta_h) N_p(cos theta_d, cos phi), M_p Gaussians
    (alpha_p, beta_p) in theta_h, N_p the pinned azimuthal terms; here
    theta_o = 0 so theta_h = theta_i, theta_d = theta_i / 2 and cos phi is the
    light's fixed azimuth (away from grazing angles, where the projection's
    1e-8 floor moves cos phi)."""
    sigma = sigma_of(ALBEDO, ROUGH)
    al, be = shape(ROUGH, TILT)
    az = (C.c_double * 7)()
    phi0 = 0.6
    for t in np.linspace(-1.5, 1.5, 31):
        single = bsdf_parts(t, 0.0, phi0)[0]
        ctd = max(math.cos(0.5 * abs(t)), 1e-3)
        L.or_azimuthal(ctd, math.cos(phi0), d3(sigma), az)
        n = np.array(list(az))
        want = gauss(be[0], t - al[0]) * n[0] + gauss(be[1], t - al[1]) * n[1:4] + gauss(be[2], t - al[2]) * n[4:7]
        # (the oracle's constants are the shader's f32 literals: ~1.4e-8 relative)
        assert np.allclose(single, want, rtol=1e-7, atol=1e-300)


def test_bsdf_is_direct_local_plus_global():
    out = (C.c_double * 3)()
    l = (0.3, 0.5, 0.81)
    for depth in (0.0, 5e-4, 1.3):
        L.or_hair_bsdf(d3(T_AXIS), d3((0.2, 0.1, 0.97)), d3(l), d3(ALBEDO), ROUGH, TILT, depth, out)
        p = (C.c_double * 10)()
        L.or_bsdf_parts(d3(T_AXIS), d3((0.2, 0.1, 0.97)), d3(l), d3(ALBEDO), ROUGH, TILT, depth, p)
        p = np.array(list(p))
        want = math.exp(-depth) * (p[0:3] + 0.7 * p[3:6]) + p[6:9]
        assert np.allclose(list(out), want, rtol=1e-14)


# ---- deep opacity lookup (shade.wesl:246-310) ----------------------------------------------

S_DOM, NEAR, FAR = 4, 0.0, 16.0  # 16 layers: one unit of light depth per layer


def dom_params(dom_scale=0.5, flags_w=1):
    """A light view that maps world (x, y) in [-1, 1]^2 onto the map (clip = world,
    w = 1) and measures depth along +z from z = 0."""
    p = N.HmParams()
    m = np.eye(4)
    p.light_clip_from_world[:] = [float(x) for x in m.T.reshape(-1)]
    p.light_eye[:] = [0.0, 0.0, 0.0, 1.0]
    p.light_forward[:] = [0.0, 0.0, 1.0, 0.0]
    p.dom_info[:] = [float(S_DOM), float(S_DOM), NEAR, FAR]
    p.filter_params[:] = [1.0, 0.2, dom_scale, 0.05]
    p.flags[:] = [1, 1, 1, flags_w]
    return p


def texel_world(x, y, z, fx=0.0, fy=0.0):
    """World point at map position (x + 0.5 + fx, y + 0.5 + fy) in texels."""
    ux = (x + 0.5 + fx) / S_DOM
    uy = (y + 0.5 + fy) / S_DOM
    return (2.0 * ux - 1.0, 1.0 - 2.0 * uy, z)


def quantize(z):
    return L.or_quantize_depth(z, NEAR, FAR)


def make_dom(fronts: dict, counts: dict):
    dom = np.zeros(S_DOM * S_DOM * 17, np.uint32)
    dom[: S_DOM * S_DOM] = 0xFFFFFFFF
    for (x, y), z in fronts.items():
        dom[y * S_DOM + x] = quantize(z)
    for (x, y), c in counts.items():
        base = S_DOM * S_DOM + (y * S_DOM + x) * 16
        dom[base:base + len(c)] = c
    return dom


def vis(p, dom, world, beta=1.0):
    return L.or_dom_visibility(C.byref(p), dom.ctypes.data, f3(world), beta)


def test_dom_one_texel_known_counts():
    p = dom_params()
    counts = [3, 5, 8, 2]
    dom = make_dom({(1, 2): 4.0}, {(1, 2): counts})
    zf = L.or_dequantize_depth(quantize(4.0), NEAR, FAR)
    thick = (FAR - NEAR) / 16.0
    # 2.5 layers behind the front (after the 0.05-layer slack): layers 0, 1 whole, half of layer 2
    z = zf + 0.05 * thick + 2.5 * thick
    assert abs(vis(p, dom, texel_world(1, 2, z)) - math.exp(-0.05 * (3 + 5 + 0.5 * 8))) < 1e-6
    # within the slack the texel sees no opacity
    assert vis(p, dom, texel_world(1, 2, zf + 0.04 * thick)) == 1.0
    # far behind: every count
    assert abs(vis(p, dom, texel_world(1, 2, zf + 20 * thick)) - math.exp(-0.05 * sum(counts))) < 1e-6
    # an empty texel, a point off the map, a point behind the light, the DOM switched off
    assert vis(p, dom, texel_world(2, 2, z)) == 1.0
    assert vis(p, dom, (1.5, 0.0, z)) == 1.0
    q = dom_params()
    q.light_clip_from_world[15] = -1.0  # clip w <= 0
    assert vis(q, dom, texel_world(1, 2, z)) == 1.0
    assert vis(dom_params(flags_w=0), dom, texel_world(1, 2, z)) == 1.0


def test_dom_eq8_shift_by_strand_removal():
    """Eq. 8: beta < 1 moves the sample -ln(beta) * dom_scale toward the light."""
    p = dom_params(dom_scale=0.5)
    dom = make_dom({(1, 1): 2.0}, {(1, 1): [4, 6, 9, 1, 7]})
    for z in (4.2, 5.7, 6.3):
        shifted = vis(p, dom, texel_world(1, 1, z), beta=math.exp(-1.0))
        assert abs(shifted - vis(p, dom, texel_world(1, 1, z - 0.5))) < 2e-6
    # beta is clamped at 1e-4
    assert vis(p, dom, texel_world(1, 1, 9.0), beta=0.0) == vis(p, dom, texel_world(1, 1, 9.0), beta=1e-4)


def test_dom_bilinear_blend_of_four_texels():
    p = dom_params()
    z = 7.3
    fronts = {(1, 1): 2.0, (2, 1): 3.0, (1, 2): 1.5, (2, 2): 4.5}
    counts = {(1, 1): [5, 5, 5], (2, 1): [1, 2, 3, 4], (1, 2): [9, 0, 9], (2, 2): [2, 2]}
    dom = make_dom(fronts, counts)
    v = {k: vis(p, dom, texel_world(*k, z)) for k in fronts}
    assert len(set(v.values())) == 4
    for tx, ty in ((0.25, 0.75), (0.5, 0.5), (0.9, 0.1)):
        want = (1 - ty) * ((1 - tx) * v[(1, 1)] + tx * v[(2, 1)]) + ty * ((1 - tx) * v[(1, 2)] + tx * v[(2, 2)])
        assert abs(vis(p, dom, texel_world(1, 1, z, tx, ty)) - want) < 2e-6
    # a uniform map reads the same everywhere inside it: the weights sum to one
    uni = make_dom({(x, y): 2.0 for x in range(4) for y in range(4)},
                   {(x, y): [3, 4] for x in range(4) for y in range(4)})
    vals = [vis(p, uni, texel_world(1, 1, z, a, b)) for a, b in ((0.1, 0.2), (0.6, 0.3), (0.99, 0.99))]
    assert max(vals) - min(vals) < 1e-6


# ---- lights (shade.wesl:312-327, Bevy's range window) -----------------------------------


def test_spot_mask_closed_forms():
    pos = np.array([0.2, 1.0, -0.3])
    for target in ((0.5, -0.5, 0.4), (0.0, 2.5, 0.1)):  # pointing down (flag) and up
        inner, outer = 15.0, 35.0
        Ld = scene.spot_light(pos, target, inner, outer, (1, 1, 1))
        light = plugin.point_light(Ld)
        axis = np.asarray(target) - pos
        axis /= np.linalg.norm(axis)
        assert bool(Ld["flags"] & 2) == (axis[1] < 0)
        side = np.cross(axis, [1.0, 0.0, 0.0])
        side /= np.linalg.norm(side)
        ci, co = math.cos(math.radians(inner)), math.cos(math.radians(outer))
        for ang in (0.0, 10.0, 15.0, 25.0, 35.0, 50.0):
            a = math.radians(ang)
            world = pos + 2.0 * (math.cos(a) * axis + math.sin(a) * side)
            m = L.or_spot_mask(C.byref(light), f3(world))
            want = min(max((math.cos(a) - co) / (ci - co), 0.0), 1.0) ** 2
            assert abs(m - want) < 2e-5, (target, ang, m, want)


def test_distance_attenuation_window():
    """Bevy's smooth range window: saturate(1 - (d^2 / r^2)^2)^2 / max(d^2, 1e-4)."""
    r2 = 9.0
    for d in (0.005, 0.5, 1.5, 2.9, 3.0, 4.0):
        want = max(0.0, 1.0 - (d * d / r2) ** 2) ** 2 / max(d * d, 1e-4)
        got = L.or_distance_attenuation(d * d, 1.0 / r2)
        assert abs(got - want) <= 1e-6 * max(want, 1.0), (d, got, want)


# ---- camera reconstruction (common.wgsl:212-229) ------------------------------------------


@pytest.mark.parametrize("projection", ["perspective", "perspective_forward_z", "orthographic"])
def test_reconstruct_then_project_round_trip(projection):
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, 320, 180, projection=projection)
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 64, [scene.key_light()])[0]
    rng = np.random.default_rng(9)
    out, proj = (C.c_float * 3)(), (C.c_float * 5)()
    for _ in range(40):
        px, py = int(rng.integers(0, 320)), int(rng.integers(0, 180))
        dist = float(rng.uniform(1.0, 6.0))
        L.or_reconstruct_world(C.byref(cam), px, py, dist, out)
        L.or_project_world(C.byref(cam), out, proj)
        assert abs(proj[2] - dist) <= 2e-5 * dist  # the view distance always round-trips
        if projection != "orthographic":
            # a perspective ray through the pixel centre lands back on it
            # (f32: with forward Z the ray starts at the 0.1 near plane, ~1e-2 px)
            assert abs(proj[0] - (px + 0.5)) < 5e-2 and abs(proj[1] - (py + 0.5)) < 5e-2


def test_screen_tangent_is_the_projected_direction():
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, 320, 180)
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 64, [scene.key_light()])[0]
    clip = np.array(view.clip_from_world, dtype=np.float64)

    def screen(w):
        c = clip @ np.append(w, 1.0)
        return np.array([(c[0] / c[3] * 0.5 + 0.5) * 320, (1.0 - (c[1] / c[3] * 0.5 + 0.5)) * 180])

    rng = np.random.default_rng(4)
    out = (C.c_float * 2)()
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max)
    for _ in range(20):
        w = centre + rng.uniform(-0.3, 0.3, 3)
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        L.or_screen_tangent(C.byref(cam), f3(w), f3(t), out)
        d = screen(w + t) - screen(w)
        assert np.allclose(list(out), d / np.linalg.norm(d), atol=2e-4)
    # a tangent along the view ray has no screen direction: (1, 0)
    eye = np.array(view.world_from_view)[:3, 3]
    along = (centre - eye) / np.linalg.norm(centre - eye)
    L.or_screen_tangent(C.byref(cam), f3(centre), f3(along * 1e-4), out)
    assert list(out) == [1.0, 0.0]


# ---- per-pixel set-up (shade.wesl:539-571) and direct_light (shade.wesl:429-521) ---------


def test_pixel_inputs_uvw_mapping_and_view():
    p = screen_params(64, 48)
    p.appearance[:] = [0.3, 0.08, 3.0, 1.0]
    p.albedo[:] = [0.4, 0.2, 0.1, 1.0]
    out = (C.c_float * 16)()
    tang = (C.c_float * 3)()
    for w_code in (0, 17, 63):
        word = L.or_pack_gb(0x400000, 200, 60, 5, 9, w_code, 40)
        L.or_pixel_inputs(C.byref(p), word, (0x400000 << 8) | 128, 10, 20, out)
        o = np.array(list(out))
        w = w_code / 63.0
        assert abs(o[12] - max(0.3 * (0.65 + 0.7 * w), 0.04)) < 1e-7
        assert np.allclose(o[9:12], np.array([0.4, 0.2, 0.1]) * (0.65 + 0.35 * w), rtol=1e-6)
        assert abs(o[13] - 128 / 255.0) < 1e-7
        L.or_oct_decode(200, 60, tang)
        assert np.allclose(o[3:6], list(tang), atol=0)
        # screen_params' projection has the orthographic last row: the view
        # vector is clip_from_world's z row (shade.wesl:562-567), here -z
        assert o[15] == 1.0
        assert np.allclose(o[6:9], [0.0, 0.0, -1.0], atol=1e-7)
    # a perspective camera: the view vector points from the pixel's world point to the eye
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, 32, 24)
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 8, [scene.key_light()])[0]
    L.or_pixel_inputs(C.byref(cam), L.or_pack_gb(0x820000, 200, 60, 5, 9, 30, 40), 0, 7, 19, out)
    o = np.array(list(out))
    assert o[15] == 0.0
    v = np.array(cam.camera_pos[:3]) - o[0:3]
    assert np.allclose(o[6:9], v / np.linalg.norm(v), atol=1e-6)
    # roughness is floored at 0.04
    p.appearance[0] = 0.01
    L.or_pixel_inputs(C.byref(p), L.or_pack_gb(0x400000, 200, 60, 5, 9, 0, 40), 0, 10, 20, out)
    assert out[12] == pytest.approx(0.04)


def test_direct_light_composition():
    """One covered pixel shaded by or_shade equals the sum the reference's
    light loops form from the pinned pieces: point lights times their colour
    and range window, spot lights times their cone, directional lights times
    their colour, the first shadow caster with the DOM optical depth
    -ln(max(visibility, 1e-6)) and every other light with none."""
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, 32, 24)
    dirs = [{"direction_to_light": [0.1, 0.9, 0.42, 0.0], "color": [0.5, 0.6, 0.7, 1.0], "flags": 0},
            {"direction_to_light": [0.6, 0.5, 0.62, 0.0], "color": [2.0, 1.8, 1.6, 1.0], "flags": 1},
            {"direction_to_light": [-0.7, 0.2, 0.68, 0.0], "color": [0.4, 0.3, 0.2, 1.0], "flags": 1}]
    for d in dirs:
        v = np.array(d["direction_to_light"][:3])
        d["direction_to_light"][:3] = (v / np.linalg.norm(v)).tolist()
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 8, dirs)[0]
    x, y, S = 16, 12, 8
    code = 0x820000  # a view distance inside the groom's light-space depth range
    word = L.or_pack_gb(code, 180, 90, 20, 30, 40, 50)
    center = np.full(32 * 24, EMPTY64, np.uint64)
    center[y * 32 + x] = word
    beta = np.full(32 * 24, 0xFFFFFFFF, np.uint32)
    beta[y * 32 + x] = (code << 8) | 200
    # a DOM whose every texel holds hair from the light's near plane on
    dom = np.zeros(S * S * 17, np.uint32)
    dom[: S * S] = 0
    dom[S * S:] = np.tile(np.array([40, 30, 20, 10] + [0] * 12, np.uint32), S * S)
    inp = (C.c_float * 16)()
    L.or_pixel_inputs(C.byref(cam), word, int(beta[y * 32 + x]), x, y, inp)
    inp = np.array(list(inp))
    world = inp[0:3]
    centre = 0.5 * (mesh.bounds_min + mesh.bounds_max)
    pts = [{"position_radius": [*(centre + [0.4, 0.8, 0.3]).tolist(), 0.0],
            "color_inverse_square_range": [0.9, 0.7, 0.5, 1.0 / 9.0], "light_custom_data": [0, 0, 0, 0],
            "flags": 0, "kind": 0},
           scene.spot_light(world + np.array([0.2, 1.0, 0.4]), world + np.array([0.05, 0.0, 0.0]), 10.0, 30.0,
                            (3.0, 2.0, 1.0))]
    pa, n_p, da, n_d = plugin.light_arrays(pts, dirs)
    shaded = np.zeros((32 * 24, 4), np.float32)
    sdepth = np.zeros(32 * 24, np.float32)
    cam_dom = N.HmParams()
    C.memmove(C.byref(cam_dom), C.byref(cam), C.sizeof(N.HmParams))
    cam_dom.dom_info[0] = cam_dom.dom_info[1] = float(S)
    L.or_shade(C.byref(cam_dom), center.ctypes.data, beta.ctypes.data, dom.ctypes.data, C.cast(pa, C.c_void_p), n_p,
               C.cast(da, C.c_void_p), n_d, None, 0, None, 1.5, shaded.ctypes.data, sdepth.ctypes.data, 0, 1, 0, 24)
    vis = L.or_dom_visibility(C.byref(cam_dom), dom.ctypes.data, f3(world), float(inp[13]))
    assert 0.0 < vis < 0.999, "the key light must see some opacity"
    key_depth = -math.log(max(vis, 1e-6))
    tangent, viewv, albedo, rough = inp[3:6], inp[6:9], inp[9:12], float(inp[12])
    tilt = float(cam.appearance[1])
    out = (C.c_double * 3)()

    def bsdf(l, depth):
        L.or_hair_bsdf(d3(tangent), d3(viewv), d3(l), d3(albedo), rough, tilt, depth, out)
        return np.array(list(out))

    want = np.zeros(3)
    for Ld in pts:
        lp = np.array(Ld["position_radius"][:3], np.float64)
        to = lp - world
        att = L.or_distance_attenuation(float(to @ to), Ld["color_inverse_square_range"][3])
        c = bsdf(to / np.linalg.norm(to), 0.0) * np.array(Ld["color_inverse_square_range"][:3]) * att
        if Ld["kind"] == 1:
            c = c * L.or_spot_mask(C.byref(plugin.point_light(Ld)), f3(world))
        want += c
    for i, d in enumerate(dirs):
        depth = key_depth if i == 1 else 0.0  # the first shadow caster only
        want += bsdf(d["direction_to_light"][:3], depth) * np.array(d["color"][:3])
    want *= 1.5
    got = shaded[y * 32 + x, :3]
    assert np.allclose(got, want, rtol=2e-4), (got, want)
    assert shaded[y * 32 + x, 3] == 1.0 and sdepth[y * 32 + x] == pytest.approx(inp[14])
    assert (shaded[np.arange(32 * 24) != y * 32 + x] == 0).all()


# ---- reconnection filter (filter.wgsl:20-135) -----------------------------------------------


def filter_case(words_colors, width=16, height=12, cons_only=()):
    """A frame with centre words (and their colours) at given pixels and
    conservative-only words at others; screen_params' camera, tangent +x."""
    f = oracle.Frame(None, width, height, 8)
    f.center[:] = EMPTY64
    f.cons[:] = EMPTY64
    f.shaded[:] = 0.0
    for (x, y), (depth, color) in words_colors.items():
        w = L.or_pack_gb(depth, 255, 128, 0, 0, 0, 63)
        f.center[y * width + x] = w
        f.cons[y * width + x] = w
        f.shaded[y * width + x] = [*color, 1.0]
    for (x, y), depth in dict(cons_only).items():
        f.cons[y * width + x] = L.or_pack_gb(depth, 255, 128, 0, 0, 0, 63)
    f.filter(screen_params(width, height))
    return f


def test_filter_two_taps_hand_weights():
    """Centre pixel and one neighbour (dx, dy): weights 1 and
    exp(-dpar^2 / (5 s_par)^2 - dperp^2 / (5 s_perp)^2) exp(-|c0 - c1|^2 / 0.9^2),
    s_par = 1, s_perp = 0.2, screen tangent (1, 0)."""
    c0, c1 = np.array([0.6, 0.3, 0.1]), np.array([0.2, 0.5, 0.4])
    p = screen_params(16, 12)
    inp, tan = (C.c_float * 16)(), (C.c_float * 2)()
    L.or_pixel_inputs(C.byref(p), L.or_pack_gb(1000, 255, 128, 0, 0, 0, 63), 0, 8, 6, inp)
    L.or_screen_tangent(C.byref(p), f3(inp[0:3]), f3(inp[3:6]), tan)
    t = np.array(list(tan), np.float64)
    assert abs(t[0] - 1.0) < 1e-4 and abs(t[1]) < 1e-2  # oct code (255, 128) is +x within a code step
    for dx, dy in ((2, 0), (-3, 0), (1, 1), (0, -1)):
        f = filter_case({(8, 6): (1000, c0), (8 + dx, 6 + dy): (1000, c1)})
        d = np.array([dx, dy], np.float64)
        d_par = float(d @ t)
        d_perp = float(np.linalg.norm(d - t * d_par))
        w1 = math.exp(-d_par**2 / 25.0 - d_perp**2 / 1.0) * math.exp(-float(((c0 - c1) ** 2).sum()) / 0.81)
        want = (c0 + w1 * c1) / (1.0 + w1)
        got = f.filtered[6 * 16 + 8]
        assert np.allclose(got[:3], want, rtol=2e-6), (dx, dy, got, want)
        assert f.filtered_depth[6 * 16 + 8] > 0


def test_filter_depth_rejection_promotion_and_closer_rule():
    c0, c1 = np.array([0.6, 0.3, 0.1]), np.array([0.2, 0.5, 0.4])
    base = 0x200000
    step = int(1.45e-3 * 16777214)  # the rejection threshold in depth codes
    far = filter_case({(8, 6): (base, c0), (10, 6): (base + step + 2, c1)})
    near = filter_case({(8, 6): (base, c0), (10, 6): (base + step - 2, c1)})
    assert np.allclose(far.filtered[6 * 16 + 8, :3], c0, rtol=1e-7)  # neighbour rejected
    assert not np.allclose(near.filtered[6 * 16 + 8, :3], c0, rtol=1e-3)  # neighbour kept
    # a conservative-only pixel is promoted: no colour weight, spatial weight only
    f = filter_case({(9, 6): (base, c0), (11, 6): (base, c1)}, cons_only={(10, 6): base})
    wa = math.exp(-1.0 / 25.0)
    want = (wa * c0 + wa * c1) / (2.0 * wa)
    assert np.allclose(f.filtered[6 * 16 + 10, :3], want, rtol=1e-6)
    assert 0.0 < f.filtered[6 * 16 + 10, 3] < 1.0
    # promoted with no centre neighbour in range: alpha 0, colour 0
    lone = filter_case({(1, 1): (base, c0)}, cons_only={(12, 9): base})
    assert (lone.filtered[9 * 16 + 12] == 0).all()
    # the pixel's conservative word nearer than its centre word by >= 5e-4: alpha 1
    closer_eps = int(6e-4 * 16777214)
    f = filter_case({(8, 6): (base + closer_eps, c0)}, cons_only={(9, 6): base})
    f2 = oracle.Frame(None, 16, 12, 8)
    f2.center[:], f2.cons[:], f2.shaded[:] = f.center, f.cons, f.shaded
    f2.cons[6 * 16 + 8] = L.or_pack_gb(base, 255, 128, 0, 0, 0, 63)
    f2.filter(screen_params(16, 12))
    assert f2.filtered[6 * 16 + 8, 3] == 1.0
    assert f.filtered[6 * 16 + 8, 3] < 1.0


# ---- shadow receiver (hair_shadow.wesl:30-75) ---------------------------------------------


def receiver_light(has_map=1.0, extinction=0.05):
    """The light view of dom_params() as a ShadowLight: near 0, far 16."""
    L = N.HmShadowLight()
    L.light_clip_from_world[:] = [float(x) for x in np.eye(4).T.reshape(-1)]
    L.light_eye[:] = [0.0, 0.0, 0.0, 1.0]
    L.light_forward[:] = [0.0, 0.0, 1.0, 0.0]
    L.depth[:] = [NEAR, FAR, extinction, has_map]
    return L


def rvis(light, dom, world):
    return L.or_hair_visibility(C.byref(light), dom.ctypes.data, S_DOM, f3(world))


def test_receiver_known_counts_without_slack():
    """A surface sits behind every strand in front of it: no 0.05-layer slack
    and no Eq. 8 shift; hits = sum count_k clamp(t - k, 0, 1)."""
    counts = [3, 5, 8, 2]
    dom = make_dom({(1, 2): 4.0}, {(1, 2): counts})
    zf = L.or_dequantize_depth(quantize(4.0), NEAR, FAR)
    thick = (FAR - NEAR) / 16.0
    Lr = receiver_light(extinction=0.07)
    for t, hits in ((0.0, 0.0), (0.25, 0.75), (1.5, 3 + 2.5), (2.5, 3 + 5 + 4), (9.0, 18)):
        v = rvis(Lr, dom, texel_world(1, 2, zf + t * thick))
        assert abs(v - math.exp(-0.07 * hits)) < 2e-6, (t, v)
    # no map yet, empty texel, off the map, behind the light
    assert rvis(receiver_light(has_map=0.0), dom, texel_world(1, 2, zf + 3 * thick)) == 1.0
    assert rvis(Lr, dom, texel_world(2, 2, zf + 3 * thick)) == 1.0
    assert rvis(Lr, dom, (1.2, 0.0, 5.0)) == 1.0
    q = receiver_light()
    q.light_clip_from_world[15] = -1.0
    assert rvis(q, dom, texel_world(1, 2, zf + 3 * thick)) == 1.0


def test_receiver_is_the_shade_lookup_without_slack_and_shift():
    """With beta = 1 and the same extinction the shade's dom_visibility at
    depth d + 0.05 layer equals the receiver at depth d (to rounding), over a
    bilinear neighbourhood of four distinct texels."""
    p = dom_params()
    fronts = {(1, 1): 2.0, (2, 1): 3.0, (1, 2): 1.5, (2, 2): 4.5}
    counts = {(1, 1): [5, 5, 5], (2, 1): [1, 2, 3, 4], (1, 2): [9, 0, 9], (2, 2): [2, 2]}
    dom = make_dom(fronts, counts)
    Lr = receiver_light(extinction=0.05)
    thick = (FAR - NEAR) / 16.0
    for z in (4.8, 6.1, 7.3):
        for tx, ty in ((0.25, 0.75), (0.6, 0.1)):
            r = rvis(Lr, dom, texel_world(1, 1, z, tx, ty))
            s = vis(p, dom, texel_world(1, 1, z + 0.05 * thick, tx, ty))
            assert abs(r - s) < 5e-6 and r < 1.0


# ---- composite (composite.wgsl:27-46, pipeline.rs:299-339) ---------------------------------


def composite_case(reverse: bool, colors, view_dists, target_depth, target_color):
    mesh = synthetic_groom("straight", 64, 1)
    view = scene.fit_camera(mesh, 8, 4, projection="perspective" if reverse else "perspective_forward_z")
    cam = plugin.Renderer.make_params(view, plugin.HairGroom(), mesh, 8, [scene.key_light()])[0]
    color = np.ascontiguousarray(colors, np.float32).reshape(32, 4)
    depth = np.ascontiguousarray(view_dists, np.float32).reshape(32)
    tc = np.ascontiguousarray(target_color, np.float32).reshape(32, 4).copy()
    td = np.ascontiguousarray(target_depth, np.float32).reshape(32).copy()
    L.or_composite(C.byref(cam), color.ctypes.data, depth.ctypes.data, tc.ctypes.data, td.ctypes.data)
    return cam, tc, td


def test_composite_depth_blend_and_discard():
    rng = np.random.default_rng(8)
    colors = rng.uniform(0.0, 1.0, (32, 4)).astype(np.float32)
    colors[:, 3] = rng.uniform(0.2, 1.0, 32)
    colors[3, 3] = 0.0009  # below the discard threshold
    dists = rng.uniform(1.0, 9.0, 32).astype(np.float32)
    dst = rng.uniform(0.0, 1.0, (32, 4)).astype(np.float32)
    # reverse Z (Bevy's infinite reverse perspective): clip depth = near / view distance
    cam, tc, td = composite_case(True, colors, dists, np.full(32, 0.0, np.float32), dst)
    near = 0.1  # the projection's near plane
    want_d = near / dists.astype(np.float64)
    hit = np.arange(32) != 3
    assert np.allclose(td[hit], want_d[hit], rtol=2e-5)
    a = colors[:, 3:4].astype(np.float64)
    want_rgb = colors[:, :3] * a + dst[:, :3] * (1.0 - a)
    want_a = colors[:, 3] + dst[:, 3] * (1.0 - colors[:, 3])
    assert np.allclose(tc[hit, :3], want_rgb[hit], rtol=1e-6, atol=1e-7)
    assert np.allclose(tc[hit, 3], want_a[hit], rtol=1e-6)
    assert td[3] == 0.0 and (tc[3] == dst[3]).all()  # discarded
    # a nearer stored depth (larger, reverse Z) keeps the target; equal passes
    stored = want_d.astype(np.float32).copy()
    stored[0::2] = np.minimum(1.0, stored[0::2] * 1.01)
    _, tc2, td2 = composite_case(True, colors, dists, stored, dst)
    assert (tc2[0::2] == dst[0::2]).all() and (td2[0::2] == stored[0::2]).all()
    assert not np.allclose(tc2[1::2], dst[1::2])
    # forward Z compares <=: a nearer (smaller) stored depth keeps the target
    _, tcf, tdf = composite_case(False, colors, dists, np.full(32, 1.0, np.float32), dst)
    assert (tdf[hit] < 1.0).all() and (tdf[hit] > 0.0).all()
    _, tcf2, _ = composite_case(False, colors, dists, np.zeros(32, np.float32), dst)
    assert (tcf2 == dst).all()
