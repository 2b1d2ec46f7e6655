"""The shading kernels' constant tables against the formulas they tabulate.

dual_scatter (shade.wesl:153-190) evaluates azimuthal at 16 fixed angles:
i < 8 backward, phi = (i + 0.5) pi / 16; i >= 8 forward,
phi = pi / 2 + (i - 7.5) pi / 16. hm_shade.h holds, per angle, cos(phi),
cos(phi / 2) = sqrt(0.5 + 0.5 cos phi), 0.6 - 0.8 cos phi and the TT / TRT
azimuthal weights exp(-3.65 cos phi - 3.98), exp(17 cos phi - 16.78)
(shade.wesl:94-116), and the sums of the TRT weight over each half.
"""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "bevy_hair_amd", "csrc", "hm_shade.h")


def _table():
    text = open(HDR).read()
    body = text[text.index("kDualAngles[16] = {"):]
    body = body[:body.index("};")]
    rows = re.findall(r"\{([^{}]+)\}", body)
    vals = np.array([[float(v.strip().rstrip("f")) for v in r.split(",")] for r in rows])
    back = float(re.search(r"kDualTrtBack = ([0-9.e+-]+)f", text).group(1))
    fwd = float(re.search(r"kDualTrtFwd = ([0-9.e+-]+)f", text).group(1))
    return vals, back, fwd


def test_dual_angles_match_formulas():
    vals, back, fwd = _table()
    assert vals.shape == (16, 5)
    i = np.arange(16)
    u = ((i & 7) + 0.5) / 8.0
    phi = np.where(i < 8, u * 0.5 * np.pi, (1.0 + u) * 0.5 * np.pi)
    c = np.cos(phi)
    want = np.stack([c, np.sqrt(0.5 + 0.5 * c), 0.6 - 0.8 * c, np.exp(-3.65 * c - 3.98), np.exp(17.0 * c - 16.78)], 1)
    # f32 rounding of each entry (the angle itself is rounded to f32 as the
    # reference computes it, which moves cos by at most ~1e-7 absolute)
    assert np.allclose(vals, want, rtol=2e-6, atol=2e-7)
    assert abs(back - want[:8, 4].sum()) <= 1e-6 * back
    assert abs(fwd - want[8:, 4].sum()) <= 1e-5 * fwd


def test_dual_angle_order_is_back_then_forward():
    vals, _, _ = _table()
    c = vals[:, 0]
    assert (c[:8] > 0).all() and (c[8:] < 0).all()
    assert (np.diff(c) < 0).all()  # phi increases from 0 to pi
