"""Shared test helpers: the same seeded inputs for the HIP path and the oracle."""
from __future__ import annotations

import numpy as np

from bevy_hair_amd.workload import Scene, make_scene  # noqa: F401

EMPTY64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def rel_err(got: np.ndarray, ref: np.ndarray, floor: float = 1e-3) -> np.ndarray:
    """|got - ref| / max(|ref|, floor): relative error with an absolute floor."""
    g, r = got.astype(np.float64), ref.astype(np.float64)
    return np.abs(g - r) / np.maximum(np.abs(r), floor)
