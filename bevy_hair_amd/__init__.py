"""MI355X-native deferred strand rasterizer (bevy_hair's hot path) behind a C ABI.

The product is libhairmi.so (HIP kernels for gfx950 + host driver, include/hair_mi.h);
this package is its Python host: ctypes binding, the HairGroom/HairMesh surface,
seeded synthetic grooms and the HAIRBAK1 groom cache.
"""
from .groom import HairMesh, synthetic_groom  # noqa: F401
from .plugin import HairGroom, Renderer  # noqa: F401
