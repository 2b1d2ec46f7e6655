"""hm_scan_noalloc (the allocation-free scan) against hm_scan through the C
ABI: two calls in a row on one GPU, the second with fewer bundles and a
capacity below its strand total, each bit for bit equal to hm_scan's refs and
indirect words for the same LOD records; more bundles than the scratch holds
are refused before anything runs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from bevy_hair_amd import _native as N
from tests.scenes import make_scene

pytestmark = pytest.mark.gpu


def _params_with_bundles(params, nb):
    p = N.HmParams()
    C.memmove(C.byref(p), C.byref(params), C.sizeof(N.HmParams))
    p.pass_mode[1] = nb
    return p


def test_scan_noalloc_matches_hm_scan(renderer):
    r = renderer
    # (workload, bundles kept, capacity cut): the second call is the smaller input
    calls = [("small_lod", None, None), ("tiny", 2, 150)]
    for name, keep_bundles, cap_cut in calls:
        sc = make_scene(name)
        w = sc.w
        print(f"\n[gpu no-alloc] {name} bundles={keep_bundles} capacity={cap_cut}", flush=True)
        g = r.upload(sc.mesh)
        try:
            nb = g.bundle_count if keep_bundles is None else keep_bundles
            cam = _params_with_bundles(sc.params[0], nb)
            r.lod(g, cam, N.PASS_CAMERA, w.control_points, w.strand_fraction)
            r.scan(g, cam, N.PASS_CAMERA)
            exp_refs, exp_total = r.refs(g, N.PASS_CAMERA)
            exp_ind = r.read(N.BUF_CAMERA_INDIRECT, np.uint32, 4, g)
            lod_ptr, lod_bytes = r.device_ptr(N.BUF_CAMERA_LOD, g)
            refs_ptr, refs_bytes = r.device_ptr(N.BUF_DOM_REFS, g)
            ind_ptr, _ = r.device_ptr(N.BUF_DOM_INDIRECT, g)
            assert nb * 16 <= lod_bytes
            capacity = refs_bytes // 16 if cap_cut is None else cap_cut
            assert capacity * 16 <= refs_bytes
            r.scan_noalloc(lod_ptr, nb, refs_ptr, capacity, ind_ptr)
            ind = r.read(N.BUF_DOM_INDIRECT, np.uint32, 4, g)
            total = int(ind[0])
            refs = r.read(N.BUF_DOM_REFS, np.uint32, g.mesh.strand_count * 4, g).reshape(-1, 4)[:total]
            # scan.wgsl:14-16 caps the list at the buffer's length
            want = min(exp_total, capacity)
            assert total == want, (total, exp_total, capacity)
            assert ind[1] == exp_ind[1] == 1 and ind[2] == exp_ind[2] == 1
            assert np.array_equal(refs, exp_refs[:want]), f"{name}: refs differ from hm_scan"
            if cap_cut is not None:
                assert exp_total > capacity, "the second call must hit the capacity cap"
        finally:
            g.free()


def test_scan_noalloc_rejects_too_many_bundles(renderer):
    st = N.lib().hm_scan_noalloc(renderer.ctx, C.c_void_p(16), N.SCAN_NOALLOC_MAX_BUNDLES + 1, C.c_void_p(16), 1,
                                 C.c_void_p(16))
    assert st == 6  # HM_ERR_CAPACITY, before any launch (the pointers are never touched)
    assert b"exceed" in N.lib().hm_last_error(renderer.ctx)
