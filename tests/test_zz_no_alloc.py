"""GPU test of trans_geom_wall_distance_static, the wall distance that allocates no device
memory (it stages cells, panels and minima in a static device array): called through the
C-ABI twice in a row, the second time with a smaller input, it gives bitwise the result of
trans_geom_wall_distance; an input beyond the array's size is rejected with TRANS_ESCRATCH
before the device is touched."""
import ctypes as C

import numpy as np
import pytest

from trans_amd import capi

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)


def _call(fn, cells, panels):
    c = np.ascontiguousarray(cells.T)
    p = np.ascontiguousarray(panels.T)
    out = np.empty(c.shape[1])
    rc = fn(c.ctypes.data_as(DP), c.shape[1], p.ctypes.data_as(DP), p.shape[1], out.ctypes.data_as(DP))
    return rc, out


def test_static_wall_distance_matches_allocating_entry():
    lib = capi.load_library()
    rng = np.random.Generator(np.random.MT19937(11))
    for n, m in ((60000, 5000), (7001, 613)):  # second call smaller: stale staging data must not leak in
        print(f"case: static wall distance {n} cells x {m} panels", flush=True)
        cells = rng.uniform(-1, 1, (n, 3))
        panels = rng.uniform(-1, 1, (m, 3))
        rc_s, got = _call(lib.trans_geom_wall_distance_static, cells, panels)
        rc_a, ref = _call(lib.trans_geom_wall_distance, cells, panels)
        assert rc_s == 0 and rc_a == 0
        assert np.array_equal(got, ref)


def test_static_wall_distance_rejects_oversize():
    lib = capi.load_library()
    n = capi.TRANS_WD_STATIC_DOUBLES // 4 + 1
    cells = np.zeros((n, 3))
    panels = np.zeros((1, 3))
    print(f"case: static wall distance oversize {n} cells (rejected on the host)", flush=True)
    rc, _ = _call(lib.trans_geom_wall_distance_static, cells, panels)
    assert rc == capi.TRANS_ESCRATCH
