"""dh_broadphase_noalloc (the broad phase on the library's fixed device
array, no allocation) against dh_broadphase through the C ABI: the same pair
lists and counts on two calls in a row, the second on a smaller scene, the
all-large path, a capacity overflow, and sizes beyond the fixed array."""
import ctypes as C

import numpy as np
import pytest

from delta_amd import _lib, engine, scene_io
from delta_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

ARGT = [C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_double, C.c_void_p, C.c_int,
        C.POINTER(C.c_int), C.c_void_p]


def run(fn, ds, cell, cap):
    L = _lib.lib()
    f = getattr(L, fn)
    f.argtypes = ARGT
    f.restype = C.c_int
    out = DeviceArray((2 * max(cap, 1),), np.int32)
    _lib.check(L.dh_memset(out.ptr, 0xFF, out.nbytes), "memset")
    n = C.c_int(-1)
    rc = f(C.byref(ds.scene.geom), ds.buf["obstacle"].ptr, cell, out.ptr, cap, C.byref(n), None)
    pairs = out.download()[:2 * max(n.value, 0)].reshape(-1, 2) if rc == 0 else None
    return rc, n.value, pairs


def test_noalloc_equals_broadphase_twice_then_smaller():
    big = scene_io.hopper_scene(num=12, sphere_level=1)
    small = engine.c1_scene()
    assert small.n < big.n <= 16384
    for hs in (big, small):
        ds = engine.DeviceScene(hs)
        for cell in (hs.cell_size(), 0.3 * hs.cell_size()):
            cap = 64 * hs.n
            rc0, n0, p0 = run("dh_broadphase", ds, cell, cap)
            rc1, n1, p1 = run("dh_broadphase_noalloc", ds, cell, cap)
            assert rc0 == 0 and rc1 == 0, _lib.last_error()
            assert n0 == n1 and n0 > 0, (hs.n, n0, n1)
            assert np.array_equal(p0, p1), hs.n


def test_noalloc_capacity_and_size_limits():
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    rc0, n0, _ = run("dh_broadphase", ds, hs.cell_size(), 64 * hs.n)
    rc, n, _ = run("dh_broadphase_noalloc", ds, hs.cell_size(), 10)
    assert rc == _lib.DH_ERR_CAPACITY and n == n0 > 10
    L = _lib.lib()
    L.dh_broadphase_noalloc.argtypes = ARGT
    g = _lib.DhGeometry(None, None, None, None, None, None, None, 16385, 100)
    cnt = C.c_int(-1)
    assert L.dh_broadphase_noalloc(C.byref(g), None, 1.0, None, 0, C.byref(cnt), None) == \
        _lib.DH_ERR_CAPACITY
    g = _lib.DhGeometry(None, None, None, None, None, None, None, 100, (1 << 20) + 1)
    assert L.dh_broadphase_noalloc(C.byref(g), None, 1.0, None, 0, C.byref(cnt), None) == \
        _lib.DH_ERR_CAPACITY
    assert "exceed" in _lib.last_error()
