"""The step's stages one at a time on the GPU against the oracle: broad
phase candidate set, batched detection on it, contact forces, and the fused
update given identical forces.  Each stage starts from identical inputs, so a
mismatch names its stage."""
import copy
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import _lib, contact, engine
from delta_amd.device import DeviceArray
from test_gpu_engine import spinning_scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def host_candidates(hs):
    lo = np.zeros((hs.n, 3)); hi = np.zeros((hs.n, 3))
    for p in range(hs.n):
        a, b = hs.tri_offset[p], hs.tri_offset[p + 1]
        v = np.stack([hs.x[3 * a:3 * b], hs.y[3 * a:3 * b], hs.z[3 * a:3 * b]], 1)
        m = hs.eps[p] + 0.5e-9
        lo[p] = v.min(0) - m
        hi[p] = v.max(0) + m
    out = []
    for i in range(hs.n):
        j = np.arange(i + 1, hs.n)
        ok = np.all(lo[i] <= hi[j], axis=1) & np.all(lo[j] <= hi[i], axis=1)
        if hs.obstacle[i]:
            ok &= hs.obstacle[j] == 0
        out += [(i, int(k)) for k in j[ok]]
    return out


def gpu_pairs(ds, hs, cell):
    L = _lib.lib()
    L.dh_broadphase.argtypes = [C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_double, C.c_void_p,
                                C.c_int, C.POINTER(C.c_int), C.c_void_p]
    cap = hs.n * 64
    out = DeviceArray((2 * cap,), np.int32)
    n = C.c_int(0)
    _lib.check(L.dh_broadphase(C.byref(ds.scene.geom), ds.buf["obstacle"].ptr, cell, out.ptr,
                               cap, C.byref(n), None), "broadphase")
    return [tuple(p) for p in out.download()[:2 * n.value].reshape(-1, 2)], out


def _scene(name):
    if name == "spinning":
        return spinning_scene()
    if name == "c1":
        return engine.c1_scene()
    if name.startswith("drum-") or name.startswith("funnel-"):
        from delta_amd import workloads
        kind, n = name.split("-")
        return getattr(workloads, kind + "_scene")(n=int(n))
    from delta_amd import scene_io
    return scene_io.hopper_scene(num=int(name.split("-")[1]), sphere_level=1)


@pytest.mark.parametrize("scene", ["spinning", "c1", "hopper-6", "hopper-12", "drum-3000",
                                   "funnel-1500"])
def test_broadphase_equals_all_pairs(scene):
    """Exact candidate set; the hopper lattices (touching spheres in wall
    prisms, obstacles first) have thousands of particles and pairs; the
    box of drum-3000 and the funnel of funnel-1500 enclose every particle,
    so their partner lists (3,000 and 1,500 pairs) span several of the
    large-particle search chunks."""
    hs = _scene(scene)
    ds = engine.DeviceScene(hs)
    got, _ = gpu_pairs(ds, hs, hs.cell_size())
    want = host_candidates(hs)
    assert len(got) == len(want), (len(got), len(want))
    assert got == want
    # a cell smaller than the particles still yields the same set (all large)
    got2, _ = gpu_pairs(ds, hs, hs.cell_size() * 0.3)
    assert got2 == got


def test_forces_and_update_match_oracle():
    hs = spinning_scene()
    ds = engine.DeviceScene(hs)
    pairs, d_pairs = gpu_pairs(ds, hs, hs.cell_size())
    L = _lib.lib()
    # detection on the GPU
    cap = 100000
    out = DeviceArray((cap * 144,), np.uint8)
    cnt = DeviceArray((len(pairs),), np.int32)
    nout = C.c_int(0)
    st = _lib.DhStats()
    L.dh_detect_pairs.argtypes = [C.c_int, C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p,
                                  C.POINTER(_lib.DhStats), C.c_void_p]
    _lib.check(L.dh_detect_pairs(_lib.DH_BF, C.byref(ds.scene.geom), d_pairs.ptr, len(pairs),
                                 out.ptr, cap, C.byref(nout), cnt.ptr, C.byref(st), None), "det")
    counts = cnt.download()
    allc = np.frombuffer(out.download().tobytes()[:144 * nout.value], dtype=contact.CONTACT_DTYPE)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    by_pair = [allc[offs[k]:offs[k + 1]] for k in range(len(pairs))]
    # forces on the GPU
    d_offs = DeviceArray.from_host(offs)
    F = DeviceArray((hs.n, 3), np.float64)
    T = DeviceArray((hs.n, 3), np.float64)
    tch = DeviceArray((hs.n,), np.int32)
    fp = _lib.DhForceParams(1e5, 0.1, 0.5, 0)
    L.dh_contacts_forces.argtypes = [C.POINTER(_lib.DhBodies), C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.POINTER(_lib.DhForceParams), C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
    _lib.check(L.dh_contacts_forces(C.byref(ds.scene.bodies), d_pairs.ptr, len(pairs), out.ptr,
                                    d_offs.ptr, C.byref(fp), F.ptr, T.ptr, tch.ptr, None), "forces")
    fo, to, tcho = ol.or_forces_scene(hs, pairs, by_pair, 1e5, 0.1, 0.5)
    assert np.array_equal(tch.download(), tcho)
    assert np.array_equal(bits(F.download()), bits(fo)), "force"
    assert np.array_equal(bits(T.download()), bits(to)), "torque"
    # fused update on the GPU with the oracle's forces
    F.upload(fo)
    T.upload(to)
    g = (C.c_double * 3)(0.0, -9.81, 0.0)
    b = ds.buf
    L.dh_integrate.argtypes = [C.POINTER(_lib.DhBodies), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_double, C.c_double * 3] + [C.c_void_p] * 7 + [C.c_void_p]
    _lib.check(L.dh_integrate(C.byref(ds.scene.bodies), F.ptr, T.ptr, tch.ptr, 1e-6, g,
                              b["x"].ptr, b["y"].ptr, b["z"].ptr, b["rx"].ptr, b["ry"].ptr,
                              b["rz"].ptr, b["tri_offset"].ptr, None), "integrate")
    orc = ol.oracle()
    for p in range(hs.n):
        if hs.obstacle[p]:
            continue
        lv = hs.linvel[p].copy(); W = hs.refangvel[p].copy(); R = hs.rot[p].copy()
        pos = hs.pos[p].copy()
        if tcho[p]:
            lv = lv + 1e-6 * (fo[p] / hs.mass[p])
            orc.or_update_angular(ol.dptr(W), ol.dptr(R), ol.dptr(np.ascontiguousarray(hs.inertia[p])),
                                  ol.dptr(np.ascontiguousarray(hs.inverse[p])),
                                  ol.dptr(np.ascontiguousarray(to[p])), C.c_double(1e-6))
        lv[1] = lv[1] + 1e-6 * -9.81
        pos = pos + 1e-6 * lv
        w = np.zeros(3)
        orc.or_update_rotation_matrix(ol.dptr(w), ol.dptr(W), ol.dptr(R), C.c_double(1e-6))
        series = (1e-6 * W) @ (1e-6 * W) < 3.0461741978671E-02
        for name, want in (("linvel", lv), ("refangvel", W), ("rot", R), ("pos", pos), ("angvel", w)):
            got = b[name].download().reshape(hs.n, -1)[p]
            if series or name in ("linvel", "refangvel", "pos"):
                assert np.array_equal(bits(got), bits(want)), (name, p)
            else:  # expmap's sin/cos branch: device vs host libm, <= 1 ulp each
                assert np.allclose(got, want, rtol=1e-14, atol=1e-300), (name, p)


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 5000, 16383, 16384, 16385, 1 << 20,
                               3_000_001])
def test_exclusive_scan(n):
    """dh_exclusive_scan against numpy: offs[k] = sum of counts[:k] and
    offs[n] the total, for sizes around the 1,024-entry blocks, 16 blocks
    and past 1,024 blocks (the block sums are scanned 1,024 at a time)."""
    L = _lib.lib()
    L.dh_exclusive_scan.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 1000, max(n, 1)).astype(np.int32)[:n]
    d_counts = DeviceArray.from_host(counts if n else np.zeros(1, np.int32))
    offs = DeviceArray((n + 1,), np.int64)
    _lib.check(L.dh_exclusive_scan(d_counts.ptr, n, offs.ptr, None), "scan")
    want = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    assert np.array_equal(offs.download(), want)


def test_broadphase_after_eager_steps_equals_step_and_all_pairs():
    """dh_broadphase on the C1 lattice before each of 24 eager dh_step calls:
    its candidate set equals the all-pairs box test on the downloaded
    geometry, and its count equals the candidate count of the step, which
    runs the same broad phase on the same geometry.  The lattice's initial
    overlaps fling the particles apart (484 pairs at step 0, none at step 8),
    so the comparison covers full and empty candidate sets."""
    hs = engine.c1_scene()
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, dt=1e-6, spring=1e5, damper=0.1, friction=0.0,
                           cell=hs.cell_size())
    cell = hs.cell_size()
    L = _lib.lib()
    counts = []
    for s in range(24):
        _lib.check(L.dh_synchronize(), "sync")
        got, _ = gpu_pairs(ds, hs, cell)
        again, _ = gpu_pairs(ds, hs, cell)
        cur = copy.copy(hs)
        cur.x, cur.y, cur.z = ds.download("x"), ds.download("y"), ds.download("z")
        want = host_candidates(cur)
        info = engine.step(ds, p)
        assert len(got) == len(want) == info.candidate_pairs, \
            (s, len(got), len(again), len(want), info.candidate_pairs)
        assert got == want and again == want, s
        counts.append(len(want))
    assert counts[0] > 400 and 0 in counts, counts
