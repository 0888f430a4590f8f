"""Particle-pair drivers on the GPU against the oracle: every contact field
bit for bit, the emission order, and the statistics."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import scenes
from delta_amd import _lib, contact, shapes
from delta_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

METHODS = {
    "bf": (contact.bf, _lib.DH_BF),
    "penalty": (contact.penalty, _lib.DH_PENALTY),
    "hybrid_batch": (contact.hybridWithPerBatchFallBack, _lib.DH_HYBRID_BATCH),
    "hybrid_triangle": (contact.hybridWithPerTriangleFallBack, _lib.DH_HYBRID_TRIANGLE),
    "penalty_stat": (contact.penaltyStat, _lib.DH_PENALTY_STAT),
    "hybrid_batch_stat": (contact.hybridBatchStat, _lib.DH_HYBRID_BATCH_STAT),
    "hybrid_triangle_stat": (contact.hybridTriangleStat, _lib.DH_HYBRID_TRIANGLE_STAT),
    "gjk": (lambda xA, yA, zA, nA, eA, fA, iA, xB, yB, zB, nB, eB, fB, iB:
            contact.gjk(nA, xA, yA, zA, eA, fA, iA, nB, xB, yB, zB, eB, fB, iB), _lib.DH_GJK),
}
FIELDS = ["x", "normal", "depth", "P", "Q", "eps_total", "master", "slave", "penetration",
          "friction", "triA", "triB"]
EPS = 1e-4


def same_contacts(a, b):
    assert len(a) == len(b), (len(a), len(b))
    for f in FIELDS:
        va, vb = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        if va.dtype == np.float64:
            assert np.array_equal(va.view(np.uint64), vb.view(np.uint64)), f
        else:
            assert np.array_equal(va, vb), f


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("case", range(len(scenes.pair_cases())))
def test_pair_driver_matches_oracle(method, case):
    name, A, B = scenes.pair_cases()[case]
    fn, code = METHODS[method]
    if code == _lib.DH_HYBRID_TRIANGLE and len(B) > len(A):
        A, B = B, A
    xa, xb = shapes.to_soa(A), shapes.to_soa(B)
    contact.reset_all_statistics()
    got = fn(*xa, len(A), EPS, True, 7, *xb, len(B), EPS, True, 9)
    st = contact._stats
    rc, want, ost = ol.or_detect(code, A, B, EPS, EPS, 1, 1, 7, 9)
    assert rc == 0
    same_contacts(got, want)
    assert list(st.newton_histogram) == list(ost.newton_histogram)
    for f in ("penalty_fails", "batch_fails", "batch_size", "triangle_checks",
              "pairs_evaluated", "fallback_rows"):
        assert getattr(st, f) == getattr(ost, f), f
    assert st.batch_error == ost.batch_error


def test_hybrid_triangle_rejects_reference_out_of_bounds():
    A = scenes.particle(0, 0.5e-3, 1, [0, 0, 0])
    B = scenes.particle(1, 0.5e-3, 2, [1e-3, 0, 0])
    with pytest.raises(_lib.DhError) as e:
        contact.hybridWithPerTriangleFallBack(*shapes.to_soa(A), len(A), EPS, 0, 0,
                                              *shapes.to_soa(B), len(B), EPS, 0, 1)
    assert e.value.code == 4


@pytest.mark.parametrize("method", ["bf", "penalty", "hybrid_batch", "hybrid_batch_stat", "gjk"])
def test_batched_pairs_match_per_pair_oracle(method):
    """dh_detect_pairs over several particle pairs: contacts grouped by pair
    in pair order, each group in reference order."""
    _batched_pairs(method)


@pytest.mark.parametrize("screen", [0, 1])
@pytest.mark.parametrize("max_tiles", [1, 16])
def test_batched_bf_through_the_survivor_queue(max_tiles, screen):
    """The same with the survivor queue forced on (every item's survivors
    solved from one queue, one per lane), with and without the binary32
    screen: bit-exact and in order."""
    from delta_amd import tuning
    tuning.set(detect_queue_min_items=0, detect_max_tiles=max_tiles, detect_fp32_screen=screen)
    try:
        _batched_pairs("bf")
    finally:
        tuning.set(**tuning.default())


def _batched_pairs(method):
    parts = [scenes.particle(1 + (i % 2), 0.5e-3, 50 + i, [0.98e-3 * i, 0.05e-3 * (i % 3), 0])
             for i in range(6)]
    parts.append(scenes.floor(-0.45e-3))
    pairs = [(i, i + 1) for i in range(5)] + [(i, 6) for i in range(6)] + [(0, 2)]
    offs = np.cumsum([0] + [len(p) for p in parts]).astype(np.int32)
    allm = np.concatenate(parts)
    x, y, z = (DeviceArray.from_host(v) for v in shapes.to_soa(allm))
    d_off = DeviceArray.from_host(offs)
    n = len(parts)
    d_eps = DeviceArray.from_host(np.full(n, EPS))
    d_fr = DeviceArray.from_host(np.ones(n, np.int32))
    d_id = DeviceArray.from_host(np.arange(n, dtype=np.int32) * 10)
    d_pairs = DeviceArray.from_host(np.array(pairs, np.int32).reshape(-1))
    g = _lib.DhGeometry(x.ptr, y.ptr, z.ptr, d_off.ptr, d_eps.ptr, d_fr.ptr, d_id.ptr, n,
                        int(offs[-1]))
    code = METHODS[method][1]
    cap = 200000
    out = DeviceArray((cap * 144,), np.uint8)
    cnt = DeviceArray((len(pairs),), np.int32)
    nout = C.c_int(0)
    st = _lib.DhStats()
    L = _lib.lib()
    L.dh_detect_pairs.argtypes = [C.c_int, C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p,
                                  C.POINTER(_lib.DhStats), C.c_void_p]
    _lib.check(L.dh_detect_pairs(code, C.byref(g), d_pairs.ptr, len(pairs), out.ptr, cap,
                                 C.byref(nout), cnt.ptr, C.byref(st), None), "detect_pairs")
    got = np.frombuffer(out.download().tobytes()[:144 * nout.value], dtype=contact.CONTACT_DTYPE)
    counts = cnt.download()
    want = []
    ost_total = ol.DhStats()
    for k, (a, b) in enumerate(pairs):
        rc, w, ost = ol.or_detect(code, parts[a], parts[b], EPS, EPS, 1, 1, 10 * a, 10 * b)
        assert rc == 0
        assert counts[k] == len(w)
        w = w.copy()
        w["pair"] = k
        want.append(w)
        for f in ("penalty_fails", "batch_fails", "batch_size", "triangle_checks",
                  "pairs_evaluated", "fallback_rows"):
            setattr(ost_total, f, getattr(ost_total, f) + getattr(ost, f))
    want = np.concatenate(want)
    same_contacts(got, want)
    assert np.array_equal(got["pair"], want["pair"])
    for f in ("penalty_fails", "batch_fails", "batch_size", "triangle_checks",
              "pairs_evaluated", "fallback_rows"):
        assert getattr(st, f) == getattr(ost_total, f), f


def _wall_touching_particle():
    """The 2,048-triangle funnel and an 80-triangle sphere pressed 0.05 mm
    into its wall at mid height."""
    from delta_amd import workloads
    funnel = shapes.cone_funnel(2e-3, 20e-3, 30.0)
    t = np.tan(np.radians(30.0))
    y = 15e-3
    inward = -np.array([1.0, -t, 0.0]) / np.hypot(1.0, t)
    c = np.array([2e-3 + y * t, y, 0.0]) + inward * 0.45e-3
    return funnel, scenes.particle(1, 0.5e-3, 61, c)


@pytest.mark.parametrize("queue", [0, 1])
@pytest.mark.parametrize("max_tiles", [1, 16])
@pytest.mark.parametrize("swap", [False, True])
def test_bf_large_particle_matches_oracle(max_tiles, swap, queue):
    """A particle with more rows than one LDS block holds (the funnel as A:
    2,048 rows, staged in blocks of 320 when one workgroup takes them all)
    and as B (2,048 columns: the unstaged pass): contacts bit-exact and in
    the reference's order."""
    from delta_amd import tuning
    A, B = _wall_touching_particle()
    if swap:
        A, B = B, A
    tuning.set(detect_max_tiles=max_tiles, detect_queue_min_items=0 if queue else 1 << 30)
    try:
        got = contact.bf(*shapes.to_soa(A), len(A), EPS, True, 7, *shapes.to_soa(B), len(B),
                         EPS, True, 9)
    finally:
        tuning.set(**tuning.default())
    rc, want, _ = ol.or_detect(_lib.DH_BF, A, B, EPS, EPS, 1, 1, 7, 9)
    assert rc == 0 and len(want) > 0
    same_contacts(got, want)
