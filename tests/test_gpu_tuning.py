"""Launch-parameter table and autotuner on the GPU: every setting of every
knob gives bit-identical results (unique pairs of the exact, mixed and
learned paths; contacts, per-pair counts and triangle checks of
particle-pair detection), and the autotuner keeps the fastest candidate it timed."""
import ctypes as C

import numpy as np
import pytest

from delta_amd import _lib, engine, learn, narrow, tuning
from delta_amd.device import DeviceArray
from test_gpu_stages import _scene, gpu_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def default_table():
    _lib.init()
    tuning.set(**tuning.default())
    yield
    tuning.set(**tuning.default())


def _results(batch, method):
    batch.run(method)
    P, Q, d, fl = batch.results()
    return P.view(np.uint64), Q.view(np.uint64), d.view(np.uint64), fl


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.fixture(scope="module")
def batch():
    _lib.init()
    b = narrow.PairBatch(200003, seed=41, with_f32=True)
    return b


@pytest.fixture(scope="module")
def model():
    _lib.init()
    b = narrow.PairBatch(1 << 16, seed=43)
    tr = learn.Trainer(seed=2)
    tr.fit(learn.Dataset(b.soa, b.n), epochs=1, batch=8192, lr=3e-3)
    tr.install()
    return tr


@pytest.mark.parametrize("method,knob,values", [
    ("bf64", "exact_block", [64, 128, 192]),
    ("hybrid64", "exact_block", [64, 192]),
    ("mixed", "mixed_wg_per_cu", [1, 3, 16, 32]),
    ("learned", "learned_wg_per_cu", [1, 5, 16, 32]),
])
def test_pair_results_identical_for_every_setting(batch, model, method, knob, values):
    want = _results(batch, method)
    for v in values:
        tuning.set(**{knob: v})
        got = _results(batch, method)
        assert _same(got, want), (method, knob, v)


def test_autotune_pairs_keeps_fastest_and_outputs_are_unchanged(batch, model):
    for method, knob, count in (("bf64", "exact_block", 4), ("mixed", "mixed_wg_per_cu", 10),
                                ("learned", "learned_wg_per_cu", 10)):
        want = _results(batch, method)
        trials = tuning.autotune_pairs(batch, method, reps=2)
        assert len(trials) == count
        assert all(t["knob"] == knob and t["ms"] > 0 for t in trials)
        best = min(trials, key=lambda t: t["ms"])
        assert tuning.get()[knob] == best["value"]
        P, Q, d, fl = batch.results()
        assert _same((P.view(np.uint64), Q.view(np.uint64), d.view(np.uint64), fl), want)


def _detect(ds, d_pairs, npairs, cap=1 << 17):
    L = _lib.lib()
    L.dh_detect_pairs.argtypes = [C.c_int, C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p,
                                  C.POINTER(_lib.DhStats), C.c_void_p]
    out = DeviceArray((cap * 144,), np.uint8)
    cnt = DeviceArray((npairs,), np.int32)
    nout = C.c_int(0)
    st = _lib.DhStats()
    _lib.check(L.dh_detect_pairs(_lib.DH_BF, C.byref(ds.scene.geom), d_pairs.ptr, npairs,
                                 out.ptr, cap, C.byref(nout), cnt.ptr, C.byref(st), None),
               "dh_detect_pairs")
    raw = out.download().tobytes()[:144 * nout.value]
    # pairs evaluated / culled depend on the tiling (the mid-phase culls
    # against the union of a tile's rows); the pairs checked do not
    return raw, cnt.download(), st.triangle_checks


@pytest.mark.parametrize("scene", ["c1", "hopper-6"])
def test_detect_tilings_identical_and_autotune(scene):
    hs = _scene(scene)
    ds = engine.DeviceScene(hs)
    pairs, d_pairs = gpu_pairs(ds, hs, hs.cell_size())
    n = len(pairs)
    assert n > 0
    want = _detect(ds, d_pairs, n)
    assert len(want[0]) > 0
    for target, cap, per_cu, qmin in ((1, 1, 0, 1 << 30), (1 << 20, 64, 0, 1 << 30),
                                      (2048, 7, 0, 1 << 30), (n * 3, 3, 0, 1 << 30),
                                      (1, 1, 1, 1 << 30), (n * 3, 3, 2, 1 << 30),
                                      (1 << 20, 64, 64, 1 << 30), (1, 1, 0, 0), (2048, 7, 4, 0),
                                      (1 << 20, 64, 0, 0)):
        tuning.set(detect_tile_target=target, detect_max_tiles=cap, detect_wg_per_cu=per_cu,
                   detect_queue_min_items=qmin)
        got = _detect(ds, d_pairs, n)
        assert got[0] == want[0], (scene, target, cap, per_cu, qmin)
        assert np.array_equal(got[1], want[1])
        assert got[2] == want[2]
    tuning.set(**tuning.default())
    cap = 1 << 17
    out = DeviceArray((cap * 144,), np.uint8)
    trials = tuning.autotune_detect(ds.scene.geom, d_pairs.ptr, n, out.ptr, cap, reps=2)
    assert [t["knob"] for t in trials] == (["detect_tile_target"] * 6 + ["detect_max_tiles"] * 5
                                           + ["detect_wg_per_cu"] * 6
                                           + ["detect_queue_min_items"] * 2
                                           + ["detect_fp32_screen"] * 2)
    t = tuning.get()
    best = [min(trials[a:b], key=lambda r: r["ms"])["value"]
            for a, b in ((0, 6), (6, 11), (11, 17), (17, 19), (19, 21))]
    assert [t["detect_tile_target"], t["detect_max_tiles"], t["detect_wg_per_cu"],
            t["detect_queue_min_items"], t["detect_fp32_screen"]] == best
    assert out.download().tobytes()[:len(want[0])] == want[0]
    assert _detect(ds, d_pairs, n)[0] == want[0]

