"""The zoltan-application kernels on the GPU against the oracle: bf and the
detection driver bit for bit; penalty and hybrid to the rounding of the
device's pow/log/log10 against glibc's."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
import zoltan_lib as zl
from delta_amd import narrow, shapes, zoltan
from delta_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

G = os.path.join(ol.ROOT, "tests", "golden", "zoltan.npz")


def bits(x):
    return np.ascontiguousarray(np.asarray(x, np.float64)).view(np.uint64)


def test_zoltan_bf_matches_oracle():
    g = np.load(G)
    A, B = g["bf_A"], g["bf_B"]
    n = len(A)
    soa = np.empty((18, n))
    for t, base in ((A, 0), (B, 9)):
        for ax in range(3):
            for v in range(3):
                soa[base + 3 * ax + v] = t[:, v, ax]
    gen = narrow.PairBatch(1 << 14, seed=5).soa.download()
    soa = np.concatenate([soa, gen], axis=1)
    n = soa.shape[1]
    d_soa = DeviceArray.from_host(np.ascontiguousarray(soa))
    P = DeviceArray.from_host(np.full((3, n), -5.0))
    Q = DeviceArray.from_host(np.full((3, n), -5.0))
    fl = DeviceArray.zeros((n,), np.uint8)
    zoltan.bf(d_soa, n, P, Q, flags=fl)
    P, Q, fl = P.download(), Q.download(), fl.download()
    early = 0
    for i in range(n):
        Ai = soa[0:9, i].reshape(3, 3).T
        Bi = soa[9:18, i].reshape(3, 3).T
        rc, P0, Q0 = zl.or_zoltan_bf(Ai, Bi)
        if rc < 0:
            early += 1
            assert fl[i] == 1 and np.all(P[:, i] == -5.0) and np.all(Q[:, i] == -5.0)
        else:
            assert np.array_equal(bits(P[:, i]), bits(P0)) and np.array_equal(bits(Q[:, i]), bits(Q0))
            assert fl[i] == (4 if rc == 1 else 0)
    assert early >= 20


@pytest.mark.parametrize("case", ["pen_a", "pen_b", "hyb_a", "hyb_b", "hyb_c"])
def test_zoltan_batches_match_oracle(case):
    g = np.load(G)
    it, nt = (int(v) for v in g[f"{case}_range"])
    kind = "penalty" if case.startswith("pen") else "hybrid"
    d, e, f = (np.ascontiguousarray(g[f"{case}_{k}"]) for k in "def")
    abc = g[f"{case}_abc"]
    p0, q0 = zl.or_batch(kind, it, nt, abc, d, e, f)
    n = d.shape[1]
    dd, de, df = (DeviceArray.from_host(x) for x in (d, e, f))
    p = DeviceArray.from_host(np.full((3, n), -7.0))
    q = DeviceArray.from_host(np.full((3, n), -7.0))
    getattr(zoltan, kind)(it, nt, abc[0:3], abc[3:6], abc[6:9], dd, de, df, p, q)
    p, q = p.download(), q.download()
    untouched = np.all(p0 == -7.0, axis=0)
    assert np.array_equal(untouched, np.all(p == -7.0, axis=0))
    scale = np.abs(np.concatenate([d, e, f])).max()
    assert np.abs(p - p0).max() <= 1e-12 * scale and np.abs(q - q0).max() <= 1e-12 * scale


def _scene(npart, seed):
    rng = shapes.Rng(seed)
    tris, pid = [], []
    for k in range(npart):
        m = shapes.jittered_icosphere(0, 0.5, 0.1, rng)
        m = m @ shapes.random_rotation(rng).T + np.array([1.02 * (k % 3), 1.02 * (k // 3), 0.0])
        tris.append(m)
        pid += [k] * len(m)
    t = np.concatenate(tris)  # (n, vertex, axis)
    return np.ascontiguousarray(t.transpose(1, 2, 0)), np.array(pid, np.int32)


@pytest.mark.parametrize("two_range", [False, True])
def test_zoltan_detection_matches_oracle(two_range):
    t, pid = _scene(6, 12)
    n = t.shape[2]
    T = ((zl.DP * 3) * 6)()
    for v in range(3):
        for ax in range(3):
            T[v][ax] = t[v, ax].ctypes.data_as(zl.DP)
    out = (zl.ZContact * 8192)()
    if two_range:
        s1, e1, s2, e2 = 0, n // 2, 0, n
    else:
        s1, e1, s2, e2 = 0, n, 0, 0
    m = ol.oracle().or_zoltan_detection(s1, e1, s2, e2, 0 if two_range else 1, T,
                                        pid.ctypes.data_as(C.POINTER(C.c_int)), out, 8192)
    want = np.frombuffer(bytes(out), dtype=zl.ZDTYPE)[:m]
    assert m >= 4
    dt = [DeviceArray.from_host(np.ascontiguousarray(t[v])) for v in range(3)]
    dpid = DeviceArray.from_host(pid)
    got = zoltan.detection(s1, e1, dt, None, dpid, 6, s2=s2 if two_range else None,
                           e2=e2 if two_range else None, cap=4)
    assert len(got) == m
    for f in ("point", "normal", "depth", "pp", "qq"):
        assert np.array_equal(bits(got[f]), bits(want[f])), f
    assert np.array_equal(got["pid"], want["pid"])
