"""GJK (gjk.cpp:302-460): the oracle's restatement against the compiled
reference's golden vectors bit for bit, against the brute-force triangle
distance between convex hulls, and the kernel header (gjk.hpp, host build)
against the oracle.  Runs without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import shapes

G = os.path.join(ol.ROOT, "tests", "golden", "gjk.npz")


def or_gjk(a, b):
    o = ol.oracle()
    o.or_gjk.restype = C.c_double
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    p = np.zeros(3)
    q = np.zeros(3)
    d = o.or_gjk(ol.dptr(a), len(a), ol.dptr(b), len(b), ol.dptr(p), ol.dptr(q))
    return d, p, q


def golden_cases():
    g = np.load(G)
    for k in range(len(g["d"])):
        a = g["a"][g["offa"][k]:g["offa"][k + 1]]
        b = g["b"][g["offb"][k]:g["offb"][k + 1]]
        yield a, b, g["p"][k], g["q"][k], g["d"][k]


def bits(x):
    return np.ascontiguousarray(np.asarray(x, np.float64)).view(np.uint64)


def test_oracle_gjk_matches_reference_golden():
    n = 0
    for a, b, p0, q0, d0 in golden_cases():
        d, p, q = or_gjk(a, b)
        assert bits(d) == bits(d0) and np.array_equal(bits(p), bits(p0)) \
            and np.array_equal(bits(q), bits(q0)), (len(a), len(b))
        n += 1
    assert n > 400


def hull_distance_bf(va, fa, vb, fb):
    """Minimum of the oracle's bfSolver over all face pairs of two closed
    triangle meshes (the distance between their surfaces)."""
    o = ol.oracle()
    best = np.inf
    P = np.zeros(3)
    Q = np.zeros(3)
    for ta in fa:
        A = np.ascontiguousarray(va[ta].T)
        for tb in fb:
            B = np.ascontiguousarray(vb[tb].T)
            o.or_bf_solver(ol.dptr(A[0].copy()), ol.dptr(A[1].copy()), ol.dptr(A[2].copy()),
                           ol.dptr(B[0].copy()), ol.dptr(B[1].copy()), ol.dptr(B[2].copy()),
                           ol.dptr(P), ol.dptr(Q))
            best = min(best, np.linalg.norm(Q - P))
    return best


@pytest.mark.parametrize("gap", [1e-3, 0.05, 0.7, 3.0])
def test_gjk_equals_surface_distance_of_separated_hulls(gap):
    """For separated convex polytopes the GJK distance is the distance between
    their surfaces, to within the iteration's 1e-6 stopping tolerance, and
    never below it (|v| is a point of the Minkowski difference)."""
    rng = shapes.Rng(int(gap * 1000) + 3)
    va, fa = shapes.icosphere(0)
    vb, fb = shapes.icosphere(0)
    va = va @ shapes.random_rotation(rng).T
    vb = vb @ shapes.random_rotation(rng).T + np.array([2.0 + gap, 0.3, -0.1])
    d, p, q = or_gjk(va, vb)
    ref = hull_distance_bf(va, fa, vb, fb)
    assert ref > 0
    assert d >= ref * (1 - 1e-12)
    assert d - ref <= 2e-6
    assert abs(np.linalg.norm(q - p) - d) <= 1e-12 * max(d, 1.0)


def test_gjk_overlapping_sets_report_contact():
    va, _ = shapes.icosphere(1)
    d, _, _ = or_gjk(va, va + np.array([0.5, 0.0, 0.0]))
    assert d <= 1e-6


def test_kernel_header_matches_oracle():
    hc = C.CDLL(os.path.join(ol.ROOT, "delta_amd", "libdh_hostcheck.so"))
    hc.hc_gjk.restype = C.c_double
    cases = list(golden_cases())
    rng = np.random.default_rng(9)
    for _ in range(200):
        a = rng.normal(size=(rng.integers(1, 60), 3))
        b = rng.normal(size=(rng.integers(1, 60), 3)) + rng.normal(size=3) * 3
        cases.append((a, b, None, None, None))
    for a, b, *_ in cases:
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        d0, p0, q0 = or_gjk(a, b)
        p = np.zeros(3)
        q = np.zeros(3)
        d = hc.hc_gjk(ol.dptr(a), len(a), ol.dptr(b), len(b), ol.dptr(p), ol.dptr(q))
        assert bits(d) == bits(d0) and np.array_equal(bits(p), bits(p0)) \
            and np.array_equal(bits(q), bits(q0))


def test_gjk_driver_emits_on_close_particles():
    """or_detect(DH_GJK): one contact per particle pair whose hulls come
    within epsA + epsB, with the contact fields of contactpoint.cpp:116-155
    and triangle indices -1; none for distant pairs."""
    import scenes
    from delta_amd import _lib
    emitted = 0
    for name, A, B in scenes.pair_cases():
        rc, cs, st = ol.or_detect(_lib.DH_GJK, A, B, 1e-4, 1e-4, 1, 1, 3, 4)
        assert rc == 0 and len(cs) <= 1 and st.pairs_evaluated == 1
        d, p, q = or_gjk(A.reshape(-1, 3), B.reshape(-1, 3))
        assert (len(cs) == 1) == (np.linalg.norm(q - p) <= 2e-4)
        if len(cs):
            c = cs[0]
            assert np.array_equal(bits(c["P"]), bits(p)) and np.array_equal(bits(c["Q"]), bits(q))
            assert c["triA"] == -1 and c["triB"] == -1 and c["master"] == 3 and c["slave"] == 4
            emitted += 1
    far = scenes.particle(1, 0.5e-3, 5, [5e-3, 0, 0])
    rc, cs, _ = ol.or_detect(_lib.DH_GJK, scenes.pair_cases()[0][1], far, 1e-4, 1e-4)
    assert rc == 0 and len(cs) == 0
    assert emitted >= 3
