"""Pins of the ISPC restatements (bf.ispc, penalty.ispc; no ispc compiler
here, so no golden vectors): hand-worked pre-pass codes, the two case-1
fall-throughs, the early returns, closed forms, an independent long-double
brute force over the 15 features, an independent dense Gauss-Newton
formulation of the penalty steps, and the host build of the kernel header
bit for bit against the C restatement."""
import math

import numpy as np
import pytest

import ispc_lib as il
from test_oracle_math import _pt_tri, _seg_seg, random_separated

LD = np.longdouble

# (vertices p1..p6, pre-pass code).  Each code was worked by hand from
# bf.ispc:9-174 (normals nb of B and na of A; for j < 3 the pre-pass reads
# nn[0][0], nn[0][1], nn[1][0] = nb.x, na.x, nb.y and the x components of
# w0 and dir columns 0..2; u, v are A's edges p2-p1, p3-p1):
#  -1: nb = (2,0,0), na = (0,2,0); j=0: a = -4, b = -4, r = 1, P = p2,
#      w = P - p4 = (0,-2,-4), D = 4*4 - 4*5 = -4, s = -16/-4 = 4 > 1.
#   2: both triangles in x = 0: every x difference is 0, a = -0, b = 0.
#   0: nb = (0,1,0), na = (0,0,-1); j < 3: a = 1, b = 1, r = 1 and P runs
#      over p2, p3, p1 with (s, t) = (1,-0), (-0,1), (-0,-0); j >= 3: a = -0,
#      b = -1, r = 0, P = p4, p5, p6 with (s, t) in the triangle.
#  -2: nb = (0,-1,2), na = (0,0,-2); j=0: a = -1, b = 0, r = -inf.
# In all four the edge-edge and vertex-face minima tie (vertex to vertex),
# so distance keeps the code (ptmin < ssmin fails, bf.ispc:874-877).
PREPASS = [
    ([[2, 0, 0], [0, 0, 0], [0, 0, 1], [0, 2, 4], [0, 0, 4], [0, 0, 3]], -1.0,
     [0, 0, 1], [0, 0, 3]),
    ([[0, 2, 0], [0, 0, 1], [0, 0, 0], [0, 2, 2], [0, 0, 2], [0, 0, 3]], 2.0,
     [0, 0, 1], [0, 0, 2]),
    ([[1, 0, 0], [0, 0, 0], [0, 1, 0], [1, 0, 2], [0, 0, 4], [1, 0, 3]], 0.0,
     [1, 0, 0], [1, 0, 2]),
    ([[0, 2, 0], [1, 0, 0], [0, 0, 0], [1, 0, 2], [0, 2, 3], [0, 0, 2]], -2.0,
     [1, 0, 0], [1, 0, 2]),
]


@pytest.mark.parametrize("V,code,P,Q", PREPASS)
def test_prepass_codes(V, code, P, Q):
    p, q, d = il.one("bf", V)
    assert d == code
    assert np.array_equal(p, P) and np.array_equal(q, Q)


def test_early_returns_leave_outputs():
    # B collinear: zero normal
    p, q, d = il.one("bf", [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [2, 2, 1]])
    assert d == -1 and np.all(p == -7) and np.all(q == -8)
    # the mixed test nb.x == 0 && na.y == 0 && na.z == 0 (bf.ispc:110): A in
    # the plane x = 0 (na along x), B in z = 5 (nb along z)
    p, q, d = il.one("bf", [[0, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 5], [1, 0, 5], [0, 1, 5]])
    assert d == -1 and np.all(p == -7) and np.all(q == -8)


def test_vertex_over_face():
    h = 0.037
    A = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    # B's first vertex over A's interior (feature 0), the others high up
    p, q, d = il.one("bf", A + [[0.2, 0.3, h], [0.25, 0.3, 1.0], [0.2, 0.5, 0.8]])
    assert d == pytest.approx(h, rel=1e-14)
    assert np.allclose(p, [0.2, 0.3, 0], atol=1e-15) and np.array_equal(q, [0.2, 0.3, h])
    # A's second vertex under B's face (feature 4)
    B = [[0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]]
    p, q, d = il.one("bf", [[-1, -1, -2], [0.3, 0.2, 1 - h], [2, -1, -3]] + B)
    assert d == pytest.approx(h, rel=1e-12)
    assert np.array_equal(p, [0.3, 0.2, 1 - h]) and np.allclose(q, [0.3, 0.2, 1.0])


def test_vertex_face_case1_falls_through():
    """bf.ispc:891-899: vertex-face feature 1 (B's second vertex over A) has
    no break, so Q is B's third vertex."""
    A = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    p5 = [0.2, 0.3, 0.5]
    p, q, d = il.one("bf", A + [[2, 2, 5], p5, [2, -1, 4]])
    assert d == pytest.approx(0.5, rel=1e-14)
    assert np.allclose(p, [0.2, 0.3, 0], atol=1e-15)
    assert np.array_equal(q, [2, -1, 4])


def test_edge_case1_falls_through():
    """bf.ispc:505-521: edge pair 1 (A's p1-p2 with B's p5-p6) runs into case
    2, whose Q lies on B's edge p6-p4 at the same parameter."""
    h = 0.25
    A = [[0, 0, 0], [1, 0, 0], [0.5, -1, 0]]
    B = [[0.4, 0, h + 3], [0.4, 1, h], [0.4, -1.5, h]]
    p, q, d = il.one("bf", A + B)
    # P from the pair-1 solve: (0.4, 0, 0) on A's p1-p2 (ssc = 0.6); case 2's
    # Q = p4 + (p6 - p4) ttc with the pair-1 ttc = 0.6
    assert np.allclose(p, [0.4, 0, 0], atol=1e-15)
    assert np.allclose(q, [0.4, -0.9, h + 3 - 1.8], atol=1e-15)
    assert d in il.CODES


def _features(A, B):
    """Long-double distances of the 6 vertex-face and 9 edge-edge features,
    in the reference's numbering."""
    A = A.astype(LD)
    B = B.astype(LD)
    vf = [_pt_tri(B[j], *A) for j in range(3)] + [_pt_tri(A[j], *B) for j in range(3)]
    ee = [_seg_seg(A[k], A[(k + 1) % 3], B[m], B[(m + 1) % 3]) for k in range(3) for m in range(3)]
    return np.array(vf, dtype=LD), np.array(ee, dtype=LD)


@pytest.mark.parametrize("scale", [1.0, 10.0])
def test_against_independent_brute_force(scale):
    """Separated pairs of nearby triangles: where one feature family wins by
    a margin, distance, P and Q follow it, with the two fall-throughs and the
    kept pre-pass code.  (Far-apart pairs, as in ttd.example, mostly tie at a
    vertex pair, where the two families agree.)"""
    rng = np.random.default_rng(11)
    pairs = random_separated(rng, 1500)
    pp = np.ascontiguousarray(
        np.stack([np.concatenate([A, B]) for A, B in pairs], axis=2) * scale)
    P, Q, d = il.run_cpu("bf", pp)
    checked = {"vf": 0, "ee": 0, "vf1": 0, "ee1": 0}
    for i in range(pp.shape[2]):
        A, B = pp[0:3, :, i], pp[3:6, :, i]
        vf, ee = _features(A, B)
        vmin, emin = float(vf.min()), float(ee.min())
        pq = np.linalg.norm(P[:, i] - Q[:, i])
        if vmin < emin * (1 - 1e-9):
            assert d[i] == pytest.approx(vmin, rel=1e-12)
            if int(np.argmin(vf)) == 1:
                assert np.array_equal(Q[:, i], B[2])
                checked["vf1"] += 1
            else:
                assert pq == pytest.approx(vmin, rel=1e-9)
                checked["vf"] += 1
        elif emin < vmin * (1 - 1e-9):
            assert d[i] in il.CODES
            if int(np.argmin(ee)) == 1:
                checked["ee1"] += 1
            else:
                assert pq == pytest.approx(emin, rel=1e-9)
                checked["ee"] += 1
    assert min(checked.values()) > 0, checked


def test_host_header_equals_restatement():
    """The kernels' header (host build) against the C restatement, bit for
    bit, with preset outputs (in/out semantics) and degenerate rows."""
    rng = np.random.default_rng(5)
    pp = il.ttd_pairs(rng, 3000, 1e-3)
    # degenerate B, degenerate A, coplanar and shared-vertex rows
    pp[3:6, :, 0] = pp[3, :, 0][None]
    pp[0:3, :, 1] = pp[0, :, 1][None]
    pp[:, 2, 2] = 0.0
    pp[3, :, 3] = pp[0, :, 3]
    for case in PREPASS:
        pp = np.concatenate([pp, np.asarray(case[0], float).reshape(6, 3, 1)], axis=2)
    pp = np.ascontiguousarray(pp)
    n = pp.shape[2]
    P0 = rng.uniform(-1, 1, (3, n))
    Q0 = rng.uniform(-1, 1, (3, n))
    d0 = rng.uniform(-1, 1, n)
    for kind in ("bf", "pen"):
        a = il.run_cpu(kind, pp, "oracle", P0, Q0, d0)
        b = il.run_cpu(kind, pp, "host", P0, Q0, d0)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), kind


def _gauss_newton(A, B, scaled):
    """The five regularised penalty steps as a dense 4x4 solve: minimise
    |P - Q|^2 + r/2 sum max(c, 0)^2 over the barycentric x (P = A0 + BA x0 +
    CA x1, Q = B0 + ED x2 + FD x3) from x = 0.33, with Hessian 2 G^T G +
    r J^T J + delta I over the active constraints; scaled(j) says whether
    delta is multiplied by r before step j."""
    BA, CA, ED, FD = A[1] - A[0], A[2] - A[0], B[1] - B[0], B[2] - B[0]
    G = np.stack([BA, CA, -ED, -FD], 1)
    H = 2 * G.T @ G
    tr = np.trace(H)
    delta = 0.1 * tr
    lam = math.sqrt(0.0125 * tr)
    r = lam * 10 ** (math.log(lam) + 14)
    Jc = np.array([[-1, 0, 0, 0], [0, -1, 0, 0], [1, 1, 0, 0], [0, 0, -1, 0], [0, 0, 0, -1],
                   [0, 0, 1, 1]], float)
    off = np.array([0, 0, -1, 0, 0, -1], float)
    x = np.full(4, 0.33)
    for j in range(5):
        if scaled(j):
            delta = r * delta
        c = Jc @ x + off
        J, m = Jc[c > 0], c[c > 0]
        grad = 2 * G.T @ ((A[0] - B[0]) + G @ x) + r * J.T @ m
        x = x - np.linalg.solve(H + r * J.T @ J + delta * np.eye(4), grad)
    return A[0] + BA * x[0] + CA * x[1], B[0] + ED * x[2] + FD * x[3]


def test_penalty_against_dense_formulation():
    """penalty.ispc's schedule (delta *= r before the fifth step only) and
    step against the dense formulation; the other two schedules (never,
    every step) are measurably different, so the check decides."""
    rng = np.random.default_rng(3)
    pp = il.ttd_pairs(rng, 200, 1e-3)
    P, Q, d = il.run_cpu("pen", pp)
    far = {"never": 0, "always": 0}
    for i in range(pp.shape[2]):
        A, B = pp[0:3, :, i], pp[3:6, :, i]
        p, q = _gauss_newton(A, B, lambda j: j == 4)
        # the hand-unrolled LDL and LAPACK round differently: 4e-12 m at most
        # here, coordinates ~3e-3 m
        assert np.allclose(P[:, i], p, rtol=0, atol=2e-11)
        assert np.allclose(Q[:, i], q, rtol=0, atol=2e-11)
        assert d[i] == pytest.approx(np.linalg.norm(P[:, i] - Q[:, i]), rel=1e-14)
        for name, sch in (("never", lambda j: False), ("always", lambda j: True)):
            p2, q2 = _gauss_newton(A, B, sch)
            if max(np.abs(p2 - P[:, i]).max(), np.abs(q2 - Q[:, i]).max()) > 1e-8:
                far[name] += 1
    # pairs already converged after four steps do not tell "never" apart
    assert far["never"] > 30 and far["always"] > 150, far
