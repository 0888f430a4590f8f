"""The GPU kernels' per-pair arithmetic (exact_math.hpp, dyn_math.hpp),
compiled for the host, against the oracle bit for bit.  Runs without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from test_oracle_golden import bits

HC = os.path.join(ol.ROOT, "delta_amd", "libdh_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC):
        pytest.fail("delta_amd/libdh_hostcheck.so is not built (python -m delta_amd.build)")
    return C.CDLL(HC)


def soa_of(A, B):
    n = len(A)
    s = np.empty((18, n))
    for t, base in ((A, 0), (B, 9)):
        for ax in range(3):
            for v in range(3):
                s[base + 3 * ax + v] = t[:, v, ax]
    return np.ascontiguousarray(s)


@pytest.mark.parametrize("method", [0, 1, 2, 5])
def test_pair_kernels_match_oracle(hc, method):
    g = np.load(os.path.join(ol.ROOT, "tests", "golden", "pair_solvers.npz"))
    rng = np.random.default_rng(2)
    A = np.concatenate([g["A"], rng.uniform(0, 1e-3, (20000, 3, 3))])
    B = np.concatenate([g["B"], rng.uniform(0, 1e-3, (20000, 3, 3))])
    s = soa_of(A, B)
    n = len(A)
    outs = []
    for lib, fn in ((ol.oracle(), "or_tri_pairs"), (hc, "hc_tri_pairs")):
        P = np.zeros(3 * n); Q = np.zeros(3 * n); d = np.zeros(n); f = np.zeros(n, np.uint8)
        getattr(lib, fn)(method, ol.dptr(s), C.c_long(n), C.c_double(2e-4), ol.dptr(P),
                         ol.dptr(Q), ol.dptr(d), f.ctypes.data_as(C.c_void_p))
        outs.append((P, Q, d, f))
    (P1, Q1, d1, f1), (P2, Q2, d2, f2) = outs
    assert np.array_equal(bits(P1), bits(P2)) and np.array_equal(bits(Q1), bits(Q2))
    assert np.array_equal(bits(d1), bits(d2)) and np.array_equal(f1, f2)


def test_dynamics_match_oracle(hc):
    g = np.load(os.path.join(ol.ROOT, "tests", "golden", "forces_dynamics.npz"))
    orc = ol.oracle()
    for i in range(len(g["h"])):
        h = C.c_double(g["h"][i])
        W1 = np.ascontiguousarray(g["ang_in"][i]).copy(); W2 = W1.copy()
        args = [ol.dptr(g[k][i].copy()) for k in ("rot", "J", "Ji", "torque")]
        hc.hc_update_angular(ol.dptr(W1), *args, h)
        orc.or_update_angular(ol.dptr(W2), *args, h)
        assert np.array_equal(bits(W1), bits(W2))
        R1 = g["rot"][i].copy(); R2 = R1.copy(); w1 = np.zeros(3); w2 = np.zeros(3)
        W = np.ascontiguousarray(g["ang_in"][i]).copy()
        hc.hc_update_rotation(ol.dptr(w1), ol.dptr(W), ol.dptr(R1), h)
        orc.or_update_rotation_matrix(ol.dptr(w2), ol.dptr(W), ol.dptr(R2), h)
        assert np.array_equal(bits(R1), bits(R2)) and np.array_equal(bits(w1), bits(w2))


def _separated_and_crossing(rng, m):
    """Pairs separated by A's plane (B lifted along A's normal to a minimum
    height log-uniform in [1e-7, 1e-3]; a quarter with B's edge nearly
    parallel to one of A's) and pairs where an edge of B pierces A's
    interior."""
    A = rng.uniform(0, 1e-3, (m, 3, 3))
    nA = np.cross(A[:, 1] - A[:, 0], A[:, 2] - A[:, 0])
    nA /= np.linalg.norm(nA, axis=1)[:, None]
    B = rng.uniform(0, 1e-3, (m, 3, 3))
    k = m // 4
    B[:k, 1] = A[:k, 1] + rng.normal(size=(k, 3)) * 1e-10
    B[:k, 2] = A[:k, 2]
    h = np.einsum("mvk,mk->mv", B - A[:, :1], nA)
    lift = 10 ** rng.uniform(-7, -3, m) - h.min(axis=1)
    B = B + (lift[:, None] * nA)[:, None, :]
    c = A.mean(axis=1)
    X = np.empty_like(A)
    hh = 10 ** rng.uniform(-7, -4, m)
    X[:, 0] = c + hh[:, None] * nA
    X[:, 1] = c - hh[:, None] * nA
    X[:, 2] = rng.uniform(0, 1e-3, (m, 3))
    return A, B, X


def test_feature_evaluation_and_certificate(hc):
    """pair_features.hpp (the learned path's FP64 stage), host build, against
    the oracle's bfSolver (bf.cpp:409-1100): evaluating bfSolver's winning
    feature reproduces its closest points bit for bit; the certificate accepts
    a feature only when its distance is within 8e-7 of the true distance,
    accepts bfSolver's own answer on separated pairs and never accepts a pair
    of intersecting triangles."""
    g = np.load(os.path.join(ol.ROOT, "tests", "golden", "pair_solvers.npz"))
    rng = np.random.default_rng(5)
    m = 8000
    As, Bs, Xs = _separated_and_crossing(rng, m)
    A = np.concatenate([g["A"], As, As])
    B = np.concatenate([g["B"], Bs, Xs])
    ng = len(g["A"])
    sep = np.zeros(len(A), bool)
    sep[ng:ng + m] = True
    cross = np.zeros(len(A), bool)
    cross[ng + m:] = True
    s = soa_of(A, B)
    n = len(A)
    code = np.zeros(n, np.int32); same = np.zeros(n, np.int32); cbf = np.zeros(n, np.int32)
    call = np.zeros(15 * n, np.int32); dall = np.zeros(15 * n); dbf = np.zeros(n)
    hc.hc_learned_probe(ol.dptr(s), C.c_long(n), *(x.ctypes.data_as(C.c_void_p) for x in
                                                   (code, same, cbf, call, dall, dbf)))
    P = np.zeros(3 * n); Q = np.zeros(3 * n); d = np.zeros(n); f = np.zeros(n, np.uint8)
    ol.oracle().or_tri_pairs(0, ol.dptr(s), C.c_long(n), C.c_double(2e-4), ol.dptr(P),
                             ol.dptr(Q), ol.dptr(d), f.ctypes.data_as(C.c_void_p))
    call = call.reshape(n, 15).astype(bool)
    dall = dall.reshape(n, 15)
    assert same.all()
    assert np.array_equal(bits(dbf), bits(d))
    assert ((code >= 0) & (code < 15)).all()
    # a feature bfSolver never picks can be undefined (NaN on degenerate
    # parallel edges); it never certifies
    nan = np.isnan(dall)
    assert not call[nan].any()
    # every feature's distance is a distance between points of the triangles
    # (bfSolver ranks features by its own rounded intermediates: on near-ties
    # another feature can come out smaller, by 1.8e-10 relative at most here,
    # on a nearly parallel edge pair)
    assert (dall[~nan] >= np.broadcast_to(d[:, None], dall.shape)[~nan] * (1 - 1e-9)).all()
    # soundness on separated pairs (bfSolver's distance is the true one there
    # up to the near-ties above)
    rel = (dall - d[:, None]) / np.where(d > 0, d, 1.0)[:, None]
    assert (rel[sep][call[sep]] <= 8e-7).all()
    # completeness: bfSolver's own answer certifies on separated pairs except
    # on some near-ties of nearly parallel edges, where a neighbouring feature
    # does
    assert cbf[sep].mean() > 0.99
    assert call[sep].any(axis=1).mean() > 0.999
    # intersecting triangles: no feature ever certifies
    assert not call[cross].any()


def _decide(hc, s, ce, cv, fn="hc_learned_decide"):
    n = s.shape[1]
    P = np.zeros(3 * n); Q = np.zeros(3 * n); st = np.zeros(n, np.int32)
    getattr(hc, fn)(ol.dptr(s), C.c_long(n), ce.ctypes.data_as(C.c_void_p),
                    cv.ctypes.data_as(C.c_void_p), ol.dptr(P), ol.dptr(Q),
                    st.ctypes.data_as(C.c_void_p))
    return P, Q, st


def host_pairs(hc, n, seed=3):
    """n pairs of the unique-pair workload from the host build of its
    generator (gen_pairs.hpp), with their kinds; asserts that every column
    index stayed inside the 18 x n block."""
    s = np.zeros(18 * n); t = np.zeros(n); k = np.zeros(n, np.uint8)
    f = hc.hc_gen_tri_pairs
    f.restype = C.c_long
    f.argtypes = [C.c_ulonglong, C.c_long, C.c_long] + [C.c_double] * 6 + [C.c_int] + \
        [C.c_void_p] * 3
    bad = f(seed, 0, n, 1e-3, 1e-7, 1e-2, 0.05, 0.01, 1e-4, 48, s.ctypes.data, t.ctypes.data,
            k.ctypes.data)
    assert bad == 0
    return s.reshape(18, n), t, k


@pytest.mark.parametrize("fn", ["hc_learned_decide", "hc_learned_decide_support"])
def test_exact_choice_is_bfsolver_bit_for_bit(hc, fn):
    """The learned path's decision (pair_features.hpp: two candidate
    features, bfSolver's rule, the acceptance test -- exact_choice() of the
    exact mode, support_choice() of the default path -- and tie folding),
    host build, against the oracle's bfSolver: whenever it accepts a pair
    its P and Q are the oracle's bit for bit, for the ideal candidates
    (bfSolver's own two minima), for the runner-up edge-edge feature and for
    random candidates (an untrained network).  Pairs: the unique-pair
    workload's distribution, the compiled reference's golden pairs and
    hand-built degenerate pairs (ties, parallel edges and faces, touching,
    intersecting).  With the ideal candidates the exact mode accepts all
    separated generated pairs but a few near-parallel ones; the support test
    accepts > 99% of the general ones and sends every near-parallel one to
    the brute force."""
    import scenes
    gen, _, kind = host_pairs(hc, 1 << 14)
    g = np.load(os.path.join(ol.ROOT, "tests", "golden", "pair_solvers.npz"))
    s = np.ascontiguousarray(np.concatenate([gen, soa_of(g["A"], g["B"]),
                                             scenes.degenerate_pairs(motions=24)], axis=1))
    n = s.shape[1]
    ngen = gen.shape[1]
    P0 = np.zeros(3 * n); Q0 = np.zeros(3 * n); d0 = np.zeros(n)
    ol.oracle().or_tri_pairs(0, ol.dptr(s), C.c_long(n), C.c_double(2e-4), ol.dptr(P0),
                             ol.dptr(Q0), ol.dptr(d0), None)
    vals = np.zeros(15 * n)
    hc.hc_feature_values(ol.dptr(s), C.c_long(n), ol.dptr(vals))
    vals = vals.reshape(n, 15)
    # first index of the least value, as bfSolver's strict < loops pick
    ve = np.where(np.isnan(vals[:, :9]), np.inf, vals[:, :9])
    vv = np.where(np.isnan(vals[:, 9:]), np.inf, vals[:, 9:])
    ideal_e = np.argmin(ve, axis=1).astype(np.int32)
    ideal_v = (9 + np.argmin(vv, axis=1)).astype(np.int32)
    second_e = np.argsort(ve, axis=1, kind="stable")[:, 1].astype(np.int32)
    rng = np.random.default_rng(7)
    rand_e = rng.integers(0, 9, n).astype(np.int32)
    rand_v = rng.integers(9, 15, n).astype(np.int32)
    P3 = P0.reshape(3, n); Q3 = Q0.reshape(3, n)
    for name, ce, cv in (("ideal", ideal_e, ideal_v), ("runner-up", second_e, ideal_v),
                         ("random", rand_e, rand_v)):
        P, Q, st = _decide(hc, s, ce, cv, fn)
        acc = st > 0
        Pa = P.reshape(3, n)[:, acc]; Qa = Q.reshape(3, n)[:, acc]
        assert np.array_equal(bits(Pa), bits(P3[:, acc])), name
        assert np.array_equal(bits(Qa), bits(Q3[:, acc])), name
        if name == "ideal" and fn == "hc_learned_decide_support":
            assert acc[:ngen][kind == 0].mean() > 0.99, acc[:ngen][kind == 0].mean()
            assert not acc[:ngen][kind == 2].any()
            assert (st[:ngen] >= 3).any()
        elif name == "ideal":
            sep = kind != 2
            assert acc[:ngen][sep].mean() > 0.985, acc[:ngen][sep].mean()
            assert acc[:ngen][kind == 0].mean() > 0.999
            # ties were resolved by evaluating a few more features
            assert (st[:ngen] >= 3).any()
            assert np.mean(np.maximum(st[:ngen] - 2, 0)) < 1.5
        if name == "random":
            assert acc.mean() > 0.05  # random guesses are still right sometimes


@pytest.mark.parametrize("fn", ["hc_screened_far", "hc_plane_apart", "hc_axes_apart"])
def test_detection_screens_are_lower_bounds(hc, fn):
    """The detection's tests ahead of its binary64 bfSolver (exact_math.hpp:
    screened_far, the binary32 screen; plane_apart, the triangle-plane test;
    axes_apart, the planes and the nine edge-pair axes) skip a triangle pair
    only when the oracle's bfSolver distance
    exceeds the margin: the recorded pair of a C1 particle triangle 0.14 mm
    above the 1 m floor triangle (a contact whose binary32 bfSolver distance
    comes out 0.40 mm, three times too far), small triangles against large
    ones at heights from 1e-6 to 1e-2 m, the unique-pair workload and the
    degenerate pairs.  The binary32 screen still skips nearly every regular
    pair beyond three margins, and the plane test every small triangle held
    flat above a large one beyond the margin."""
    import scenes
    margin = 2e-4
    reg = np.array([[[-0.25007118952381935, -0.066117445166873748, -0.13037171044944115],
                     [-0.25020061591625897, -0.066418704175281645, -0.13048558368355279],
                     [-0.25010106787159503, -0.066278463038880575, -0.13027561019189846]],
                    [[0.53306809858186077, -0.65156380079126563, 0.17061959377314012],
                     [-0.51184027586962633, 0.11783143606624863, -0.22414886892219765],
                     [0.034872830782655906, -0.23586428204518717, -0.048686634857703082]]])
    rng = np.random.default_rng(11)
    m = 20000
    small = rng.uniform(-5e-4, 5e-4, (m, 3, 3)) + rng.uniform(-0.4, 0.4, (m, 1, 3)) * [1, 0, 1]
    big = np.array([[-0.5, 0.0, -0.5], [0.5, 0.0, -0.5], [0.0, 0.0, 0.6]])
    lift = 10.0 ** rng.uniform(-6, -2, m)
    small[:, :, 1] += (lift - small[:, :, 1].min(axis=1))[:, None]
    from delta_amd import shapes
    R = shapes.random_rotation(shapes.Rng(5))
    pairs = [(reg[0], reg[1])] + [((small[i] @ R.T), (big @ R.T)) for i in range(m)]
    s_big = np.zeros((18, len(pairs)))
    for i, (A, B) in enumerate(pairs):
        s_big[:9, i] = A.T.reshape(-1)
        s_big[9:, i] = B.T.reshape(-1)
    gen, _, kind = host_pairs(hc, 1 << 14)
    s = np.ascontiguousarray(np.concatenate([s_big, gen, scenes.degenerate_pairs(motions=8)], axis=1))
    n = s.shape[1]
    far = np.zeros(n, np.uint8)
    getattr(hc, fn)(ol.dptr(s), C.c_long(n), C.c_double(margin), far.ctypes.data_as(C.c_void_p))
    P = np.zeros(3 * n); Q = np.zeros(3 * n); d = np.zeros(n)
    ol.oracle().or_tri_pairs(0, ol.dptr(s), C.c_long(n), C.c_double(2 * margin), ol.dptr(P),
                             ol.dptr(Q), ol.dptr(d), None)
    assert d[0] <= margin and not far[0]
    assert np.all(d[far.astype(bool)] > margin)
    g = slice(len(pairs), len(pairs) + gen.shape[1])
    beyond = d[g] > 3 * margin
    sb = slice(1, len(pairs))
    if fn == "hc_screened_far":
        # effective on the regular pairs (the generated workload) and on
        # most small-over-large pairs beyond three margins
        assert far[g][beyond].mean() > 0.95, far[g][beyond].mean()
        assert far[sb][d[sb] > 3 * margin].mean() > 0.9
    else:
        # the large triangle's plane separates every small one lifted
        # entirely above it beyond the margin
        assert far[sb][lift > 2 * margin].all()
    if fn == "hc_axes_apart":
        # the edge axes add to the planes: among the regular pairs beyond
        # three margins, most of those the planes leave
        pl = np.zeros(n, np.uint8)
        hc.hc_plane_apart(ol.dptr(s), C.c_long(n), C.c_double(margin),
                          pl.ctypes.data_as(C.c_void_p))
        assert np.all(far[pl.astype(bool)])
        left = beyond & ~pl[g].astype(bool)
        assert left.sum() > 100 and far[g][left].mean() > 0.5, far[g][left].mean()


def test_device_sincos_is_correctly_rounded(hc):
    """cr_sincos.hpp, the device's sin and cos in expmap's large-angle
    branch (host build; the device build gives the same bits): correctly
    rounded against a 60-digit Taylor evaluation on 300 arguments from 0.17
    to 1e4, and equal to the host libm's on more than 99.5% of 1e5 arguments
    (the host libm, whose values the reference uses, is itself not correctly
    rounded on about 0.15% of them; the device libm differs from it on
    about 3.5%)."""
    import math
    from decimal import Decimal, getcontext
    getcontext().prec = 70
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899")

    def exact(x, want_cos):
        d = Decimal(x)
        k = int((d / (pi / 2)).to_integral_value())
        r = d - k * (pi / 2)
        sr, t, j = r, r, 1
        while abs(t) > Decimal(10) ** -66:
            t = -t * r * r / ((2 * j) * (2 * j + 1))
            sr += t
            j += 1
        cr, t, j = Decimal(1), Decimal(1), 1
        while abs(t) > Decimal(10) ** -66:
            t = -t * r * r / ((2 * j - 1) * (2 * j))
            cr += t
            j += 1
        q = k % 4
        if want_cos:
            return (cr, -sr, -cr, sr)[q]
        return (sr, cr, -sr, -cr)[q]

    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(0.17, 1.0, 100), rng.uniform(1.0, 10.0, 100),
                         10 ** rng.uniform(1, 4, 100)])
    n = len(xs)
    s = np.zeros(n); c = np.zeros(n)
    hc.hc_cr_sincos(ol.dptr(xs), C.c_long(n), ol.dptr(s), ol.dptr(c))
    for x, sv, cv in zip(xs, s, c):
        for got, want_cos in ((sv, False), (cv, True)):
            e = exact(float(x), want_cos)
            # correctly rounded: no other double is nearer
            err = abs(Decimal(got) - e)
            up = abs(Decimal(math.nextafter(got, math.inf)) - e)
            dn = abs(Decimal(math.nextafter(got, -math.inf)) - e)
            assert err <= up and err <= dn, (x, want_cos)
    xs = np.ascontiguousarray(np.concatenate([rng.uniform(0.17, 10.0, 50000),
                                              10 ** rng.uniform(1, 5, 50000)]))
    n = len(xs)
    s = np.zeros(n); c = np.zeros(n)
    hc.hc_cr_sincos(ol.dptr(xs), C.c_long(n), ol.dptr(s), ol.dptr(c))
    ms = np.array([math.sin(v) for v in xs]); mc = np.array([math.cos(v) for v in xs])
    assert np.mean(s == ms) > 0.995 and np.mean(c == mc) > 0.995
