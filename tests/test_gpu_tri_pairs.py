"""Unique-pair kernels on the GPU against the oracle, bit for bit.

The pairs come from the seeded HBM generator (the benchmark's distribution:
log-uniform separation 1e-7..1e-2 m, near-parallel and intersecting pairs)
plus hand-made degenerate cases."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import narrow

pytestmark = pytest.mark.gpu


def oracle_pairs(method, soa, eps_total=2e-4):
    n = soa.shape[1]
    P = np.zeros((3, n)); Q = np.zeros((3, n)); d = np.zeros(n); f = np.zeros(n, np.uint8)
    rc = ol.oracle().or_tri_pairs(method, ol.dptr(np.ascontiguousarray(soa)), C.c_int64(n),
                                  C.c_double(eps_total), ol.dptr(P), ol.dptr(Q), ol.dptr(d),
                                  f.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return P, Q, d, f


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def degenerate_soa():
    rng = np.random.default_rng(11)
    n = 4096
    A = rng.uniform(0, 1e-3, (n, 3, 3))
    B = rng.uniform(0, 1e-3, (n, 3, 3))
    k = n // 8
    B[:k] = A[:k] + np.array([2e-5, 0, 0])       # coplanar parallel offset
    B[k:2 * k, 0] = A[k:2 * k, 0]                 # shared vertex
    B[2 * k:3 * k, :2] = A[2 * k:3 * k, :2]       # shared edge
    A[3 * k:4 * k, 2] = A[3 * k:4 * k, 0]         # repeated vertex (zero area)
    B[4 * k:5 * k, 2] = 0.5 * (B[4 * k:5 * k, 0] + B[4 * k:5 * k, 1])  # collinear
    B[5 * k:6 * k] = A[5 * k:6 * k][:, [1, 2, 0]]  # identical triangles
    s = np.empty((18, n))
    for t, base in ((A, 0), (B, 9)):
        for ax in range(3):
            for v in range(3):
                s[base + 3 * ax + v] = t[:, v, ax]
    return s


@pytest.fixture(scope="module")
def gen_batch():
    return narrow.PairBatch(1 << 16, seed=3)


@pytest.mark.parametrize("method", ["bf64", "penalty64", "hybrid64", "penalty_adapt64", "bf64_exact"])
def test_generated_pairs_bitexact(gen_batch, method):
    soa = gen_batch.soa.download()
    gen_batch.run(method)
    P, Q, d, f = gen_batch.results()
    Po, Qo, do, fo = oracle_pairs(narrow.METHODS[method], soa)
    assert np.array_equal(bits(d), bits(do)), np.count_nonzero(bits(d) != bits(do))
    assert np.array_equal(bits(P), bits(Po))
    assert np.array_equal(bits(Q), bits(Qo))
    assert np.array_equal(f, fo)


@pytest.mark.parametrize("method", ["bf64", "penalty64", "hybrid64", "penalty_adapt64", "bf64_exact"])
def test_degenerate_pairs_bitexact(method):
    soa = degenerate_soa()
    b = narrow.PairBatch.from_host(soa)
    b.run(method)
    P, Q, d, f = b.results()
    Po, Qo, do, fo = oracle_pairs(narrow.METHODS[method], soa)
    assert np.array_equal(bits(d), bits(do))
    assert np.array_equal(bits(P), bits(Po))
    assert np.array_equal(bits(Q), bits(Qo))
    assert np.array_equal(f, fo)


def test_generator_realises_target_distance(gen_batch):
    """General pairs are placed by bisection at the target separation; the
    brute-force distance must reproduce it (a property of the full-size data
    that needs no oracle)."""
    gen_batch.run("bf64")
    d = gen_batch.d.download()
    tgt = gen_batch.target.download()
    kind = gen_batch.kind.download()
    g = kind == 0
    rel = np.abs(d[g] - tgt[g]) / tgt[g]
    assert np.median(rel) < 1e-9 and np.quantile(rel, 0.999) < 1e-6
    par = kind == 1
    assert np.all(d[par] >= 0.999 * tgt[par])
    assert np.all(np.isfinite(d))


@pytest.mark.parametrize("n,kind", [(1, "general"), (100, "general"), (1 << 17, "general"),
                                    (1 << 15, "far-parallel")])
def test_mixed_fp32_screen_with_fp64_recheck(n, kind):
    """DH_PAIR_MIXED: a pair is finished in binary32 only when its distance
    is bounded within 1e-6 relative in binary64 on the binary32 triangles
    and clear of the recheck band; every other pair is recomputed in
    binary64 and equals bfSolver bit for bit.  Per pair the distance is
    within 1e-6 relative of the oracle's, and the contact set is the brute
    force's.  Sets: the unique-pair workload, and pairs that are all
    near-parallel (edges within 1e-4 rad) at 1e-3..1e-2 m, far outside the
    band."""
    import ctypes as C
    from delta_amd import _lib
    eps = 2e-4
    params = None
    if kind == "far-parallel":
        params = _lib.DhGenParams(1e-3, 1e-3, 1e-2, 1.0, 0.0, 1e-4, 48)
    b = narrow.PairBatch(n, seed=17, params=params)
    soa = b.soa.download()
    P0, Q0, d0, _ = oracle_pairs(0, soa, eps)
    b.make_f32()
    b.run("mixed", eps_total=eps)
    P, Q, d, fl = b.results()
    cnt = _lib.DhPairCounters()
    _lib.check(_lib.lib().dh_tri_pairs_counters(C.byref(cnt)), "counters")
    re = (fl & 2) != 0
    assert cnt.screened + cnt.rechecked == n and cnt.rechecked == re.sum()
    assert np.array_equal(bits(d[re]), bits(d0[re]))
    assert np.array_equal(bits(P[:, re]), bits(P0[:, re]))
    assert np.array_equal(bits(Q[:, re]), bits(Q0[:, re]))
    # everything within 2*eps (so every contact) went through binary64
    assert re[d0 < 2 * eps].all()
    assert np.array_equal(d <= eps, d0 <= eps)
    # every distance within 1e-6 relative of the brute force
    assert (d0 > 0).all()
    rel = np.abs(d - d0) / d0
    assert rel.max() <= 1e-6, rel.max()
    if n == 1 << 17:
        assert (~re).mean() > 0.1, (~re).mean()


def test_mixed_needs_binary32_copy():
    from delta_amd import _lib
    b = narrow.PairBatch(64, seed=1)
    with pytest.raises(_lib.DhError):
        b.run("mixed")


def test_exact_mode_flags_generated_intersections(gen_batch):
    """The generator pushes 1% of the B triangles through A (kind 2): the exact
    mode reports them with d = 0, and none of the separated pairs."""
    gen_batch.run("bf64_exact")
    _, _, d, f = gen_batch.results()
    kind = gen_batch.kind.download()
    hit = (f & 16) != 0
    assert hit[kind == 2].mean() > 0.99
    assert not hit[kind != 2].any()
    assert (d[hit] == 0).all()
