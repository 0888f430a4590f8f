"""The learned narrow phase on the GPU: network forward and gradient against
the numpy restatement, training, and the certified results against the FP64
brute force (distance within 1e-6 relative for every pair, identical contact
set outside a 1e-9 m band around the threshold)."""
import ctypes as C

import numpy as np
import pytest

import mlp_ref
from delta_amd import _lib, learn, narrow
from delta_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def train_set():
    b = narrow.PairBatch(1 << 20, seed=11)
    return b, learn.Dataset(b.soa, b.n)


def test_forward_matches_numpy(train_set):
    _, ds = train_set
    tr = learn.Trainer(seed=4)
    out = tr.forward(ds, 4096)
    X = ds.X.download()[:4096]
    ref = mlp_ref.forward(learn.init_params(4), X)[4]
    assert np.allclose(out[:, :20], ref[:, :20], rtol=2e-2, atol=2e-2)
    assert np.median(np.abs(out[:, :20] - ref[:, :20])) < 1e-4


def test_gradient_matches_numpy(train_set):
    _, ds = train_set
    tr = learn.Trainer(seed=5)
    g, loss = tr.grad_at(ds, 128, 1024)
    X = ds.X.download()[128:1152]
    feat = ds.feat.download()[128:1152].astype(np.int64)
    bary = ds.bary.download()[128:1152]
    flag = ds.flag.download()[128:1152]
    lref, gref = mlp_ref.loss_and_grad(learn.init_params(5), X, feat, bary, flag)
    assert loss == pytest.approx(lref, rel=1e-3)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 1e-2, rel
    for lo, hi in ((learn.W1, learn.B1), (learn.W2, learn.B2), (learn.W3, learn.B3),
                   (learn.B3, learn.N_PARAMS)):
        seg = np.linalg.norm(g[lo:hi] - gref[lo:hi]) / (np.linalg.norm(gref[lo:hi]) + 1e-30)
        assert seg < 3e-2, (lo, seg)


def run_learned(batch):
    batch.run("learned")
    L = _lib.lib()
    cnt = _lib.DhPairCounters()
    _lib.check(L.dh_tri_pairs_counters(C.byref(cnt)), "counters")
    return batch.results(), cnt


def test_trained_network_certifies_and_matches_brute_force(train_set):
    _, ds = train_set
    tr = learn.Trainer(seed=0)
    l0 = tr.grad_at(ds, 0, 8192)[1]
    tr.fit(ds, epochs=6, batch=4096, lr=3e-3)
    l1 = tr.grad_at(ds, 0, 8192)[1]
    assert l1 < 0.5 * l0
    test = narrow.PairBatch(1 << 16, seed=3)
    test.run("bf64")
    P0, Q0, d0, _ = test.results()
    tr.install()
    (P, Q, d, flags), cnt = run_learned(test)
    rel = np.abs(d - d0) / d0
    assert rel.max() <= 1e-6
    margin = 2e-4
    band = np.abs(d0 - margin) <= 1e-9
    assert np.array_equal((d <= margin)[~band], (d0 <= margin)[~band])
    cert = (flags & 8) != 0
    assert cnt.certified == cert.sum() and cnt.rechecked == (~cert).sum()
    # uncertified pairs come from the brute force itself: bit-identical
    assert np.array_equal(d[~cert].view(np.uint64), d0[~cert].view(np.uint64))
    trained_frac = cert.mean()
    assert trained_frac > 0.75
    # the same path with an untrained network certifies far fewer pairs
    un = learn.Trainer(seed=0)
    un.install()
    (_, _, d_u, fl_u), _ = run_learned(test)
    assert np.abs(d_u - d0).max() / d0.min() < np.inf  # still correct
    assert (np.abs(d_u - d0) / d0).max() <= 1e-6
    assert trained_frac > ((fl_u & 8) != 0).mean() + 0.2
    tr.install()


@pytest.mark.parametrize("n", [1, 63, 65, 5000, 100003])
def test_learned_ragged_sizes(train_set, n):
    """Sizes that leave partial chunks, partial workgroups and a single pair:
    every result within 1e-6 of the brute force, uncertified ones bitwise,
    and the per-workgroup tallies add up to n."""
    _, ds = train_set
    tr = learn.Trainer(seed=0)
    tr.fit(ds, epochs=1, batch=8192, lr=3e-3)
    tr.install()
    test = narrow.PairBatch(n, seed=29)
    test.run("bf64")
    _, _, d0, _ = test.results()
    (_, _, d, flags), cnt = run_learned(test)
    cert = (flags & 8) != 0
    assert ((flags & 8) != 0).sum() + ((flags & 2) != 0).sum() == n
    assert cnt.certified + cnt.rechecked == n
    assert (np.abs(d - d0) / d0).max() <= 1e-6
    assert np.array_equal(d[~cert].view(np.uint64), d0[~cert].view(np.uint64))


def test_network_warm_start_of_the_penalty_newton(train_set):
    """DH_PAIR_LEARNED_PENALTY starts the penalty Newton iteration from the
    network's barycentric outputs; DH_PAIR_PENALTY_CERT from the reference's
    0.33.  Both certify or fall back, so both are correct; the trained start
    needs fewer Newton steps and certifies more pairs."""
    _, ds = train_set
    tr = learn.Trainer(seed=0)
    tr.fit(ds, epochs=6, batch=4096, lr=3e-3)
    tr.install()
    test = narrow.PairBatch(1 << 16, seed=3)
    test.run("bf64")
    _, _, d0, _ = test.results()
    stats = {}
    for m in ("learned_penalty", "penalty_cert"):
        test.run(m)
        L = _lib.lib()
        cnt = _lib.DhPairCounters()
        _lib.check(L.dh_tri_pairs_counters(C.byref(cnt)), "counters")
        _, _, d, fl = test.results()
        cert = (fl & 8) != 0
        assert (np.abs(d - d0) / d0).max() <= 1e-6
        assert np.array_equal(d[~cert].view(np.uint64), d0[~cert].view(np.uint64))
        assert cnt.certified == cert.sum() and cnt.rechecked == (~cert).sum()
        stats[m] = (cnt.features_fp64 / test.n, cert.mean())
    (its_w, cert_w), (its_c, cert_c) = stats["learned_penalty"], stats["penalty_cert"]
    assert its_w < its_c, stats
    assert cert_w >= cert_c, stats
