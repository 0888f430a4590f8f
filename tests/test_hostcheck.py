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
