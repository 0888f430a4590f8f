"""ctypes helpers for the ISPC kernels (bf.ispc, penalty.ispc): the oracle
restatement, the host build of the kernel math and the GPU entry points, all
over vertex-major SoA tables pp[v][k][i].  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

import oracle_lib as ol

DP = C.POINTER(C.c_double)
T3 = DP * 3
T63 = T3 * 6
HC = os.path.join(ol.ROOT, "delta_amd", "libdh_hostcheck.so")
CODES = (2.0, -2.0, -1.0, 0.0)


def tables(pp):
    """pp: (6, 3, n) float64 -> pp[v][k] pointer table."""
    assert pp.dtype == np.float64 and pp.flags.c_contiguous
    t = T63()
    for v in range(6):
        for k in range(3):
            t[v][k] = pp[v, k].ctypes.data_as(DP)
    return t


def rows(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return T3(*(a[k].ctypes.data_as(DP) for k in range(3)))


def outputs(n, p0=-7.0, q0=-8.0, d0=-9.0):
    return np.full((3, n), p0), np.full((3, n), q0), np.full(n, d0)


def run_cpu(kind, pp, impl="oracle", P=None, Q=None, d=None):
    """kind 'bf' or 'pen'; impl 'oracle' (C restatement) or 'host' (host
    build of the kernel header).  P, Q, d: in/out (defaults: sentinels)."""
    n = pp.shape[2]
    if P is None:
        P, Q, d = outputs(n)
    P, Q, d = P.copy(), Q.copy(), d.copy()
    if impl == "oracle":
        fn = getattr(ol.oracle(), "or_ispc_" + kind)
        fn(C.c_uint(n), tables(pp), rows(P), rows(Q), d.ctypes.data_as(DP))
    else:
        lib = C.CDLL(HC)
        lib.hc_ispc(C.c_int(1 if kind == "pen" else 0), C.c_uint(n), tables(pp), rows(P),
                    rows(Q), d.ctypes.data_as(DP))
    return P, Q, d


def one(kind, V, impl="oracle"):
    """V: (6, 3) vertices of one pair -> (P, Q, distance) with sentinels in."""
    pp = np.ascontiguousarray(np.asarray(V, float).reshape(6, 3, 1))
    P, Q, d = run_cpu(kind, pp, impl)
    return P[:, 0], Q[:, 0], d[0]


def ttd_pairs(rng, n, length=10.0):
    """The generator of ttd.example:16-45: A's vertices uniform in
    [0, length)^3, B's in [2 length, 3 length)^3 (drand48 replaced by numpy)."""
    a = rng.uniform(0, length, (3, 3, n))
    b = rng.uniform(0, length, (3, 3, n)) + 2 * length
    return np.ascontiguousarray(np.concatenate([a, b]))
