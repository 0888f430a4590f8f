"""ctypes helpers for the zoltan-application kernels: oracle restatements and
the GPU entry points.  Test infrastructure only."""
import ctypes as C

import numpy as np

import oracle_lib as ol

DP = C.POINTER(C.c_double)
TABLE3 = DP * 3


class ZContact(C.Structure):
    _fields_ = [("pid", C.c_int * 2), ("color", C.c_int * 2), ("point", C.c_double * 3),
                ("normal", C.c_double * 3), ("depth", C.c_double), ("pp", C.c_double * 3),
                ("qq", C.c_double * 3)]


ZDTYPE = np.dtype([("pid", "i4", 2), ("color", "i4", 2), ("point", "f8", 3), ("normal", "f8", 3),
                   ("depth", "f8"), ("pp", "f8", 3), ("qq", "f8", 3)])
assert C.sizeof(ZContact) == ZDTYPE.itemsize


def table(arr):
    """(3, n) float64 array -> table of three row pointers (keeps arr alive)."""
    assert arr.dtype == np.float64 and arr.flags.c_contiguous
    return TABLE3(*(arr[k].ctypes.data_as(DP) for k in range(3)))


def or_zoltan_bf(A, B):
    """A, B: (3 vertices, 3 coords). Returns (rc, P, Q)."""
    a = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(B.T)
    P = np.full(3, np.nan)
    Q = np.full(3, np.nan)
    rc = ol.oracle().or_zoltan_bf(*(ol.dptr(np.ascontiguousarray(r)) for r in a),
                                  *(ol.dptr(np.ascontiguousarray(r)) for r in b),
                                  ol.dptr(P), ol.dptr(Q))
    return rc, P, Q


def or_batch(kind, it, nt, abc, d, e, f, p0=-7.0):
    n = d.shape[1]
    p = np.full((3, n), p0)
    q = np.full((3, n), p0)
    a = np.ascontiguousarray(abc[0:3])
    b = np.ascontiguousarray(abc[3:6])
    c = np.ascontiguousarray(abc[6:9])
    fn = getattr(ol.oracle(), "or_zoltan_" + kind)
    it_t = C.c_uint if kind == "penalty" else C.c_int
    fn(it_t(it), it_t(nt), ol.dptr(a), ol.dptr(b), ol.dptr(c), table(d), table(e), table(f),
       table(p), table(q))
    return p, q
