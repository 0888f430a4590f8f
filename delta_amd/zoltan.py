"""The reference's zoltan-application contact kernels (zoltan/bf.cpp,
penalty.cpp, hybrid.cpp, contact.cpp) over the C ABI: the same names and
argument meaning, with device arrays (DeviceArray of shape (3, n): row k is
coordinate k, the reference's `d[k][i]`).  No CPU fallback."""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

DP = C.c_void_p
TABLE3 = C.c_void_p * 3


class ZoltanContact(C.Structure):
    """zoltan/contact.h:34-43."""
    _fields_ = [("pid", C.c_int * 2), ("color", C.c_int * 2), ("point", C.c_double * 3),
                ("normal", C.c_double * 3), ("depth", C.c_double), ("pp", C.c_double * 3),
                ("qq", C.c_double * 3)]


CONTACT_DTYPE = np.dtype([("pid", "i4", 2), ("color", "i4", 2), ("point", "f8", 3),
                          ("normal", "f8", 3), ("depth", "f8"), ("pp", "f8", 3),
                          ("qq", "f8", 3)])
assert C.sizeof(ZoltanContact) == CONTACT_DTYPE.itemsize

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = _lib.lib()
    L.dh_zoltan_bf.argtypes = [DP, C.c_int64, DP, DP, DP, DP, DP]
    L.dh_zoltan_penalty.argtypes = [C.c_uint, C.c_uint, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), TABLE3, TABLE3,
                                    TABLE3, TABLE3, TABLE3, DP]
    L.dh_zoltan_hybrid.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), TABLE3, TABLE3,
                                   TABLE3, TABLE3, TABLE3, DP]
    T6 = (C.c_void_p * 3) * 6
    L.dh_zoltan_detection.argtypes = [C.c_int, C.c_int, T6, DP, DP, C.c_int, DP, C.c_int,
                                      C.POINTER(C.c_int), DP]
    L.dh_zoltan_detection2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, T6, DP, DP, C.c_int,
                                       DP, C.c_int, C.POINTER(C.c_int), DP]
    for f in ("dh_zoltan_bf", "dh_zoltan_penalty", "dh_zoltan_hybrid", "dh_zoltan_detection",
              "dh_zoltan_detection2"):
        getattr(L, f).restype = C.c_int
    _bound = True


def _rows(a):
    """(3, n) DeviceArray -> table of its three row addresses."""
    n = a.shape[1]
    item = np.dtype(a.dtype).itemsize
    return TABLE3(*(a.ptr + k * n * item for k in range(3)))


def _vec3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def bf(soa, n, P, Q, d=None, flags=None, stream=None):
    """zoltan/bf.cpp:60-952 over n pairs in the dh_tri_pairs layout."""
    _bind()
    _lib.check(_lib.lib().dh_zoltan_bf(soa.ptr, n, P.ptr, Q.ptr, d.ptr if d is not None else None,
                                       flags.ptr if flags is not None else None, stream),
               "dh_zoltan_bf")


def penalty(it, nt, a, b, c, d, e, f, p, q, stream=None):
    """zoltan/penalty.cpp:27-208."""
    _bind()
    _lib.check(_lib.lib().dh_zoltan_penalty(it, nt, _vec3(a), _vec3(b), _vec3(c), _rows(d),
                                            _rows(e), _rows(f), _rows(p), _rows(q), stream),
               "dh_zoltan_penalty")


def hybrid(it, nt, a, b, c, d, e, f, p, q, stream=None):
    """zoltan/hybrid.cpp:37-132 (serial semantics)."""
    _bind()
    _lib.check(_lib.lib().dh_zoltan_hybrid(it, nt, _vec3(a), _vec3(b), _vec3(c), _rows(d),
                                           _rows(e), _rows(f), _rows(p), _rows(q), stream),
               "dh_zoltan_hybrid")


def _t_table(t):
    """t: list of 3 (or 6) DeviceArrays of shape (3, n) (vertex v: rows x, y, z)."""
    T = ((C.c_void_p * 3) * 6)()
    for v, arr in enumerate(t[:3]):
        rows = _rows(arr)
        for ax in range(3):
            T[v][ax] = rows[ax]
    return T


def detection(s, e, t, tid, pid, npid, s2=None, e2=None, cap=1 << 16, stream=None):
    """zoltan/contact.cpp:56-146 (all-to-all over [s, e)) or, with s2/e2,
    149-240 ([s, e) against [s2, e2)).  Returns the kept contacts in
    discovery order as a structured array (CONTACT_DTYPE)."""
    _bind()
    L = _lib.lib()
    while True:
        out = DeviceArray((max(cap, 1) * CONTACT_DTYPE.itemsize,), np.uint8)
        n = C.c_int(0)
        T = _t_table(t)
        tptr = tid.ptr if tid is not None else None
        if s2 is None:
            rc = L.dh_zoltan_detection(s, e, T, tptr, pid.ptr, npid, out.ptr, cap, C.byref(n),
                                       stream)
        else:
            rc = L.dh_zoltan_detection2(s, e, s2, e2, T, tptr, pid.ptr, npid, out.ptr, cap,
                                        C.byref(n), stream)
        if rc == _lib.DH_ERR_CAPACITY:
            cap = n.value
            continue
        _lib.check(rc, "dh_zoltan_detection")
        raw = out.download().tobytes()[:n.value * CONTACT_DTYPE.itemsize]
        return np.frombuffer(raw, dtype=CONTACT_DTYPE).copy()
