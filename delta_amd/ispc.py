"""The reference's ISPC kernels ispc_bf (delta/contact/detection/bf.ispc:3-6)
and ispc_pen (penalty.ispc:154-174) over the C ABI: the same names and
argument meaning, with device arrays (DeviceArray of shape (3, n): row k is
coordinate k, the reference's `pp1[k][i]`).  No CPU fallback."""
import ctypes as C

from . import _lib
from .zoltan import TABLE3, _rows

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = _lib.lib()
    for f in ("dh_ispc_bf", "dh_ispc_pen"):
        getattr(L, f).argtypes = [C.c_uint] + [TABLE3] * 8 + [C.c_void_p, C.c_void_p]
        getattr(L, f).restype = C.c_int
    _bound = True


def _call(name, n, pp, P, Q, distance, stream):
    _bind()
    if len(pp) != 6:
        raise ValueError("six vertex arrays (pp1..pp6)")
    _lib.check(getattr(_lib.lib(), name)(n, *(_rows(a) for a in pp), _rows(P), _rows(Q),
                                         distance.ptr, stream), name)


def ispc_bf(n, pp, P, Q, distance, stream=None):
    """bf.ispc:3-937: pp = [pp1..pp6] (triangle A = pp1..pp3, B = pp4..pp6).
    P, Q and distance are in/out as in the reference (see dh_ispc_bf)."""
    _call("dh_ispc_bf", n, pp, P, Q, distance, stream)


def ispc_pen(n, pp, P, Q, distance, stream=None):
    """penalty.ispc:154-174: five-step penalty, distance = |P - Q|."""
    _call("dh_ispc_pen", n, pp, P, Q, distance, stream)
