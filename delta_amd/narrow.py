"""Unique triangle-pair narrow phase (the isolated narrow-phase workload):
seeded generation in HBM and the per-pair kernels of dh_tri_pairs.

Layout (include/delta_hip.h): soa[k*n + i], k = 0..2 x of A's vertices,
3..5 y of A, 6..8 z of A, 9..17 the same for B; outputs P[c*n+i], Q[c*n+i],
d[i], flags[i]."""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

METHODS = {
    "bf64": _lib.DH_PAIR_BF64,
    "penalty64": _lib.DH_PAIR_PENALTY64,
    "hybrid64": _lib.DH_PAIR_HYBRID64,
    "mixed": _lib.DH_PAIR_MIXED,
    "learned": _lib.DH_PAIR_LEARNED,
    "penalty_adapt64": _lib.DH_PAIR_PENALTY_ADAPT64,
    "bf64_exact": _lib.DH_PAIR_BF64_EXACT,
    "learned_penalty": _lib.DH_PAIR_LEARNED_PENALTY,
    "penalty_cert": _lib.DH_PAIR_PENALTY_CERT,
}

# algorithmic cost per pair (operation counts of the reference's arithmetic,
# sqrt and division counted as one flop; bytes = 18 inputs + P, Q, d + flag)
FLOP_BF = 1212
FLOP_PENALTY4 = 876
BYTES_F64 = 18 * 8 + 7 * 8 + 1
BYTES_F32_BULK = 18 * 4 + 7 * 4 + 1


def default_gen_params(**kw):
    p = _lib.DhGenParams(1e-3, 1e-7, 1e-2, 0.05, 0.01, 1e-4, 48)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class PairBatch:
    """n unique pairs resident in HBM plus output buffers."""

    def __init__(self, n, seed=3, params=None, with_f32=False):
        self.n = int(n)
        L = _lib.lib()
        _lib.init()
        self.soa = DeviceArray((18, self.n), np.float64)
        self.target = DeviceArray((self.n,), np.float64)
        self.kind = DeviceArray((self.n,), np.uint8)
        p = params or default_gen_params()
        _lib.check(L.dh_gen_tri_pairs(seed, self.n, C.byref(p), self.soa.ptr, self.target.ptr,
                                      self.kind.ptr, None), "dh_gen_tri_pairs")
        self.soa32 = None
        if with_f32:
            self.make_f32()
        self.P = DeviceArray((3, self.n), np.float64)
        self.Q = DeviceArray((3, self.n), np.float64)
        self.d = DeviceArray((self.n,), np.float64)
        self.flags = DeviceArray((self.n,), np.uint8)

    @classmethod
    def from_host(cls, soa):
        self = cls.__new__(cls)
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        assert soa.shape[0] == 18
        self.n = soa.shape[1]
        _lib.init()
        self.soa = DeviceArray.from_host(soa)
        self.target = None
        self.kind = None
        self.soa32 = None
        self.P = DeviceArray((3, self.n), np.float64)
        self.Q = DeviceArray((3, self.n), np.float64)
        self.d = DeviceArray((self.n,), np.float64)
        self.flags = DeviceArray((self.n,), np.uint8)
        return self

    def make_f32(self):
        if self.soa32 is None:
            self.soa32 = DeviceArray((18, self.n), np.float32)
            _lib.check(_lib.lib().dh_convert_f32(self.soa.ptr, self.soa32.ptr, 18 * self.n,
                                                 None), "dh_convert_f32")
        return self.soa32

    def run(self, method, eps_total=2e-4, stream=None):
        m = METHODS[method] if isinstance(method, str) else method
        s32 = self.soa32.ptr if self.soa32 is not None else None
        _lib.check(_lib.lib().dh_tri_pairs(m, self.soa.ptr, s32, self.n, eps_total, self.P.ptr,
                                           self.Q.ptr, self.d.ptr, self.flags.ptr, stream),
                   f"dh_tri_pairs({method})")

    def results(self):
        return self.P.download(), self.Q.download(), self.d.download(), self.flags.download()
