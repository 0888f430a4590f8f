"""delta::contact::detection mirror over the HIP library.

Each function has the reference's name, argument order and meaning
(delta/contact/detection/bf.h:65-81, penalty.h:62-115, hybrid.h:57-170):
coordinates are flat per-axis arrays x[3t+k] of a particle's triangle soup,
`numberOfTrianglesOf...` is a triangle count, epsilon is the particle margin.
They return a numpy structured array of contact records (the contactpoint
fields plus triA, triB) in the reference's emission order.

The statistics the reference keeps in static globals (penalty.cpp:33-48,
hybrid.cpp:27-57) are kept module-wide here with the same getters; like the
reference's they are not thread-safe, and cleanHybridStatistics resets only
the fail counters (hybrid.cpp:36-41)."""
import ctypes as C

import numpy as np

from . import _lib

CONTACT_DTYPE = np.dtype([("x", "f8", 3), ("normal", "f8", 3), ("depth", "f8"), ("P", "f8", 3),
                          ("Q", "f8", 3), ("eps_total", "f8"), ("master", "i4"),
                          ("slave", "i4"), ("penetration", "i4"), ("friction", "i4"),
                          ("triA", "i4"), ("triB", "i4"), ("pair", "i4"), ("pad_", "i4")])
assert CONTACT_DTYPE.itemsize == 144

_stats = _lib.DhStats()


def _arr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _detect(method, xA, yA, zA, nA, epsA, frA, idA, xB, yB, zB, nB, epsB, frB, idB):
    L = _lib.lib()
    xs = [_arr(v) for v in (xA, yA, zA, xB, yB, zB)]
    for v, n in zip(xs, (nA, nA, nA, nB, nB, nB)):
        if v.size < 3 * n:
            raise ValueError("coordinate array shorter than 3 * numberOfTriangles")
    cap = max(16, min(int(nA) * int(nB), 1 << 16))
    st = _stats
    while True:
        buf = np.zeros(cap, CONTACT_DTYPE)
        n = C.c_int(0)
        saved = _lib.DhStats.from_buffer_copy(st)
        rc = L.dh_detect_pair(method, xs[0].ctypes.data, xs[1].ctypes.data, xs[2].ctypes.data,
                              int(nA), C.c_double(epsA), int(bool(frA)), int(idA),
                              xs[3].ctypes.data, xs[4].ctypes.data, xs[5].ctypes.data, int(nB),
                              C.c_double(epsB), int(bool(frB)), int(idB), buf.ctypes.data, cap,
                              C.byref(n), C.byref(st))
        if rc == _lib.DH_ERR_CAPACITY:
            C.memmove(C.addressof(st), C.addressof(saved), C.sizeof(st))
            cap = n.value
            continue
        _lib.check(rc, "dh_detect_pair")
        return buf[:n.value].copy()


def _sig(f):
    f.restype = C.c_int
    return f


def _bind():
    L = _lib.lib()
    L.dh_detect_pair.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                 C.c_int, C.POINTER(C.c_int), C.POINTER(_lib.DhStats)]
    L.dh_detect_pair.restype = C.c_int


def bf(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA, particleA,
       xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB, particleB):
    """bf.cpp:28-149: contact for every triangle pair with d <= epsA+epsB."""
    _bind()
    return _detect(_lib.DH_BF, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA,
                   particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB,
                   particleB)


def penalty(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA, particleA,
            xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB, particleB):
    """penalty.cpp:113-241: four fixed Newton steps, contact if d < eps."""
    _bind()
    return _detect(_lib.DH_PENALTY, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA,
                   frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                   frictionB, particleB)


def hybridWithPerBatchFallBack(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA,
                               particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                               frictionB, particleB):
    """hybrid.cpp:338-487: penalty per row of A; the row is redone with
    bfSolver when more than 3*nB/32 of its solves failed."""
    _bind()
    return _detect(_lib.DH_HYBRID_BATCH, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA,
                   frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                   frictionB, particleB)


def hybridWithPerTriangleFallBack(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA,
                                  particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                                  frictionB, particleB):
    """hybrid.cpp:59-200: any failed solve redoes its row; rows run to nTriB
    as in the reference (nTriB > nTriA raises: the reference reads out of
    bounds there)."""
    _bind()
    return _detect(_lib.DH_HYBRID_TRIANGLE, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA,
                   frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                   frictionB, particleB)


def penaltyStat(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA, particleA,
                xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB, particleB):
    """penalty.cpp:50-111: adaptive solver with the iteration histogram."""
    _bind()
    return _detect(_lib.DH_PENALTY_STAT, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA,
                   frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB,
                   frictionB, particleB)


def hybridBatchStat(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA, particleA,
                    xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB, particleB):
    """hybrid.cpp:489-627: row minimum only, batch statistics."""
    _bind()
    return _detect(_lib.DH_HYBRID_BATCH_STAT, xA, yA, zA, numberOfTrianglesOfGeometryA,
                   epsilonA, frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB,
                   epsilonB, frictionB, particleB)


def hybridTriangleStat(xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA, particleA,
                       xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB, particleB):
    """hybrid.cpp:202-336: row minimum only, triangle statistics."""
    _bind()
    return _detect(_lib.DH_HYBRID_TRIANGLE_STAT, xA, yA, zA, numberOfTrianglesOfGeometryA,
                   epsilonA, frictionA, particleA, xB, yB, zB, numberOfTrianglesOfGeometryB,
                   epsilonB, frictionB, particleB)


# statistics getters (penalty.h:47-53, hybrid.h:130-170)
def gjk(numberOfTrianglesOfGeometryA, xA, yA, zA, epsilonA, frictionA, particleA,
        numberOfTrianglesOfGeometryB, xB, yB, zB, epsilonB, frictionB, particleB):
    """gjk.cpp:25-100 (count-first argument order of gjk.h:32-48): GJK over
    every vertex of both particles, at most one contact, when |P-Q| <=
    epsA+epsB; its triangle indices are -1.  The reference's wrapper copies
    only nTri vertices and then reads 3*nTri points from that buffer, past its
    end; this one uses all 3*nTri vertices."""
    _bind()
    return _detect(_lib.DH_GJK, xA, yA, zA, numberOfTrianglesOfGeometryA, epsilonA, frictionA,
                   particleA, xB, yB, zB, numberOfTrianglesOfGeometryB, epsilonB, frictionB,
                   particleB)


def cleanPenaltyStatistics():
    for i in range(17):
        _stats.newton_histogram[i] = 0


def getPenaltyIterationStatistics():
    return [int(v) for v in _stats.newton_histogram]


def cleanHybridStatistics():
    _stats.penalty_fails = 0
    _stats.batch_fails = 0


def getPenaltyFails():
    return int(_stats.penalty_fails)


def getBatchFails():
    return int(_stats.batch_fails)


def getBatchSize():
    return int(_stats.batch_size)


def getBatchError():
    return float(_stats.batch_error)


def reset_all_statistics():
    """Zero every counter (the reference has no single call for this)."""
    C.memset(C.addressof(_stats), 0, C.sizeof(_stats))


def engine_counters():
    """Counters with no reference getter: nTriA*nTriB per call (the
    reference's tri-cmp, Engine.cpp:451), the reference's evaluations,
    fallback rows, and the brute-force pairs the box test skipped."""
    return {"triangle_checks": int(_stats.triangle_checks),
            "pairs_evaluated": int(_stats.pairs_evaluated),
            "fallback_rows": int(_stats.fallback_rows),
            "pairs_culled": int(_stats.pairs_culled)}
