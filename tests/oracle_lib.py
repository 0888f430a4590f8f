"""ctypes access to the CPU checkers: oracle/liboracle.so (the C restatement)
and oracle/_ref/libdelta_ref.so (the reference's own sources, built in the
development container only).  Test infrastructure; never imported by the
product package."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_OMP_SO = os.path.join(ROOT, "oracle", "liboracle_omp.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libdelta_ref.so")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class DhContact(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("normal", C.c_double * 3), ("depth", C.c_double),
                ("P", C.c_double * 3), ("Q", C.c_double * 3), ("eps_total", C.c_double),
                ("master", C.c_int32), ("slave", C.c_int32), ("penetration", C.c_int32),
                ("friction", C.c_int32), ("triA", C.c_int32), ("triB", C.c_int32),
                ("pair", C.c_int32), ("pad_", C.c_int32)]


assert C.sizeof(DhContact) == 144


class DhStats(C.Structure):
    _fields_ = [("newton_histogram", C.c_int64 * 17), ("penalty_fails", C.c_int64),
                ("batch_fails", C.c_int64), ("batch_size", C.c_int64),
                ("batch_error", C.c_double), ("triangle_checks", C.c_int64),
                ("pairs_evaluated", C.c_int64), ("fallback_rows", C.c_int64),
                ("pairs_culled", C.c_int64)]


class DhForceParams(C.Structure):
    _fields_ = [("spring", C.c_double), ("damper", C.c_double), ("friction", C.c_double),
                ("is_sphere", C.c_int)]


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        _oracle = C.CDLL(ORACLE_SO)
    return _oracle


_oracle_omp = None


def oracle_omp():
    """oracle/liboracle_omp.so: the same restatement with the OpenMP CPU
    baseline drivers (oracle/cpu_baseline.c)."""
    global _oracle_omp
    if _oracle_omp is None:
        if not os.path.exists(ORACLE_OMP_SO):
            raise RuntimeError("oracle/liboracle_omp.so is not built (make -C oracle oracle)")
        _oracle_omp = C.CDLL(ORACLE_OMP_SO)
    return _oracle_omp


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_detect.restype = C.c_int
    return _ref


def contacts_to_array(buf, n):
    """dh_contact list -> structured numpy copy."""
    dt = np.dtype([("x", "f8", 3), ("normal", "f8", 3), ("depth", "f8"), ("P", "f8", 3),
                   ("Q", "f8", 3), ("eps_total", "f8"), ("master", "i4"), ("slave", "i4"),
                   ("penetration", "i4"), ("friction", "i4"), ("triA", "i4"), ("triB", "i4"),
                   ("pair", "i4"), ("pad_", "i4")])
    assert dt.itemsize == 144
    return np.frombuffer(bytes(C.string_at(C.addressof(buf), 144 * n)), dtype=dt).copy()


def _pair_ptrs(tA, tB):
    """tA, tB: (3,3) arrays [vertex][axis] -> six pointers to 3 doubles."""
    a = np.ascontiguousarray(np.asarray(tA, dtype=np.float64).T)
    b = np.ascontiguousarray(np.asarray(tB, dtype=np.float64).T)
    return a, b, [dptr(a[0]), dptr(a[1]), dptr(a[2]), dptr(b[0]), dptr(b[1]), dptr(b[2])]


def or_bf_solver(tA, tB):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    edge = oracle().or_bf_solver(*ptrs, dptr(P), dptr(Q))
    return P, Q, edge


def ref_bf_solver(tA, tB):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    ref().ref_bf_solver(*ptrs, dptr(P), dptr(Q))
    return P, Q


def or_penalty(tA, tB, max_err):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    f = C.c_int(0)
    oracle().or_penalty_solver(*ptrs, dptr(P), dptr(Q), C.c_double(max_err), C.byref(f))
    return P, Q, f.value


def ref_penalty(tA, tB, max_err):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    f = C.c_int(0)
    ref().ref_penalty_solver(*ptrs, dptr(P), dptr(Q), C.c_double(max_err), C.byref(f))
    return P, Q, f.value


def or_penalty_adaptive(tA, tB, max_err):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    its = C.c_int(-1)
    oracle().or_penalty_solver_adaptive(*ptrs, dptr(P), dptr(Q), C.c_double(max_err),
                                        C.byref(its))
    return P, Q, its.value


def ref_penalty_adaptive(tA, tB, max_err):
    a, b, ptrs = _pair_ptrs(tA, tB)
    P = np.zeros(3)
    Q = np.zeros(3)
    its = C.c_int(-1)
    ref().ref_penalty_solver_adaptive(*ptrs, dptr(P), dptr(Q), C.c_double(max_err),
                                      C.byref(its))
    return P, Q, its.value


def _mesh_ptrs(m):
    """m: (ntri, 3, 3) [tri][vertex][axis] -> x, y, z arrays of 3*ntri."""
    m = np.asarray(m, dtype=np.float64)
    xs = [np.ascontiguousarray(m[:, :, k].reshape(-1)) for k in range(3)]
    return xs


def or_detect(method, mA, mB, epsA, epsB, fricA=0, fricB=0, idA=0, idB=1, cap=None):
    xa = _mesh_ptrs(mA)
    xb = _mesh_ptrs(mB)
    nA, nB = len(mA), len(mB)
    cap = cap if cap is not None else max(1, nA * nB)
    buf = (DhContact * cap)()
    n = C.c_int(0)
    st = DhStats()
    rc = oracle().or_detect(method, dptr(xa[0]), dptr(xa[1]), dptr(xa[2]), nA,
                            C.c_double(epsA), fricA, idA, dptr(xb[0]), dptr(xb[1]),
                            dptr(xb[2]), nB, C.c_double(epsB), fricB, idB, buf, cap,
                            C.byref(n), C.byref(st))
    return rc, contacts_to_array(buf, min(n.value, cap)), st


def ref_detect(method, mA, mB, epsA, epsB, fricA=0, fricB=0, idA=0, idB=1):
    xa = _mesh_ptrs(mA)
    xb = _mesh_ptrs(mB)
    nA, nB = len(mA), len(mB)
    cap = max(1, nA * nB)
    buf = (DhContact * cap)()
    r = ref()
    r.ref_clean_stats()
    h0 = (C.c_longlong * 17)()
    pf0, bf0, bs0 = C.c_longlong(), C.c_longlong(), C.c_longlong()
    be0 = C.c_double()
    # cleanHybridStatistics leaves _batchSize and _batchError untouched
    # (hybrid.cpp:36-41): record them before the call
    r.ref_get_stats(h0, C.byref(pf0), C.byref(bf0), C.byref(bs0), C.byref(be0))
    n = r.ref_detect(method, dptr(xa[0]), dptr(xa[1]), dptr(xa[2]), nA, C.c_double(epsA),
                     fricA, idA, dptr(xb[0]), dptr(xb[1]), dptr(xb[2]), nB, C.c_double(epsB),
                     fricB, idB, buf, cap)
    h = (C.c_longlong * 17)()
    pf, bfl, bs = C.c_longlong(), C.c_longlong(), C.c_longlong()
    be = C.c_double()
    r.ref_get_stats(h, C.byref(pf), C.byref(bfl), C.byref(bs), C.byref(be))
    stats = dict(hist=list(h), penalty_fails=pf.value, batch_fails=bfl.value,
                 batch_size=bs.value, batch_error=be.value, batch_size_before=bs0.value,
                 batch_error_before=be0.value)
    return contacts_to_array(buf, n), stats


class DhBodiesHost(C.Structure):
    _fields_ = [("pos", DP), ("refpos", DP), ("linvel", DP), ("angvel", DP),
                ("refangvel", DP), ("rot", DP), ("inertia", DP), ("inverse", DP),
                ("mass", DP), ("obstacle", IP), ("material", IP), ("n", C.c_int)]


class OrScene(C.Structure):
    _fields_ = [("n", C.c_int), ("tri_offset", IP), ("x", DP), ("y", DP), ("z", DP),
                ("rx", DP), ("ry", DP), ("rz", DP), ("eps", DP), ("friction", IP),
                ("id", IP), ("b", DhBodiesHost)]


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(IP)


def or_step(hs, method, spring, damper, friction, dt, gravity, omp_threads=None):
    """One oracle step on a delta_amd.engine.HostScene, in place.  Returns
    (contacts, checks, candidate pairs).  With omp_threads: the
    reference-shaped OpenMP step of oracle/cpu_baseline.c (brute force only,
    rows of each particle pair in parallel; a timing driver)."""
    for f in ("x", "y", "z", "rx", "ry", "rz", "eps", "pos", "refpos", "linvel", "angvel",
              "refangvel", "rot", "inertia", "inverse", "mass"):
        setattr(hs, f, np.ascontiguousarray(getattr(hs, f), dtype=np.float64))
    b = DhBodiesHost(dptr(hs.pos), dptr(hs.refpos), dptr(hs.linvel), dptr(hs.angvel),
                     dptr(hs.refangvel), dptr(hs.rot), dptr(hs.inertia), dptr(hs.inverse),
                     dptr(hs.mass), _ip(hs.obstacle), _ip(hs.material), hs.n)
    sc = OrScene(hs.n, _ip(hs.tri_offset), dptr(hs.x), dptr(hs.y), dptr(hs.z), dptr(hs.rx),
                 dptr(hs.ry), dptr(hs.rz), dptr(hs.eps), _ip(hs.friction), _ip(hs.id), b)
    fp = DhForceParams(spring, damper, friction, 0)
    g = (C.c_double * 3)(*gravity)
    checks = C.c_int64(0)
    npairs = C.c_int(0)
    if omp_threads:
        L = oracle_omp()
        L.or_step_rows_omp.restype = C.c_int64
        nc = L.or_step_rows_omp(C.byref(sc), C.byref(fp), C.c_double(dt), g, C.byref(checks),
                                C.byref(npairs), C.c_int(omp_threads))
        return nc, checks.value, npairs.value
    L = oracle()
    L.or_step.restype = C.c_int64
    nc = L.or_step(C.byref(sc), method, C.byref(fp), C.c_double(dt), g, C.byref(checks),
                   C.byref(npairs))
    return nc, checks.value, npairs.value


def or_forces_scene(hs, pairs, contacts_by_pair, spring, damper, friction):
    """Per-particle force/torque and touched flags from per-pair contact
    lists, ascending partner order (the step's accumulation rule)."""
    n = hs.n
    fp = DhForceParams(spring, damper, friction, 0)
    L = oracle()
    per = []
    for k, (a, b) in enumerate(pairs):
        cs = contacts_by_pair[k]
        if len(cs) == 0:
            continue
        buf = (DhContact * len(cs)).from_buffer_copy(np.ascontiguousarray(cs).tobytes())
        fA = np.zeros(3); tA = np.zeros(3); fB = np.zeros(3); tB = np.zeros(3)
        pos, ang, lin = hs.pos, hs.angvel, hs.linvel
        L.or_contacts_forces(buf, len(cs), 0, dptr(np.ascontiguousarray(pos[a])),
                             dptr(np.ascontiguousarray(ang[a])), dptr(np.ascontiguousarray(lin[a])),
                             C.c_double(hs.mass[a]), dptr(np.ascontiguousarray(pos[b])),
                             dptr(np.ascontiguousarray(ang[b])), dptr(np.ascontiguousarray(lin[b])),
                             C.c_double(hs.mass[b]), C.byref(fp), dptr(fA), dptr(tA))
        L.or_contacts_forces(buf, len(cs), 1, dptr(np.ascontiguousarray(pos[b])),
                             dptr(np.ascontiguousarray(ang[b])), dptr(np.ascontiguousarray(lin[b])),
                             C.c_double(hs.mass[b]), dptr(np.ascontiguousarray(pos[a])),
                             dptr(np.ascontiguousarray(ang[a])), dptr(np.ascontiguousarray(lin[a])),
                             C.c_double(hs.mass[a]), C.byref(fp), dptr(fB), dptr(tB))
        per.append((a, b, fA, tA, fB, tB))
    force = np.zeros((n, 3)); torque = np.zeros((n, 3)); touched = np.zeros(n, np.int32)
    for p in range(n):
        contrib = [(a, fB, tB) for (a, b, fA, tA, fB, tB) in per if b == p]
        contrib += [(b, fA, tA) for (a, b, fA, tA, fB, tB) in per if a == p]
        contrib.sort(key=lambda c: c[0])
        for _, f, t in contrib:
            force[p] = force[p] + f
            torque[p] = torque[p] + t
        touched[p] = 1 if contrib else 0
    return force, torque, touched
