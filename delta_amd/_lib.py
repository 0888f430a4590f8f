"""ctypes binding of libdelta_hip.so (include/delta_hip.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdelta_hip.so")

DH_OK = 0
DH_ERR_CAPACITY = 3

# particle-pair methods
DH_BF, DH_PENALTY, DH_HYBRID_BATCH, DH_HYBRID_TRIANGLE = 0, 1, 2, 3
DH_PENALTY_STAT, DH_HYBRID_BATCH_STAT, DH_HYBRID_TRIANGLE_STAT, DH_GJK = 4, 5, 6, 7
DH_BF_LEARNED = 8
# unique-pair methods
DH_PAIR_BF64, DH_PAIR_PENALTY64, DH_PAIR_HYBRID64 = 0, 1, 2
DH_PAIR_MIXED, DH_PAIR_LEARNED, DH_PAIR_PENALTY_ADAPT64 = 3, 4, 5
DH_PAIR_BF64_EXACT, DH_PAIR_LEARNED_PENALTY, DH_PAIR_PENALTY_CERT = 6, 7, 8
DH_FLAG_FAILED, DH_FLAG_RECHECK, DH_FLAG_EDGE, DH_FLAG_CERTIFIED, DH_FLAG_INTERSECT = 1, 2, 4, 8, 16

VP = C.c_void_p


class DhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"delta_hip error {code}: {msg}")
        self.code = code


class DhContact(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("normal", C.c_double * 3), ("depth", C.c_double),
                ("P", C.c_double * 3), ("Q", C.c_double * 3), ("eps_total", C.c_double),
                ("master", C.c_int32), ("slave", C.c_int32), ("penetration", C.c_int32),
                ("friction", C.c_int32), ("triA", C.c_int32), ("triB", C.c_int32),
                ("pair", C.c_int32), ("pad_", C.c_int32)]


class DhStats(C.Structure):
    _fields_ = [("newton_histogram", C.c_int64 * 17), ("penalty_fails", C.c_int64),
                ("batch_fails", C.c_int64), ("batch_size", C.c_int64),
                ("batch_error", C.c_double), ("triangle_checks", C.c_int64),
                ("pairs_evaluated", C.c_int64), ("fallback_rows", C.c_int64),
                ("pairs_culled", C.c_int64)]


class DhGenParams(C.Structure):
    _fields_ = [("box", C.c_double), ("sep_min", C.c_double), ("sep_max", C.c_double),
                ("frac_parallel", C.c_double), ("frac_intersect", C.c_double),
                ("max_angle", C.c_double), ("bisection_steps", C.c_int)]


class DhPairCounters(C.Structure):
    _fields_ = [("screened", C.c_int64), ("rechecked", C.c_int64),
                ("features_fp64", C.c_int64), ("certified", C.c_int64)]


class DhGeometry(C.Structure):
    _fields_ = [("x", VP), ("y", VP), ("z", VP), ("tri_offset", VP), ("eps", VP),
                ("friction", VP), ("id", VP), ("nparticles", C.c_int), ("ntri_total", C.c_int)]


class DhBodies(C.Structure):
    _fields_ = [("pos", VP), ("refpos", VP), ("linvel", VP), ("angvel", VP),
                ("refangvel", VP), ("rot", VP), ("inertia", VP), ("inverse", VP),
                ("mass", VP), ("obstacle", VP), ("material", VP), ("n", C.c_int)]


class DhForceParams(C.Structure):
    _fields_ = [("spring", C.c_double), ("damper", C.c_double), ("friction", C.c_double),
                ("is_sphere", C.c_int)]


assert C.sizeof(DhContact) == 144

_SIGS = {
    "dh_init": (C.c_int, []),
    "dh_last_error": (C.c_char_p, []),
    "dh_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "dh_malloc": (C.c_int, [C.POINTER(VP), C.c_size_t]),
    "dh_free": (C.c_int, [VP]),
    "dh_memcpy_h2d": (C.c_int, [VP, VP, C.c_size_t]),
    "dh_memcpy_d2h": (C.c_int, [VP, VP, C.c_size_t]),
    "dh_memset": (C.c_int, [VP, C.c_int, C.c_size_t]),
    "dh_synchronize": (C.c_int, []),
    "dh_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "dh_event_create": (C.c_int, [C.POINTER(VP)]),
    "dh_event_destroy": (C.c_int, [VP]),
    "dh_event_record": (C.c_int, [VP, VP]),
    "dh_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), VP, VP]),
    "dh_tri_pairs": (C.c_int, [C.c_int, VP, VP, C.c_int64, C.c_double, VP, VP, VP, VP, VP]),
    "dh_tri_pairs_counters": (C.c_int, [C.POINTER(DhPairCounters)]),
    "dh_gen_tri_pairs": (C.c_int, [C.c_uint64, C.c_int64, C.POINTER(DhGenParams), VP, VP, VP,
                                   VP]),
    "dh_convert_f32": (C.c_int, [VP, VP, C.c_int64, VP]),
}

# entry points declared in include/delta_hip.h that are looked up lazily (the
# test of exported symbols reads the header itself)
_lib = None


def lib():
    """Load libdelta_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m delta_amd.build` "
                "(the HIP library is required: there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().dh_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != DH_OK:
        raise DhError(rc, f"{what}: {last_error()}")
    return rc


def init():
    check(lib().dh_init(), "dh_init")
