"""Launch-parameter table and autotuner of libdelta_hip (dh_tuning_*,
dh_autotune_*).

The table's knobs (persistent workgroups per CU of the learned and mixed
passes, the FP64 per-pair block size, the row tiling of particle-pair
detection) change how work is spread over the GPU and never the results.
`autotune_pairs` / `autotune_detect` time every candidate on the caller's
own device-resident inputs and keep the fastest; `save` / `load` keep a
table as JSON next to the workload it was tuned on."""
import ctypes as C
import json

from . import _lib

KNOBS = {0: "learned_wg_per_cu", 1: "mixed_wg_per_cu", 2: "exact_block",
         3: "detect_tile_target", 4: "detect_max_tiles", 5: "detect_wg_per_cu",
         6: "detect_queue_min_items", 7: "detect_fp32_screen"}


class DhTuning(C.Structure):
    _fields_ = [("learned_wg_per_cu", C.c_int32), ("mixed_wg_per_cu", C.c_int32),
                ("exact_block", C.c_int32), ("detect_tile_target", C.c_int32),
                ("detect_max_tiles", C.c_int32), ("detect_wg_per_cu", C.c_int32),
                ("detect_queue_min_items", C.c_int32), ("detect_fp32_screen", C.c_int32)]


class DhTuneTrial(C.Structure):
    _fields_ = [("knob", C.c_int32), ("value", C.c_int32), ("ms", C.c_float)]


_bound = False


def _L():
    global _bound
    L = _lib.lib()
    if not _bound:
        for name in ("dh_tuning_default", "dh_tuning_get"):
            getattr(L, name).argtypes = [C.POINTER(DhTuning)]
            getattr(L, name).restype = C.c_int
        L.dh_tuning_set.argtypes = [C.POINTER(DhTuning)]
        L.dh_tuning_set.restype = C.c_int
        L.dh_autotune_tri_pairs.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                            C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int, C.POINTER(DhTuneTrial), C.c_int,
                                            C.POINTER(C.c_int), C.c_void_p]
        L.dh_autotune_tri_pairs.restype = C.c_int
        L.dh_autotune_detect.argtypes = [C.c_int, C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_int, C.c_int, C.POINTER(DhTuneTrial),
                                         C.c_int, C.POINTER(C.c_int), C.c_void_p]
        L.dh_autotune_detect.restype = C.c_int
        _bound = True
    return L


def _as_dict(t):
    return {name: int(getattr(t, name)) for name, _ in DhTuning._fields_}


def default():
    t = DhTuning()
    _lib.check(_L().dh_tuning_default(C.byref(t)), "dh_tuning_default")
    return _as_dict(t)


def get():
    t = DhTuning()
    _lib.check(_L().dh_tuning_get(C.byref(t)), "dh_tuning_get")
    return _as_dict(t)


def set(**values):
    """Sets the named knobs, keeps the others; raises DhError on a value out
    of range."""
    cur = get()
    unknown = [k for k in values if k not in cur]
    if unknown:
        raise KeyError(f"unknown knobs {unknown}")
    cur.update(values)
    t = DhTuning(**cur)
    _lib.check(_L().dh_tuning_set(C.byref(t)), "dh_tuning_set")
    return get()


def _trials(buf, n):
    return [{"knob": KNOBS[buf[i].knob], "value": int(buf[i].value), "ms": float(buf[i].ms)}
            for i in range(n)]


def autotune_pairs(batch, method, eps_total=2e-4, reps=3, stream=None):
    """Tunes the knob of `method` on a narrow.PairBatch; returns the trials.
    The batch's outputs then hold the chosen configuration's results."""
    from .narrow import METHODS
    m = METHODS[method] if isinstance(method, str) else method
    buf = (DhTuneTrial * 32)()
    n = C.c_int(0)
    s32 = batch.soa32.ptr if batch.soa32 is not None else None
    _lib.check(_L().dh_autotune_tri_pairs(m, batch.soa.ptr, s32, batch.n, eps_total,
                                          batch.P.ptr, batch.Q.ptr, batch.d.ptr, batch.flags.ptr,
                                          reps, buf, len(buf), C.byref(n), stream),
               f"dh_autotune_tri_pairs({method})")
    return _trials(buf, n.value)


def autotune_detect(geom, pairs_dev, npairs, out_dev, cap, method=_lib.DH_BF, reps=3,
                    stream=None):
    """Tunes the detection tiling on device pairs (an int32 DeviceArray of
    2*npairs particle indices); `out_dev` must hold `cap` >= every contact."""
    buf = (DhTuneTrial * 32)()
    n = C.c_int(0)
    _lib.check(_L().dh_autotune_detect(method, C.byref(geom), pairs_dev, npairs, out_dev, cap,
                                       reps, buf, len(buf), C.byref(n), stream),
               "dh_autotune_detect")
    return _trials(buf, n.value)


def save(path, workload=None):
    with open(path, "w") as f:
        json.dump({"workload": workload, "tuning": get()}, f, indent=1)


def load(path):
    with open(path) as f:
        return set(**json.load(f)["tuning"])
