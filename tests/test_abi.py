"""libdelta_hip.so loads and exports every entry point include/delta_hip.h
declares (no compute: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "delta_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dh_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "dh_tri_pairs" in names and "dh_detect_pair" in names and "dh_step" in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "delta_amd", "libdelta_hip.so")
    if not os.path.exists(path):
        pytest.fail("delta_amd/libdelta_hip.so is not built (python -m delta_amd.build)")
    lib = C.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_package_import_does_not_touch_gpu():
    import delta_amd
    assert hasattr(delta_amd, "lib")


def test_tuning_table_validates_without_gpu():
    """dh_tuning_* touch no device: defaults, a round trip and rejection of
    out-of-range values."""
    from delta_amd import _lib, tuning
    d = tuning.default()
    assert d == {"learned_wg_per_cu": 6, "mixed_wg_per_cu": 8, "exact_block": 256,
                 "detect_tile_target": 4096, "detect_max_tiles": 16, "detect_wg_per_cu": 0,
                 "detect_queue_min_items": 32768, "detect_fp32_screen": 0}
    try:
        assert tuning.set(exact_block=128, learned_wg_per_cu=3)["exact_block"] == 128
        assert tuning.get()["learned_wg_per_cu"] == 3
        for bad in ({"exact_block": 100}, {"exact_block": 320}, {"mixed_wg_per_cu": 0},
                    {"learned_wg_per_cu": 33}, {"detect_max_tiles": 65},
                    {"detect_tile_target": 0}, {"detect_wg_per_cu": -1},
                    {"detect_wg_per_cu": 65}, {"detect_queue_min_items": -1},
                    {"detect_fp32_screen": 2}):
            with pytest.raises(_lib.DhError):
                tuning.set(**bad)
        assert tuning.get()["exact_block"] == 128  # a rejected set changes nothing
        with pytest.raises(KeyError):
            tuning.set(block=64)
    finally:
        tuning.set(**d)
    assert tuning.get() == d


def test_noalloc_broadphase_scratch_fits_its_limits():
    """The fixed array of dh_broadphase_noalloc holds the buffers of the
    largest accepted broad phase exactly, and every smaller one (host-only
    layout arithmetic, no device)."""
    path = os.path.join(ROOT, "delta_amd", "libdelta_hip.so")
    lib = C.CDLL(path)
    f = lib.dh_broadphase_noalloc_scratch_bytes
    f.restype = C.c_size_t
    f.argtypes = [C.c_int, C.c_int]
    full = f(16384, 1 << 20)
    assert full == 8523264
    for n, t in [(2, 2), (65, 5122), (1000, 80000), (4097, 1 << 20), (16384, 0), (8191, 999999)]:
        assert 0 < f(n, t) <= full, (n, t)
