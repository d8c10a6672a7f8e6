"""The C-ABI library loads and exports every function include/*.h declares (CPU only)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    names = set()
    for h in ("c_superintervals.h", "si_device.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"^typedef[^;]*;", "", txt, flags=re.M)   # function-pointer types
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "defined"):
                names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import superintervals_amd._native as N
    names = declared_functions()
    assert "indexSuperIntervals" in names and "siSearchFill" in names
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header():
    import superintervals_amd._native as N
    assert set(declared_functions()) <= set(N.exported_symbols()) | {
        "createIndexResult", "createKeyResult", "createItemResult"}


def test_no_gpu_errors_cleanly():
    """Without a GPU, creating an index fails with an error instead of a crash."""
    import superintervals_amd._native as N
    try:
        import torch
        if torch.cuda.is_available():
            return
    except ImportError:
        pass
    h = N.lib.siIndexCreate(4)
    assert not h
    assert b"device" in N.lib.siGetLastError()


def test_tuning_knobs_round_trip():
    """Every knob the autotuner saves and restores reaches the library (host-side state,
    no GPU call), and an unknown knob name is an error rather than silently ignored."""
    import pytest
    from superintervals_amd import autotune
    saved = autotune.current()
    assert set(saved) == set(autotune.KNOBS)
    assert "lane_steps" in saved and saved["lane_steps"] >= 1
    try:
        autotune.apply({"lane_steps": 3, "tile": 512})
        now = autotune.current()
        assert now["lane_steps"] == 3 and now["tile"] == 512
        with pytest.raises(KeyError):
            autotune.apply({"lane_step": 3})
    finally:
        autotune.apply(saved)
    assert autotune.current() == saved
