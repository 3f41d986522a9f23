"""The C-ABI library loads and exports every entry point include/wfe_hip.h
declares (no GPU calls)."""
import re
from pathlib import Path

import wfe_amd
from wfe_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "wfe_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wfe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "wfe_create" in names and "wfe_step" in names and "wfe_case_load" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = wfe_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_abi_version():
    assert wfe_amd.lib().wfe_abi_version() == 1


def test_errors_are_reported_not_raised():
    L = wfe_amd.lib()
    assert L.wfe_step(None, 1.0, 0.0) != 0
    assert b"null" in L.wfe_last_error()


def test_oversize_grid_is_rejected_before_any_array_is_read():
    """Every kernel addresses a 3-D array with a 32-bit byte offset, so an FP64
    field in the device layout must stay below 4 GiB (2^29 padded points).
    The size check runs before the base-state/metric arrays are looked at,
    so null arrays are fine here and no GPU call is made."""
    import ctypes as C
    L = wfe_amd.lib()
    g = _lib.wfe_grid(2000, 1000, 300, 3, 1000.0, 1000.0, 100.0)  # pitch 2048 -> 6.3e8 points
    p = _lib.wfe_params()
    b = _lib.wfe_base()
    m = _lib.wfe_metric()
    h = C.c_void_p()
    rc = L.wfe_create(C.byref(g), C.byref(p), C.byref(b), C.byref(m), 0, C.byref(h))
    assert rc == -1 and not h.value
    assert b"4 GiB" in L.wfe_last_error()
    # just below the limit: the size check passes and the null arrays are reported
    g = _lib.wfe_grid(1500, 1000, 300, 3, 1000.0, 1000.0, 100.0)  # pitch 1536 -> 4.7e8
    rc = L.wfe_create(C.byref(g), C.byref(p), C.byref(b), C.byref(m), 0, C.byref(h))
    assert rc == -1 and b"missing" in L.wfe_last_error()
