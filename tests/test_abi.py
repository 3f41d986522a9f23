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
