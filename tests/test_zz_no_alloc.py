"""The allocation-free diagnostics entry point (wfe_diagnose_fields_noalloc)
returns exactly what the allocating one (wfe_diagnose_fields) returns, call
after call on the shared per-GPU workspace, and rejects grids beyond its
fixed workspace before touching the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from test_output_path import host_diagnose


def _call(L, fn, gr, prec, dev, moist):
    from wfe_amd import _lib
    out = _lib.wfe_diag()
    rc = getattr(L, fn)(C.byref(gr), prec, dev["w"].ptr, dev["thp"].ptr, dev["pip"].ptr,
                        dev["qc"].ptr if moist else None, dev["qr"].ptr if moist else None,
                        None, C.byref(out))
    _lib.check(rc, fn)
    return {f: getattr(out, f) for f, _ in _lib.wfe_diag._fields_}


@pytest.mark.gpu
@pytest.mark.parametrize("precision,moist", [("mixed", True), ("fp64", True), ("fp64", False)])
def test_noalloc_equals_the_allocating_entry_point_on_consecutive_calls(precision, moist):
    from wfe_amd import Grid, _lib
    L = _lib.lib()
    rng = np.random.default_rng(23)
    # second call smaller than the first: stale records of the first call
    # beyond the second one's workgroups must not leak into its result
    for nx, ny, nz in ((150, 90, 40), (37, 13, 9)):
        g = Grid(nx, ny, nz, 3, 1000.0, 1000.0, 100.0)
        scale = 10.0 if nx == 150 else 1.0  # the larger grid holds the larger extremes
        ft = np.float64 if precision == "fp64" else np.float32
        host = {f: (rng.standard_normal(g.npts) * s * scale).astype(
                    np.float64 if f == "pip" else ft)
                for f, s in (("w", 3.0), ("thp", 2.0), ("pip", 1e-3), ("qc", 1e-3),
                             ("qr", 1e-3))}
        dev = {f: _lib.DeviceArray(a) for f, a in host.items()}
        gr = _lib.wfe_grid(nx, ny, nz, 3, 1000.0, 1000.0, 100.0)
        prec = _lib.WFE_PREC_FP64 if precision == "fp64" else _lib.WFE_PREC_MIXED
        a = _call(L, "wfe_diagnose_fields", gr, prec, dev, moist)
        b = _call(L, "wfe_diagnose_fields_noalloc", gr, prec, dev, moist)
        host = {f: d.numpy().astype(np.float64) for f, d in dev.items()}
        print(precision, moist, (nx, ny, nz), b, flush=True)
        assert b == a
        assert b == host_diagnose(g, host, moist)


def test_noalloc_rejects_a_grid_beyond_its_workspace():
    """ceil(2000/64) * ceil(1000/4) * ceil(101/8) = 32 * 250 * 13 workgroups
    exceed the 16384 records of the workspace: WFE_E_ARG before any GPU call
    (the pointers are never dereferenced)."""
    from wfe_amd import _lib
    L = _lib.lib()
    out = _lib.wfe_diag()
    gr = _lib.wfe_grid(2000, 1000, 100, 3, 1000.0, 1000.0, 100.0)
    rc = L.wfe_diagnose_fields_noalloc(C.byref(gr), _lib.WFE_PREC_MIXED, 64, 64, 64, None, None,
                                       None, C.byref(out))
    assert rc == -1
    assert b"fixed workspace" in L.wfe_last_error()
