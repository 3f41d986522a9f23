"""Output path on the device: the driver's interior diagnostics as a device
reduction (diagnose, src/main.cpp:202-240), interior packing of fields in
the reference's output order with asynchronous copies (Writer::write_field,
src/io/writer.cpp:89-107), and the driver's asynchronous writer against its
synchronous (reference) path."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import kitchen_sink, perturbed_state

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "wfe_amd" / "lib" / "wfe"
SUPERCELL = ROOT / "cases" / "wk82_supercell.ini"


def host_diagnose(g, fields, moist):
    """The host loops of src/main.cpp:202-240 over the interior, written with
    the same comparisons (NaN never enters an extreme; it clears finite)."""
    d = dict(wmax=0.0, thmin=0.0, thmax=0.0, pipmax=0.0, qcmax=0.0, qrmax=0.0, finite=1)
    w = g.interior(fields["w"], g.nz + 1).ravel()
    if not np.isfinite(w).all():
        d["finite"] = 0
    aw = np.abs(w[~np.isnan(w)])
    if aw.size:
        d["wmax"] = max(0.0, float(aw.max()))
    th = g.interior(fields["thp"], g.nz).ravel()
    if not np.isfinite(th).all():
        d["finite"] = 0
    th = th[~np.isnan(th)]
    if th.size:
        d["thmin"] = min(0.0, float(th.min()))
        d["thmax"] = max(0.0, float(th.max()))
    p = np.abs(g.interior(fields["pip"], g.nz).ravel())
    if not np.isfinite(p).all():
        d["finite"] = 0
    p = p[~np.isnan(p)]
    if p.size:
        d["pipmax"] = max(0.0, float(p.max()))
    if moist:
        qc = g.interior(fields["qc"], g.nz).ravel()
        qc = qc[~np.isnan(qc)]
        if qc.size:
            d["qcmax"] = max(0.0, float(qc.max()))
        qr = g.interior(fields["qr"], g.nz).ravel()
        if not np.isfinite(qr).all():
            d["finite"] = 0
        qr = qr[~np.isnan(qr)]
        if qr.size:
            d["qrmax"] = max(0.0, float(qr.max()))
    return d


def _stepped(precision, nsteps=3):
    from wfe_amd import Integrator
    c = kitchen_sink(nx=40, ny=30, nz=22)
    gi = Integrator.from_case(c, precision)
    st = perturbed_state(c, 7)
    for f, a in st.items():
        gi.set(f, a)
    dt = c.scalar("dt")
    for s in range(nsteps):
        gi.step(dt, s * dt)
    return c, gi


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_device_diagnose_equals_the_host_loops(precision):
    c, gi = _stepped(precision)
    g = c.grid
    fields = {f: gi.get(f) for f in ("w", "thp", "pip", "qc", "qr")}
    dev = gi.diagnose()
    host = host_diagnose(g, fields, True)
    print(precision, dev, flush=True)
    assert dev == host
    assert dev["wmax"] > 0 and dev["qrmax"] >= 0 and dev["finite"] == 1


@pytest.mark.gpu
def test_device_diagnose_nan_and_inf_rules():
    """NaN never enters an extreme but clears finite; Inf enters; qc is not
    part of the finite test (main.cpp:226-239)."""
    c, gi = _stepped("mixed", 1)
    g = c.grid
    f3 = {f: gi.get(f) for f in ("w", "thp", "pip", "qc", "qr")}
    ng = g.ng
    idx = lambda i, j, k: ((k + ng) * g.NY + (j + ng)) * g.NX + (i + ng)  # noqa: E731
    cases = [
        ("w", idx(3, 4, 5), np.nan),
        ("thp", idx(7, 2, 1), -np.inf),
        ("qr", idx(1, 1, 1), np.inf),
        ("qc", idx(2, 2, 2), np.nan),
    ]
    for name, cell, val in cases:
        a = f3[name].copy()
        a[cell] = val
        gi.set(name, a)
        fields = dict(f3)
        fields[name] = a
        dev = gi.diagnose()
        host = host_diagnose(g, fields, True)
        print(name, val, dev, flush=True)
        assert dev == host, (name, dev, host)
        assert dev["finite"] == (0 if name != "qc" else 1)
        gi.set(name, f3[name])
    assert gi.diagnose()["finite"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_packed_fields_equal_the_downloaded_interior(precision):
    c, gi = _stepped(precision)
    g = c.grid
    for f in ("u", "v", "w", "thp", "pip", "qv", "qc", "qr"):
        nl = g.nz + 1 if f == "w" else g.nz
        want = g.interior(gi.get(f), nl).astype(np.float32).ravel()
        got = gi.packed(f)
        assert got.tobytes() == want.tobytes(), f
    rain = gi.get("rain").reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
    assert gi.packed("rain").tobytes() == rain.astype(np.float32).ravel().tobytes()


@pytest.mark.gpu
def test_async_copies_hold_the_state_of_the_pack_while_steps_go_on():
    """A pack enqueued after step n and waited for after two more steps holds
    step n's state, in pinned and in ordinary host memory."""
    from wfe_amd import _lib
    L = _lib.lib()
    c, gi = _stepped("mixed", 1)
    g = c.grid
    dt = c.scalar("dt")
    n = g.nx * g.ny * g.nz
    want = {f: g.interior(gi.get(f), g.nz).astype(np.float32).ravel() for f in ("u", "thp")}
    p = C.c_void_p()
    _lib.check(L.wfe_host_alloc(n * 4, C.byref(p)), "host alloc")
    try:
        pinned = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n,))
        plain = np.empty(n, dtype=np.float32)
        _lib.check(L.wfe_pack_field_async(gi._h, 0, C.cast(p, C.POINTER(C.c_float)), n), "pack")
        _lib.check(L.wfe_pack_field_async(gi._h, 3, plain.ctypes.data_as(C.POINTER(C.c_float)),
                                          n), "pack")
        t = C.c_longlong()
        _lib.check(L.wfe_output_mark(gi._h, C.byref(t)), "mark")
        gi.step(dt, dt)
        gi.step(dt, 2 * dt)
        _lib.check(L.wfe_output_wait(gi._h, t.value), "wait")
        assert pinned.tobytes() == want["u"].tobytes()
        assert plain.tobytes() == want["thp"].tobytes()
        after = g.interior(gi.get("u"), g.nz).astype(np.float32).ravel()
        assert after.tobytes() != want["u"].tobytes()
        # a bad count and an unknown ticket are errors, not faults
        assert L.wfe_pack_field_async(gi._h, 0, C.cast(p, C.POINTER(C.c_float)), n - 1) == -1
        assert L.wfe_output_wait(gi._h, t.value + 1000) == -1
    finally:
        _lib.check(L.wfe_host_free(p), "host free")


@pytest.mark.gpu
@pytest.mark.parametrize("precision,moist", [("mixed", True), ("fp64", True), ("mixed", False)])
def test_stateless_diagnose_over_caller_device_fields(precision, moist):
    """wfe_diagnose_fields over caller-owned device arrays in the reference
    layout against the host loops; ragged sizes."""
    from wfe_amd import Grid, _lib
    L = _lib.lib()
    rng = np.random.default_rng(11)
    for nx, ny, nz in ((37, 13, 9), (70, 5, 17)):
        g = Grid(nx, ny, nz, 3, 1000.0, 1000.0, 100.0)
        ft = np.float64 if precision == "fp64" else np.float32
        host = {f: (rng.standard_normal(g.npts) * s).astype(np.float64 if f == "pip" else ft)
                for f, s in (("w", 3.0), ("thp", 2.0), ("pip", 1e-3), ("qc", 1e-3),
                             ("qr", 1e-3))}
        dev = {f: _lib.DeviceArray(a) for f, a in host.items()}
        want = host_diagnose(g, {f: a.astype(np.float64) for f, a in host.items()}, moist)
        out = _lib.wfe_diag()
        gr = _lib.wfe_grid(nx, ny, nz, 3, 1000.0, 1000.0, 100.0)
        prec = _lib.WFE_PREC_FP64 if precision == "fp64" else _lib.WFE_PREC_MIXED
        rc = L.wfe_diagnose_fields(C.byref(gr), prec, dev["w"].ptr, dev["thp"].ptr,
                                   dev["pip"].ptr, dev["qc"].ptr if moist else None,
                                   dev["qr"].ptr if moist else None, None, C.byref(out))
        _lib.check(rc, "wfe_diagnose_fields")
        got = {f: getattr(out, f) for f, _ in _lib.wfe_diag._fields_}
        assert got == want, (nx, ny, nz, got, want)


def test_stateless_diagnose_rejects_bad_arguments():
    from wfe_amd import _lib
    L = _lib.lib()
    out = _lib.wfe_diag()
    gr = _lib.wfe_grid(8, 8, 8, 3, 1.0, 1.0, 1.0)
    assert L.wfe_diagnose_fields(None, 1, None, None, None, None, None, None, C.byref(out)) == -1
    assert L.wfe_diagnose_fields(C.byref(gr), 1, None, None, None, None, None, None,
                                 C.byref(out)) == -1
    assert b"null" in L.wfe_last_error()
    gr.ng = 2
    assert L.wfe_diagnose_fields(C.byref(gr), 1, 8, 8, 8, None, None, None, C.byref(out)) == -1


def _driver(args, env_extra, cwd):
    import os
    env = dict(os.environ, **env_extra)
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=300, cwd=cwd, env=env)


@pytest.mark.gpu
def test_driver_async_output_matches_the_synchronous_path(tmp_path):
    """The driver with the device diagnostics and the asynchronous writer
    prints the same diagnostic lines and writes the same bytes as with
    whole-field downloads and host loops (WFE_SYNC_OUTPUT=1)."""
    small = dict(nx=30, ny=22, nz=18, dx=2000, dy=2000, dz=600, dt=6, t_end=72,
                 bubble_xc=30000, bubble_yc=22000, bubble_xr=8000, bubble_yr=8000,
                 out_interval=18, diag_interval=12, precision="mixed")
    outs = {}
    for mode, env in (("async", {}), ("sync", {"WFE_SYNC_OUTPUT": "1"})):
        d = tmp_path / mode
        d.mkdir()
        env["WFE_ACOUSTIC_KERNEL"] = "fused-64x16"
        r = _driver([SUPERCELL, *[f"{k}={v}" for k, v in {**small, "out_dir": d}.items()]],
                    env, d)
        print(mode, r.stdout[-1500:], r.stderr[-800:], flush=True)
        assert r.returncode == 0, r.stderr
        outs[mode] = (d, [l for l in r.stdout.splitlines() if l.startswith("adim")])
    (da, la), (ds, ls) = outs["async"], outs["sync"]
    assert la == ls and len(la) == 6
    files = sorted(p.name for p in ds.iterdir() if p.suffix == ".bin")
    assert files == sorted(p.name for p in da.iterdir() if p.suffix == ".bin")
    assert "rain_000012.bin" in files and "w_000012.bin" in files
    for name in files:
        assert (da / name).read_bytes() == (ds / name).read_bytes(), name
