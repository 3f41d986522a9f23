"""The standalone driver wfe_amd/lib/wfe, the counterpart of the reference's
src/main.cpp: .ini in, the reference's meta.json and <var>_%06d.bin out,
exit codes 0 / 1 (setup) / 2 (NaN) / 3 (|w| guard) (main.cpp:353-472,
writer.cpp:27-107)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import rel_err

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "wfe_amd" / "lib" / "wfe"
SUPERCELL = ROOT / "cases" / "wk82_supercell.ini"

# A small moist supercell: 24 x 20 x 18 at 2 km, 10 steps of 6 s.
SMALL = dict(nx=24, ny=20, nz=18, dx=2000, dy=2000, dz=600, dt=6, t_end=60,
             bubble_xc=24000, bubble_yc=20000, bubble_xr=8000, bubble_yr=8000)


def _run(args, **kw):
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=300, **kw)


def test_driver_rejects_a_missing_config():
    r = _run([ROOT / "cases" / "no_such_case.ini"])
    assert r.returncode == 1
    assert "case setup failed" in r.stderr


def test_driver_rejects_a_bad_grid():
    r = _run([ROOT / "cases" / "warm_bubble.ini", "nx=2"])
    assert r.returncode == 1
    assert "grid too small" in r.stderr


@pytest.mark.gpu
def test_driver_runs_a_case_and_matches_the_library(tmp_path, monkeypatch):
    """The driver steps a moist case, prints the adim / bitti lines, writes
    meta.json and every field at each output step, and its last output equals
    the same steps taken through the library directly."""
    from wfe_amd import Case, Integrator

    monkeypatch.setenv("WFE_ACOUSTIC_KERNEL", "fused-64x16")
    over = {**SMALL, "out_dir": tmp_path, "out_interval": 30, "diag_interval": 30,
            "precision": "fp64"}
    r = _run([SUPERCELL, *[f"{k}={v}" for k, v in over.items()]], cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "adim     10" in r.stdout and "bitti: 10 adim" in r.stdout

    meta = json.loads((tmp_path / "meta.json").read_text())
    nx, ny, nz = SMALL["nx"], SMALL["ny"], SMALL["nz"]
    assert (meta["nx"], meta["ny"], meta["nz"]) == (nx, ny, nz)
    assert meta["vars"]["w"] == nz + 1 and meta["vars"]["qr"] == nz and meta["vars"]["rain"] == 1
    for step in (0, 5, 10):
        for v, nl in meta["vars"].items():
            if v == "rain" and step == 0:
                continue
            f = tmp_path / f"{v}_{step:06d}.bin"
            assert f.stat().st_size == 4 * nx * ny * nl, f.name
    assert (tmp_path / "zc.bin").stat().st_size == 4 * nx * ny * nz

    c = Case(SUPERCELL, **SMALL)
    gi = Integrator.from_case(c, "fp64")
    for s in range(10):
        gi.step(SMALL["dt"], s * SMALL["dt"])
    g = c.grid
    for v in ("u", "v", "w", "thp", "pip", "qv", "qc", "qr"):
        nl = nz + 1 if v == "w" else nz
        drv = np.fromfile(tmp_path / f"{v}_000010.bin", dtype=np.float32).reshape(nl, ny, nx)
        lib = g.interior(gi.get(v), nl).astype(np.float32)
        assert rel_err(drv, lib) <= 1e-6, v
    rain = np.fromfile(tmp_path / "rain_000010.bin", dtype=np.float32).reshape(ny, nx)
    r2 = gi.get("rain").reshape(g.NY, g.NX)[g.ng:g.ng + ny, g.ng:g.ng + nx].astype(np.float32)
    assert rel_err(rain, r2) <= 1e-6


@pytest.mark.gpu
def test_driver_w_guard_exits_3_with_emergency_output(tmp_path):
    """|w|max above w_abort at a 10-step check stops the run with exit code 3
    after writing the current state (main.cpp:429-437)."""
    over = {**SMALL, "out_dir": tmp_path, "out_interval": 60, "diag_interval": 60,
            "w_abort": 1e-6}
    r = _run([SUPERCELL, *[f"{k}={v}" for k, v in over.items()]], cwd=tmp_path)
    assert r.returncode == 3, (r.stdout[-1000:], r.stderr[-1000:])
    assert "exceeded the limit" in r.stderr
    assert (tmp_path / "w_000010.bin").stat().st_size == 4 * 24 * 20 * 19


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_integrator_init_path_matches_the_case_path(precision):
    """wfe::Integrator::init (grid, base-state and metric views, DynParams;
    the reference's init, integrator.hpp:18-42) with the initial state
    uploaded through state().<field>.upload steps bit-identically to the
    context the case builder creates in one call (terrain, open boundaries,
    moisture, PD, Kessler)."""
    exe = ROOT / "wfe_amd" / "lib" / "wfe_integrator_check"
    args = [SUPERCELL, "steps=3", f"precision={precision}", "nx=24", "ny=20", "nz=18",
            "dx=2000", "dy=2000", "dz=600", "dt=6", "terrain=gauss", "mtn_h0=600",
            "mtn_a=8000", "mtn_xc=24000", "mtn_yc=20000", "pd_moist=on", "coriolis_f=1e-4",
            "bubble_xc=24000", "bubble_yc=20000", "bubble_xr=8000", "bubble_yr=8000"]
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "identical" in r.stdout


def test_integrator_check_rejects_file_mode(tmp_path):
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", nx=12, ny=10, nz=8, dx=12000.0, top=5000.0, hours=0)
    exe = ROOT / "wfe_amd" / "lib" / "wfe_integrator_check"
    r = subprocess.run([str(exe), str(ini)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "file-mode" in r.stderr


@pytest.mark.gpu
def test_driver_profile_table_and_launch_sync(tmp_path):
    """WFE_PROF=1 prints the reference's section table from per-kernel times
    (integrator.cpp:113-122); WFE_SYNC=1 waits after every launch (step
    graphs off) and gives the same diagnostics as the default run."""
    import os
    over = {**SMALL, "out_dir": tmp_path, "out_interval": 60, "diag_interval": 30}
    args = [SUPERCELL, *[f"{k}={v}" for k, v in over.items()]]
    base = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=300, cwd=tmp_path)
    prof = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=300, cwd=tmp_path, env=dict(os.environ, WFE_PROF="1"))
    sync = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=300, cwd=tmp_path, env=dict(os.environ, WFE_SYNC="1"))
    print(prof.stdout[-1500:], flush=True)
    for r in (base, prof, sync):
        assert r.returncode == 0, r.stderr
    assert "profil (WFE_PROF=1" in prof.stdout
    for section in ("sinir kosullari", "kutle akisi+div", "yavas tendency", "akustik dongu",
                    "kessler"):
        assert section in prof.stdout, section
    lines = lambda r: [l for l in r.stdout.splitlines() if l.startswith("adim")]  # noqa: E731
    assert lines(base) == lines(prof) == lines(sync)
