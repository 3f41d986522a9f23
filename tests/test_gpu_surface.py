"""Surface/PBL/soil/radiation column physics (SfcPBL, src/physics/surface.cu)
on the GPU against the oracle, both built from the same regional case file
by their own case assembly, with physics = simple (src/main.cpp:364-377)."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import pin_shapes, rel_err
from oracle_lib import OracleCase

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
STATE = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")
SURFACE = ("tsk", "pblh", "t2m", "u10")


def _case(tmp_path, pbl, **kw):
    from wfe_amd.lam_case import write_lam_case
    args = dict(nx=40, ny=30, nz=20, dx=12000.0, top=15000.0, seed=9, hours=3, bdy_interval_h=3)
    args.update(kw)
    return write_lam_case(tmp_path / f"lam_{pbl}", extra={"physics": "simple", "pbl": pbl},
                          **args)


def _compare(g, gi, om, tol_state, tol_sfc, label):
    errs = {}
    for f in STATE:
        nzl = g.nz + 1 if f == "w" else g.nz
        errs[f] = rel_err(g.interior(gi.get(f), nzl), g.interior(om.get(f), nzl))
    for f in SURFACE:
        a = gi.get(f).reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
        b = om.get(f).reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
        errs[f] = rel_err(a, b)
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items()
           if not v <= (tol_sfc if k in SURFACE else tol_state)}
    assert not bad, (label, bad)
    return errs


@pytest.mark.parametrize("pbl", ["local", "nonlocal"])
def test_surface_physics_fp64_matches_the_oracle(tmp_path, pbl):
    from wfe_amd import Case, Integrator
    ini = _case(tmp_path, pbl)
    c = Case(ini)
    assert c.scalar("sfc_on") == 1 and c.scalar("sfc_nonlocal") == (pbl == "nonlocal")
    om = OracleCase(ini).model()
    gi = Integrator.from_case(c, "fp64")
    g = c.grid
    dt = c.scalar("dt")
    # 6 steps starting at 00 UTC: night; then jump the clock to noon so the
    # short-wave branch and a convective boundary layer are exercised
    for s in range(6):
        gi.step(dt, s * dt)
        om.step(dt, s * dt)
    _compare(g, gi, om, 1e-10, 1e-10, f"{pbl} night")
    t0 = 12 * 3600.0
    for s in range(6):
        gi.step(dt, t0 + s * dt)
        om.step(dt, t0 + s * dt)
    e = _compare(g, gi, om, 1e-10, 1e-10, f"{pbl} noon")
    pblh = gi.get("pblh").reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
    if pbl == "nonlocal":
        assert pblh.max() > 100.0  # a daytime mixed layer somewhere
    else:
        assert np.all(pblh == 0)
    assert e["tsk"] < 1e-10


def test_surface_physics_mixed_precision_tracks_the_oracle(tmp_path):
    from wfe_amd import Case, Integrator
    ini = _case(tmp_path, "nonlocal")
    c = Case(ini)
    om = OracleCase(ini).model()
    gi = Integrator.from_case(c, "mixed")
    dt = c.scalar("dt")
    for s in range(10):
        gi.step(dt, 12 * 3600.0 + s * dt)
        om.step(dt, 12 * 3600.0 + s * dt)
    _compare(c.grid, gi, om, 2e-4, 1e-4, "mixed noon")


def test_surface_physics_under_step_graphs_matches_eager_steps(tmp_path, monkeypatch):
    """The hour of day is a device scalar written before each step, so the
    replayed step graph follows the clock."""
    from wfe_amd import Case, Integrator
    pin_shapes(monkeypatch)
    ini = _case(tmp_path, "nonlocal")
    c = Case(ini)
    dt = c.scalar("dt")
    outs = []
    for graph in ("1", "0"):
        monkeypatch.setenv("WFE_GRAPH", graph)
        gi = Integrator.from_case(c, "fp64")
        for s in range(8):
            gi.step(dt, 9 * 3600.0 + s * 1800.0)
        outs.append({f: gi.get(f) for f in ("thp", "u", "tsk", "t2m")})
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f


def test_driver_writes_the_surface_fields(tmp_path):
    ini = _case(tmp_path, "local")
    out = tmp_path / "run"
    r = subprocess.run([str(ROOT / "wfe_amd" / "lib" / "wfe"), str(ini), "t_end=600",
                        "out_interval=300", "diag_interval=300", f"out_dir={out}"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-800:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "fizik: yuzey + yerel-K PBL" in r.stdout
    meta = json.loads((out / "meta.json").read_text())
    for f in SURFACE:
        assert meta["vars"][f] == 1
        assert (out / f"{f}_000010.bin").stat().st_size == 4 * 40 * 30
    t2 = np.fromfile(out / "t2m_000010.bin", dtype=np.float32)
    assert np.isfinite(t2).all() and 200 < t2.min() and t2.max() < 330


def test_ensembles_with_surface_physics_follow_the_clock(tmp_path, monkeypatch):
    """Members with surface physics in an ensemble, stepped with their times
    (one member on another clock), equal the same members stepped alone, bit
    for bit; without times the ensemble step is refused rather than run
    with a stale hour of day."""
    pin_shapes(monkeypatch)
    from wfe_amd import Case, Ensemble, Integrator
    from wfe_amd._lib import WfeError
    cases = [Case(_case(tmp_path, "local", hours=0)),
             Case(_case(tmp_path, "nonlocal", nx=32, ny=24, nz=20, seed=5, hours=0))]
    dts = [c.scalar("dt") for c in cases]
    t0 = [9 * 3600.0, 15 * 3600.0]
    alone = [Integrator.from_case(c, "mixed") for c in cases]
    for s in range(6):
        for gi, dt, t in zip(alone, dts, t0):
            gi.step(dt, t + s * 1800.0)
    members = [Integrator.from_case(c, "mixed") for c in cases]
    ens = Ensemble(members)
    with pytest.raises(WfeError, match="surface physics"):
        ens.step(dts)
    for s in range(6):
        ens.step(dts, [t + s * 1800.0 for t in t0])
    ens.sync()
    for q, (a, b) in enumerate(zip(alone, members)):
        for f in STATE + ("tsk", "t2m", "pblh"):
            np.testing.assert_array_equal(b.get(f), a.get(f), err_msg=f"member {q} {f}")
    del ens
