"""The reference's regression gates (tools/run_tests.py:72-101) on the gfx950
path, at full size, in the mixed-precision mode (the reference's gates are
for its FP32 build) -- plus the FP64 mode where the guarantee is exact.

Each case runs to the reference's t_end and evaluates the gate on the last
diagnostic, as run_tests.py does with the last stdout line.
"""
import math

import numpy as np
import pytest

from helpers import case

pytestmark = pytest.mark.gpu


def run(c, precision="mixed", nsteps=None, every=None):
    g = c.grid
    dt = c.scalar("dt")
    n = nsteps or int(math.ceil(c.scalar("t_end") / dt))
    gi = c.integrator(precision)
    for s in range(n):
        gi.step(dt, s * dt)
        if every and (s + 1) % every == 0:
            print(f"  step {s + 1}/{n} |w|max={gi.absmax('w'):.3f}", flush=True)
    gi.sync()
    return gi, n


def diag(c, gi):
    g = c.grid
    w = g.interior(gi.get("w"), g.nz + 1)
    th = g.interior(gi.get("thp"))
    pip = g.interior(gi.get("pip"))
    d = {"wmax": float(np.abs(w).max()), "thmin": float(min(0.0, th.min())),
         "thmax": float(max(0.0, th.max())), "pipmax": float(np.abs(pip).max())}
    if c.params.moisture:
        d["qrmax"] = float(g.interior(gi.get("qr")).max() * 1000)
        d["qcmax"] = float(g.interior(gi.get("qc")).max() * 1000)
    return d


def test_warm_bubble_gates_and_mirror_symmetry():
    c = case("warm_bubble")
    gi, n = run(c, every=200)
    assert n == 667
    d = diag(c, gi)
    print("warm_bubble", d, flush=True)
    assert 16.0 <= d["wmax"] <= 18.0
    assert 1.0 <= d["thmax"] <= 1.3
    g = c.grid
    u = g.interior(gi.get("u")).astype(np.float32)
    v = g.interior(gi.get("v")).astype(np.float32)
    sym = float(np.sqrt(((u - np.transpose(v, (0, 2, 1))) ** 2).mean()) /
                (np.sqrt((u ** 2).mean()) + 1e-12))
    print("symmetry", sym, flush=True)
    assert sym < 1e-4


def test_straka_gates():
    c = case("straka")
    gi, n = run(c, every=400)
    d = diag(c, gi)
    g = c.grid
    th = g.interior(gi.get("thp")).astype(np.float32)
    sfc = th[0, g.ny // 2, :]
    cold = np.where(sfc < -1.0)[0]
    front = (cold.max() + 0.5) * c.grid.dx - 25600.0 if len(cold) else 0.0
    print("straka", d, "front", front, flush=True)
    assert -9.6 <= d["thmin"] <= -8.5
    assert 13500 <= front <= 15500


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_schaer_rest_stays_at_rest(precision):
    c = case("schaer_rest")
    gi, n = run(c, precision)
    d = diag(c, gi)
    print("schaer_rest", precision, d, flush=True)
    assert d["wmax"] < 1e-6 and abs(d["thmax"]) < 1e-6


def test_schaer_mountain_wave():
    c = case("schaer")
    gi, n = run(c, every=2400)
    d = diag(c, gi)
    print("schaer", d, flush=True)
    assert 1.2 <= d["wmax"] <= 2.5


def test_bubble_outflow():
    c = case("bubble_outflow")
    gi, n = run(c, every=400)
    d = diag(c, gi)
    print("bubble_outflow", d, flush=True)
    assert d["thmax"] < 0.01 and d["wmax"] < 1.0
    assert d["pipmax"] < 1e-3


def test_wk82_supercell():
    c = case("wk82_supercell")
    gi, n = run(c, every=480)
    d = diag(c, gi)
    print("wk82", d, flush=True)
    assert 25.0 <= d["wmax"] <= 60.0
    assert d["qrmax"] > 0.2


def test_moist_blob_positivity():
    c = case("moist_blob")
    gi, n = run(c, every=1000)
    g = c.grid
    qv = g.interior(gi.get("qv")).astype(np.float32)
    print("moist_blob min qv", float(qv.min()), "max", float(qv.max()), flush=True)
    assert float(qv.min()) >= -1e-6
