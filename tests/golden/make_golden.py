"""Writes the golden vectors of the CPU oracle (tests/golden/*.npz).

They are regression anchors: the oracle is pinned by the closed forms and
invariants of tests/test_oracle*.py, and these files record what it returns
on three small seeded cases, so that any later change to it shows up.
Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

from helpers import case, interiors, kitchen_sink, perturbed_state  # noqa: E402
from oracle_lib import OracleModel  # noqa: E402

STATE = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")
CASES = {
    "kitchen_sink_2steps": (lambda: kitchen_sink(nx=12, ny=10, nz=10, dx=3000, dy=3000,
                                                 mtn_xc=18000, mtn_yc=15000, bubble_xc=15000,
                                                 bubble_yc=12000, bubble_zc=1200), 2),
    "warm_bubble_3steps": (lambda: case("warm_bubble", nx=16, ny=4, nz=12, dx=500, dy=500,
                                        dz=500, bubble_xc=4000, bubble_yc=1000,
                                        bubble_zc=2000), 3),
    "schaer_2steps": (lambda: case("schaer", nx=24, ny=4, nz=12, mtn_xc=12000), 2),
}


def run(name):
    make, nsteps = CASES[name]
    c = make()
    st = perturbed_state(c, 17)
    om = OracleModel.from_case(c)
    for f, a in st.items():
        om.set(f, a)
    dt = c.scalar("dt")
    for s in range(nsteps):
        om.step(dt, s * dt)
    g = c.grid
    return interiors(g, {f: om.get(f) for f in STATE})


if __name__ == "__main__":
    for name in CASES:
        out = run(name)
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(name, {f: float(np.abs(a).max()) for f, a in out.items()})
