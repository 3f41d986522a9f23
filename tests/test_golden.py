"""Committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle on three small seeded cases).

The CPU test holds the oracle to the vectors it wrote, so a later change to
the oracle cannot pass silently. The GPU tests hold the HIP path to the same
vectors without running the oracle at all, with the tolerances of the
full-step parity tests (relative L2 over the interior).
"""
from pathlib import Path

import numpy as np
import pytest

from helpers import rel_err

GOLDEN = Path(__file__).resolve().parent / "golden"

import sys  # noqa: E402
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402

STATE = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")
MIXED_TOL = {"u": 2e-5, "v": 2e-5, "w": 2e-5, "thp": 2e-5, "pip": 1e-4, "qv": 1e-4,
             "qc": 1e-4, "qr": 1e-4}


def _golden(name):
    with np.load(GOLDEN / f"{name}.npz") as z:
        return {f: z[f] for f in z.files}


def _check(name, got, tol):
    ref = _golden(name)
    assert set(ref) == set(STATE)
    errs = {f: rel_err(got[f], ref[f]) for f in STATE}
    print(name, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    bad = {f: e for f, e in errs.items() if not e <= (tol[f] if isinstance(tol, dict) else tol)}
    assert not bad, f"{name}: {bad}"


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_golden(name):
    _check(name, make_golden.run(name), 1e-12)


def _gpu_run(name, precision):
    from wfe_amd import Integrator
    from helpers import interiors, perturbed_state

    make, nsteps = make_golden.CASES[name]
    c = make()
    gi = Integrator.from_case(c, precision)
    for f, a in perturbed_state(c, 17).items():
        gi.set(f, a)
    dt = c.scalar("dt")
    for s in range(nsteps):
        gi.step(dt, s * dt)
    return interiors(c.grid, {f: gi.get(f) for f in STATE})


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_gpu_fp64_matches_golden(name):
    _check(name, _gpu_run(name, "fp64"), 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_gpu_mixed_matches_golden(name):
    _check(name, _gpu_run(name, "mixed"), MIXED_TOL)
