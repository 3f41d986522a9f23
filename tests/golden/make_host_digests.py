"""Writes tests/golden/host_builder_digests.json: SHA-256 digests of every
FP64 array the native host case builder produces (Metric, BaseState, the
initial state and the Davies boundary weights and targets) for the
idealized case files, the test grid that switches everything on, an
ensemble-perturbed regional case and the full 330 x 220 x 45 regional case.

The digests pin the host builders bit for bit, so a restructuring of them
that changes any value by one ulp shows up.
Run from the repository root: python tests/golden/make_host_digests.py
"""
from __future__ import annotations

import hashlib
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

from helpers import case, kitchen_sink  # noqa: E402
from wfe_amd.lam_case import write_lam_case  # noqa: E402

OUT = HERE / "host_builder_digests.json"
ARRAYS = ("thb", "pib", "rhob", "dthbdz", "thbw", "rhobw", "thvb", "thvbw", "qvb", "ub", "vb",
          "zeta_w", "zeta_c", "dzeta_c", "dzeta_w", "h", "hx", "hy", "hx_u", "hy_v", "jac",
          "fcor", "mapf", "u", "v", "w", "thp", "pip", "qv", "qc", "qr", "bdy_wgt") + tuple(
    f"bdy_{s}_{v}" for s in ("lo", "hi") for v in ("u", "v", "thp", "pip", "qv"))
IDEALIZED = ("cfg1_dry_bubble", "cfg2_supercell", "cfg5_adjoint_dry", "warm_bubble", "schaer",
             "schaer_rest", "straka", "moist_blob", "bubble_outflow", "wk82_supercell")


def digests(c) -> dict:
    out = {}
    for name in ARRAYS:
        try:
            a = c.array(name)
        except KeyError:
            continue
        out[name] = hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    out["zt"] = repr(c.scalar("zt"))
    return out


def regional(tmp: Path, tag: str, **kw):
    ini = write_lam_case(tmp / tag, **kw)
    from wfe_amd import Case
    return Case(ini)


def cases(tmp: Path):
    for name in IDEALIZED:
        yield name, lambda n=name: case(n)
    yield "kitchen_sink", lambda: kitchen_sink()
    yield "regional_small", lambda: regional(tmp, "small", nx=24, ny=18, nz=16, dx=12000.0,
                                             top=12000.0, seed=7, hours=2, bdy_interval_h=1)
    yield "regional_member_perturbed", lambda: regional(
        tmp, "member", nx=30, ny=20, nz=12, dx=12000.0, top=10000.0, seed=11, low_noise_K=0.5,
        extra={"ic_perturb": 0.3, "ens_seed": 5})
    yield "regional_330x220x45", lambda: regional(tmp, "full", nx=330, ny=220, nz=45)


def compute() -> dict:
    with tempfile.TemporaryDirectory() as d:
        return {name: digests(make()) for name, make in cases(Path(d))}


if __name__ == "__main__":
    OUT.write_text(json.dumps(compute(), indent=1, sort_keys=True) + "\n")
    print("wrote", OUT)
