"""Shared test helpers: case construction, seeded perturbations, error norms."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from wfe_amd import Case, Integrator
from oracle_lib import OracleModel

ROOT = Path(__file__).resolve().parent.parent
CASES = ROOT / "cases"

# A small grid exercising every switch of the hot path at once: terrain,
# open x and y, WK82 moisture with PD transport and Kessler, Coriolis,
# diffusion, Rayleigh, w damping and the quarter-circle (u, v) base wind.
KITCHEN_SINK = dict(
    nx=24, ny=20, nz=18, dx=2000, dy=2000, dz=600, dt=6, acoustic_ns=6,
    profile="wk82", rh_max=0.95, wind_profile="quarter_circle", wind_us=20, wind_zs=5000,
    wind_offset=-8, wind_offset_v=-6, terrain="gauss", mtn_h0=600, mtn_a=8000,
    mtn_xc=24000, mtn_yc=20000, bc_x="open", bc_y="open", pd_moist="on",
    coriolis_f=1e-4, diff_K=15, w_damping="on", rayleigh_zd=7000, rayleigh_alpha=0.02,
    bubble_dtheta=3, bubble_xc=20000, bubble_yc=18000, bubble_zc=1400, bubble_xr=8000,
    bubble_yr=8000, bubble_zr=1400,
)


def case(name: str, **over) -> Case:
    return Case(CASES / f"{name}.ini", **over)


def kitchen_sink(**over) -> Case:
    kw = dict(KITCHEN_SINK)
    kw.update(over)
    return case("wk82_supercell", **kw)


def perturbed_state(c: Case, seed: int, amp: dict | None = None) -> dict:
    """Initial state of the case plus seeded smooth-ish noise on every
    prognostic (interior and faces), so all code paths see non-trivial data."""
    rng = np.random.default_rng(seed)
    g = c.grid
    amp = amp or {"u": 0.5, "v": 0.5, "w": 0.2, "thp": 0.3, "pip": 2e-5,
                  "qv": 5e-4, "qc": 3e-4, "qr": 3e-4}
    st = {}
    moist = bool(c.params.moisture)
    for f in ("u", "v", "w", "thp", "pip", "qv", "qc", "qr"):
        a = c.array(f).copy()
        a3 = g.as3d(a)
        ng = g.ng
        if f in ("qv", "qc", "qr") and not moist:
            st[f] = a
            continue
        nzl = g.nz + (1 if f == "w" else 0)
        nxl = g.nx + (1 if f == "u" else 0)
        nyl = g.ny + (1 if f == "v" else 0)
        noise = rng.standard_normal((nzl, nyl, nxl)) * amp[f]
        if f == "w":
            noise[0] = 0.0       # surface w is diagnostic
            noise[-1] = 0.0      # rigid lid w(nz) = 0
        sl = a3[ng:ng + nzl, ng:ng + nyl, ng:ng + nxl]
        sl += noise
        if f in ("qv", "qc", "qr"):
            np.maximum(sl, 0.0, out=sl)
        st[f] = a
    return st


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b||_2 / ||b||_2 (absolute when b is zero)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def interiors(g, st: dict) -> dict:
    out = {}
    for f, a in st.items():
        nzl = g.nz + 1 if f == "w" else g.nz
        out[f] = g.interior(a, nzl)
    return out


def oracle_and_gpu(c: Case, precision: str, state: dict | None = None):
    """Oracle and GPU integrator started from the same state."""
    om = OracleModel.from_case(c)
    gi = Integrator.from_case(c, precision)
    if state:
        for f, a in state.items():
            om.set(f, a)
            gi.set(f, a)
    return om, gi


def pin_shapes(monkeypatch) -> None:
    """Contexts compared bit for bit run the same kernel shapes: each context
    otherwise tunes its own on its first step by timing, and the substep
    variants differ in rounding."""
    monkeypatch.setenv("WFE_ACOUSTIC_KERNEL", "split-64x16")
    monkeypatch.setenv("WFE_TENDENCY_LEVELS", "2")
    monkeypatch.setenv("WFE_TENDENCY_TILE", "64")
    monkeypatch.setenv("WFE_KESSLER_TILE", "32x16")
