"""Pins of the CPU oracle by independent means (no GPU).

Each test checks one part of the oracle against something that does not
share its code: a direct numpy restatement written from the reference's
kernel semantics (ghost fill, Davies relaxation), or what the mathematics
fixes (5th-order convergence of the flux-form advection against the exact
derivative). Each is chosen so that a plausible slip -- a swapped fill
order, an off-by-one ghost, a wrong sign or a dropped factor -- fails it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import case, kitchen_sink, perturbed_state
from oracle_lib import OracleModel


# --------------------------------------------------------------- ghost fill
def _fill_reference_semantics(a3, g, open_x, open_y, kind):
    """Ghost fill restated from src/dynamics/kernels.cu:692-835: vertical
    rule over every (j, i) column, then x over every (k, j) row, then y over
    every (k, i); `kind` is "c" (cell), "u", "v" or "w"."""
    ng, nx, ny, nz = g.ng, g.nx, g.ny, g.nz
    f = a3.copy()
    K = lambda k: k + ng  # noqa: E731
    if kind == "w":  # kernels.cu:709-723: lid 0, odd reflection about w(0), w(nz)
        f[K(nz)] = 0.0
        w0 = f[K(0)].copy()
        for m in range(1, ng + 1):
            f[K(-m)] = 2 * w0 - f[K(m)]
            f[K(nz + m)] = -f[K(nz - m)]
    else:  # kernels.cu:695-707: zero gradient, ng below and ng + 1 above
        bot, top = f[K(0)].copy(), f[K(nz - 1)].copy()
        for m in range(1, ng + 1):
            f[K(-m)] = bot
        for m in range(1, ng + 2):
            f[K(nz - 1 + m)] = top
    ilast = nx if kind == "u" else nx - 1
    jlast = ny if kind == "v" else ny - 1
    NXr, NYr = f.shape[2], f.shape[1]
    for m in range(ng):
        if open_x:  # kernels.cu:749-759; the write past the row end is dropped
            f[:, :, ng - 1 - m] = f[:, :, ng]
            if ng + ilast + 1 + m < NXr:
                f[:, :, ng + ilast + 1 + m] = f[:, :, ng + ilast]
        else:  # kernels.cu:725-733
            f[:, :, ng + nx + m] = f[:, :, ng + m]
            f[:, :, ng - 1 - m] = f[:, :, ng + nx - 1 - m]
    for m in range(ng):
        if open_y:  # kernels.cu:761-771
            f[:, ng - 1 - m, :] = f[:, ng, :]
            if ng + jlast + 1 + m < NYr:
                f[:, ng + jlast + 1 + m, :] = f[:, ng + jlast, :]
        else:  # kernels.cu:735-745
            f[:, ng + ny + m, :] = f[:, ng + m, :]
            f[:, ng - 1 - m, :] = f[:, ng + ny - 1 - m, :]
    return f


@pytest.mark.parametrize("bc", ["periodic", "open", "mixed"])
def test_ghost_fill_matches_direct_restatement(bc):
    bx = "open" if bc in ("open", "mixed") else "periodic"
    by = "open" if bc == "open" else "periodic"
    c = kitchen_sink(nx=9, ny=7, nz=6, bc_x=bx, bc_y=by)
    g = c.grid
    om = OracleModel.from_case(c)
    rng = np.random.default_rng(11)
    st = {}
    for f in ("u", "v", "w", "thp", "pip", "qv", "qc", "qr"):
        a = rng.standard_normal(g.npts)  # ghosts start as garbage too
        st[f] = a
        om.set(f, a)
    om.apply_bcs()
    kinds = {"u": "u", "v": "v", "w": "w"}
    for f, a in st.items():
        want = _fill_reference_semantics(g.as3d(a), g, bx == "open", by == "open",
                                         kinds.get(f, "c"))
        got = g.as3d(om.get(f))
        np.testing.assert_array_equal(got, want, err_msg=f)


# ------------------------------------------------------ Davies relaxation
def test_davies_relaxation_matches_closed_form():
    """T += w (lo (1 - tf) + hi tf - f) on the reference's index ranges
    (kernels.cu:292-304, 981-995) with w = cos^2(pi/2 d/width)/tau on the
    rim and 1/nudge_tau inside (boundary.cpp:24-40)."""
    width, tau, ntau, tf = 3, 500.0, 4000.0, 0.3
    c = kitchen_sink(nx=10, ny=8, nz=5)
    g = c.grid
    st = perturbed_state(c, 5)
    rng = np.random.default_rng(6)
    lo = [rng.standard_normal(g.npts) for _ in range(5)]
    hi = [rng.standard_normal(g.npts) for _ in range(5)]
    wgt = np.zeros(g.NX * g.NY)
    for j in range(g.ny):
        for i in range(g.nx):
            d = min(i, g.nx - 1 - i, j, g.ny - 1 - j)
            w = 1.0 / ntau
            if d < width:
                w = max(w, math.cos(0.5 * math.pi * d / width) ** 2 / tau)
            wgt[(j + g.ng) * g.NX + i + g.ng] = w

    def tend(active):
        om = OracleModel.from_case(c)
        for f, a in st.items():
            om.set(f, a)
        om.set("bdy_wgt", wgt)
        for v in range(5):
            om.set(f"bdy_lo_{v}", lo[v])
            om.set(f"bdy_hi_{v}", hi[v])
        om.set_bdy(active, tf)
        om.apply_bcs()
        om.tendencies(6.0)
        return {f: om.get(f) for f in ("tu", "tv", "tthp", "tpip", "tqv")}

    t0, t1 = tend(0), tend(1)
    # the state the relaxation sees is the ghost-filled one
    om = OracleModel.from_case(c)
    for f, a in st.items():
        om.set(f, a)
    om.apply_bcs()
    w2 = wgt.reshape(g.NY, g.NX)
    for v, (tn, fn, iext, jext) in enumerate([("tu", "u", 1, 0), ("tv", "v", 0, 1),
                                              ("tthp", "thp", 0, 0), ("tpip", "pip", 0, 0),
                                              ("tqv", "qv", 0, 0)]):
        f3 = g.as3d(om.get(fn))
        want = g.as3d(t0[tn]).copy()
        L3, H3 = g.as3d(lo[v]), g.as3d(hi[v])
        for j in range(g.ny + jext):
            for i in range(g.nx + iext):
                w = w2[min(j, g.ny - 1) + g.ng, min(i, g.nx - 1) + g.ng]
                ks = slice(g.ng, g.ng + g.nz)
                jj, ii = j + g.ng, i + g.ng
                want[ks, jj, ii] += w * ((1 - tf) * L3[ks, jj, ii] + tf * H3[ks, jj, ii]
                                         - f3[ks, jj, ii])
        got = g.as3d(t1[tn])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15, err_msg=tn)


# ----------------------------------------------------- advection accuracy
def _advection_error(nx: int) -> float:
    L = 40000.0
    dx = L / nx
    c = case("warm_bubble", nx=nx, ny=3, nz=6, dx=dx, dy=dx, dz=500, bubble_dtheta=0,
             wind_u=10.0)
    g = c.grid
    om = OracleModel.from_case(c)
    x = (np.arange(g.NX) - g.ng + 0.5) * dx
    th = 0.7 * np.sin(2 * np.pi * x / L) + 0.2 * np.cos(6 * np.pi * x / L)
    dth = 0.7 * (2 * np.pi / L) * np.cos(2 * np.pi * x / L) \
        - 0.2 * (6 * np.pi / L) * np.sin(6 * np.pi * x / L)
    a = np.broadcast_to(th, (g.NZ, g.NY, g.NX)).copy()
    om.set("thp", a.ravel())
    om.apply_bcs()
    om.tendencies(1.0)
    t = g.interior(om.get("tthp"))
    exact = -10.0 * dth[g.ng:g.ng + nx]
    return float(np.abs(t - exact[None, None, :]).max() / np.abs(exact).max())


def test_flux_form_advection_converges_at_fifth_order():
    # uniform u over flat terrain and a horizontally homogeneous base state:
    # div = 0 and T_theta = -u dtheta'/dx exactly; the flux difference of
    # the 5th-order interface values converges as dx^5.
    e1, e2 = _advection_error(32), _advection_error(64)
    assert e1 < 5e-3
    rate = math.log2(e1 / e2)
    assert rate > 4.6, (e1, e2, rate)
