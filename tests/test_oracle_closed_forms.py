"""Closed-form pins of the CPU oracle (no GPU, no golden files).

Every model here is assembled from numpy arrays written in this file (a
constant or hand-built base state and metric), never from the product's host
builders, and every expected value comes from the mathematics of the scheme:

- an f-plane inertial oscillation advanced by whole RK3 steps against the
  WS2002 amplification polynomial 1 + z + z^2/2 + z^3/6, z = -i f dt
  (src/dynamics/integrator.cpp:71-111, kernels.cu:340-345,378-383);
- a vertical acoustic mode through the off-centred implicit column against
  its 2 x 2 closed form (kernels.cu:531-645; the (1 +- beta)/2 weights);
- a horizontal acoustic mode on a uniformly scaled map (m != 1) with the
  divergence damping pi* = pi' + smdiv (pi' - pi'prev), one substep, several
  substeps and a whole large step (kernels.cu:480-527, 496-512, 929-965);
- one-level Kessler forms: the Klemp-Wilhelmson terminal velocity and one
  sedimentation subcycle, the surface rain flux, and the rain evaporation
  rate (src/physics/kessler.cu:47-75,100-116).

Each case is chosen so that a plausible slip in the oracle fails it: a
flipped Coriolis sign, a wrong RK3 stage fraction, an inverted or dropped
map factor, swapped off-centring weights, a flipped divergence-damping sign,
a wrong evaporation or fall-speed constant, a density ratio upside down.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import OracleModel

NG = 3
CP, RD = 1004.5, 287.04          # src/core/constants.hpp
CV = CP - RD
GRAV, LV, P00 = 9.81, 2.5e6, 1.0e5


class _Grid:
    """Padded layout of src/core/grid.hpp:22-36, restated for the tests."""

    def __init__(self, nx, ny, nz, dx, dy, dz):
        self.nx, self.ny, self.nz, self.ng = nx, ny, nz, NG
        self.dx, self.dy, self.dz = dx, dy, dz
        self.NX, self.NY, self.NZ = nx + 2 * NG, ny + 2 * NG, nz + 1 + 2 * NG
        self.npts = self.NX * self.NY * self.NZ

    def as3d(self, a):
        return a.reshape(self.NZ, self.NY, self.NX)


def _uniform_model(g: _Grid, params: dict, *, th=300.0, pi=1.0, rho=1.0, mapf=1.0,
                   fcor=0.0, levels=None, base_cols=None) -> OracleModel:
    """Oracle over flat ground with a constant base state (thb = thv =
    th, pib = pi, rho) unless `base_cols` gives per-level profiles, geometric
    levels of spacing dz (or `levels`, the w-level heights k = -ng..NZ-ng-1),
    J = 1, and uniform map factor and Coriolis parameter."""
    n3, n2, n1 = g.npts, g.NX * g.NY, g.NZ
    kk = np.arange(n1) - NG
    zw = kk * g.dz if levels is None else np.asarray(levels, dtype=float)
    zc = np.empty(n1)
    zc[:-1] = 0.5 * (zw[:-1] + zw[1:])
    zc[-1] = zw[-1] + 0.5 * (zw[-1] - zw[-2])
    dzc = np.empty(n1)
    dzc[:-1] = zw[1:] - zw[:-1]
    dzc[-1] = dzc[-2]
    dzw = np.empty(n1)
    dzw[1:] = zc[1:] - zc[:-1]
    dzw[0] = dzw[1]
    zt = float(zw[NG + g.nz])

    def col(v):
        a = np.broadcast_to(np.asarray(v, dtype=float).reshape(-1, 1, 1), (n1, g.NY, g.NX))
        return np.ascontiguousarray(a).ravel()

    b = base_cols or {}
    thc = col(b.get("thb", np.full(n1, th)))
    arrays = {
        "thb": thc, "thvb": thc, "thbw": col(b.get("thbw", np.full(n1, th))),
        "thvbw": col(b.get("thbw", np.full(n1, th))),
        "pib": col(b.get("pib", np.full(n1, pi))), "rhob": col(b.get("rhob", np.full(n1, rho))),
        "rhobw": col(b.get("rhobw", np.full(n1, rho))), "dthbdz": np.zeros(n3),
        "qvb": np.zeros(n3), "ub": np.zeros(n1), "vb": np.zeros(n1),
        "zeta_w": zw, "zeta_c": zc, "dzeta_c": dzc, "dzeta_w": dzw,
        "h": np.zeros(n2), "hx": np.zeros(n2), "hy": np.zeros(n2), "hx_u": np.zeros(n2),
        "hy_v": np.zeros(n2), "jac": np.ones(n2), "fcor": np.full(n2, fcor),
        "mapf": np.full(n2, mapf),
    }
    p = dict(diff_K=0.0, coriolis_use_ub=1, rayleigh_zd=-1.0, rayleigh_alpha=0.0,
             acoustic_ns=6, acoustic_beta=0.2, acoustic_smdiv=0.1, bc_x_open=0, bc_y_open=0,
             cstar=30.0, moisture=0, microphysics=0, pd_moist=0, w_damping=0)
    p.update(params)
    return OracleModel(g, zt, p, arrays)


# ------------------------------------------------------ inertial oscillation
@pytest.mark.parametrize("f,dt,nsteps", [(1e-4, 60.0, 12), (1.2e-3, 100.0, 9)])
def test_inertial_oscillation_follows_rk3_amplification(f, dt, nsteps):
    """Uniform (u, v) on an f-plane: every advective, pressure and damping
    term vanishes, so du/dt = f v, dv/dt = -f u, i.e. Z = u + i v obeys
    Z' = -i f Z. The WS2002 stages {dt/3, dt/2, dt} each restarted from
    time n give Z_{n+1} = R(z) Z_n with R = 1 + z + z^2/2 + z^3/6, z = -i f dt
    (Wicker and Skamarock 2002, eq. 3.3), whatever the substep count."""
    g = _Grid(6, 4, 5, 1000.0, 1000.0, 500.0)
    om = _uniform_model(g, {"acoustic_ns": 4}, fcor=f)
    u0, v0 = 7.0, -3.0
    om.set("u", np.full(g.npts, u0))
    om.set("v", np.full(g.npts, v0))
    z = -1j * f * dt
    R = 1 + z + z * z / 2 + z ** 3 / 6
    Z = complex(u0, v0)
    for n in range(nsteps):
        om.step(dt)
        Z *= R
        u = g.as3d(om.get("u"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + g.nx]
        v = g.as3d(om.get("v"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + g.nx]
        np.testing.assert_allclose(u, Z.real, rtol=0, atol=1e-12 * abs(Z), err_msg=f"u {n}")
        np.testing.assert_allclose(v, Z.imag, rtol=0, atol=1e-12 * abs(Z), err_msg=f"v {n}")
    # The clockwise sense (northern hemisphere, f > 0): the phase of Z falls.
    assert np.angle(Z / complex(u0, v0)) < 0


# ----------------------------------------------------- vertical acoustic mode
def _vertical_mode_map(a, dz, dtau, beta, th, pi):
    """One off-centred substep of w' = -cp th d(pi')/dz, pi'' = -(Rd/cv) pi
    dw/dz for the mode w_k = W sin(a k), pi'_k = P cos(a (k + 1/2)): with
    s = 2 sin(a/2)/dz, new-level weight c1 = (1 + beta)/2, old c2 = (1 - beta)/2,
      W1 = W0 + al s (c1 P1 + c2 P0),  P1 = P0 - ga s (c1 W1 + c2 W0),
    al = dtau cp th, ga = dtau Rd pi / cv. Returns the 2 x 2 matrix."""
    c1, c2 = 0.5 * (1 + beta), 0.5 * (1 - beta)
    s = 2 * math.sin(a / 2) / dz
    al, ga = dtau * CP * th, dtau * RD * pi / CV
    lhs = np.array([[1.0, -al * s * c1], [ga * s * c1, 1.0]])
    rhs = np.array([[1.0, al * s * c2], [-ga * s * c2, 1.0]])
    return np.linalg.solve(lhs, rhs)


def _vertical_mode_model(mode, beta, smdiv, nz=16, dz=250.0, th=300.0, pi=0.95):
    g = _Grid(4, 3, nz, 1000.0, 1000.0, dz)
    om = _uniform_model(g, {"acoustic_beta": beta, "acoustic_smdiv": smdiv}, th=th, pi=pi)
    a = math.pi * mode / nz
    k = np.arange(g.NZ) - NG
    w = np.where((k >= 0) & (k <= nz), VW0 * np.sin(a * k), 0.0)
    p = VP0 * np.cos(a * (k + 0.5))
    col = lambda c: np.ascontiguousarray(np.broadcast_to(c[:, None, None], (g.NZ, g.NY, g.NX)))  # noqa
    om.set("w", col(w).ravel())
    om.set("pip", col(p).ravel())
    return g, om, a


VW0, VP0 = 0.8, 3e-4


@pytest.mark.parametrize("mode,beta,nsub,smdiv", [(1, 0.2, 1, 0.1), (2, 0.2, 2, 0.1),
                                                  (3, 0.5, 5, 0.0), (1, 0.0, 4, 0.0)])
def test_vertical_acoustic_mode_matches_two_by_two_closed_form(mode, beta, nsub, smdiv):
    """A horizontally uniform w-pi' mode between the rigid ground and lid
    (w = 0 at k = 0 and nz) over a constant base state: buoyancy and the
    horizontal step vanish, and the tridiagonal solve must reproduce the
    off-centred two-level scheme exactly. With beta > 0 the mode damps; with
    the weights swapped it would grow. (With smdiv != 0 the stale pi'prev
    ghost cells force the edge faces from the third substep on; that case
    has its own test below.)"""
    nz, dz, dtau, th, pi = 16, 250.0, 1.0, 300.0, 0.95
    g, om, a = _vertical_mode_model(mode, beta, smdiv, nz, dz, th, pi)
    W0, P0 = VW0, VP0
    om.acoustic(dtau, nsub)
    M = np.linalg.matrix_power(_vertical_mode_map(a, dz, dtau, beta, th, pi), nsub)
    W1, P1 = M @ np.array([W0, P0])
    kk = np.arange(nz + 1)
    w_got = g.as3d(om.get("w"))[NG:NG + nz + 1, NG:NG + g.ny, NG:NG + g.nx]
    p_got = g.as3d(om.get("pip"))[NG:NG + nz, NG:NG + g.ny, NG:NG + g.nx]
    np.testing.assert_allclose(w_got, (W1 * np.sin(a * kk))[:, None, None] + 0 * w_got,
                               rtol=0, atol=1e-12 * W0)
    np.testing.assert_allclose(p_got, (P1 * np.cos(a * (kk[:nz] + 0.5)))[:, None, None]
                               + 0 * p_got, rtol=0, atol=1e-12 * P0)
    # energy-like norm: damped for beta > 0, conserved to O(dtau^2) at beta = 0
    ratio = np.abs(np.linalg.eigvals(M)).max()
    assert ratio < 1.0 if beta > 0 else abs(ratio - 1.0) < 1e-12


def test_stale_piprev_ghosts_force_the_edge_faces_on_the_third_substep():
    """pi'prev is copied from time n once per stage (integrator.cpp:94) and
    the substeps refresh only its interior (kernels.cu:640-644) while pi'
    itself gets its ghosts back after every column solve (:947-963). On the
    third substep of a horizontally uniform mode the west face i = 0 (and the
    south face j = 0) therefore sees
      pi*(0) - pi*(-1) = smdiv (pi'_n - pi'_1)   (periodic ghosts),
    so u(0) = -dtau cp th smdiv (pi'_n - pi'_1) / dx, while every other face
    stays at rest."""
    beta, smdiv, dtau, th, pi, nz, dz = 0.2, 0.1, 1.0, 300.0, 0.95, 16, 250.0
    g, om, a = _vertical_mode_model(1, beta, smdiv, nz, dz, th, pi)
    om.acoustic(dtau, 3)
    M1 = _vertical_mode_map(a, dz, dtau, beta, th, pi)
    P1 = (M1 @ np.array([VW0, VP0]))[1]
    kk = np.arange(nz)
    d_pi = (VP0 - P1) * np.cos(a * (kk + 0.5))  # pi'_n - pi'_1 per level
    u = g.as3d(om.get("u"))[NG:NG + nz, NG:NG + g.ny, NG:NG + g.nx]
    v = g.as3d(om.get("v"))[NG:NG + nz, NG:NG + g.ny, NG:NG + g.nx]
    want0 = -dtau * CP * th * smdiv * d_pi
    np.testing.assert_allclose(u[:, :, 0], np.broadcast_to((want0 / g.dx)[:, None],
                               (nz, g.ny)), rtol=1e-11, atol=1e-16)
    np.testing.assert_allclose(v[:, 0, :], np.broadcast_to((want0 / g.dy)[:, None],
                               (nz, g.nx)), rtol=1e-11, atol=1e-16)
    assert np.abs(u[:, :, 1:]).max() == 0.0 and np.abs(v[:, 1:, :]).max() == 0.0


# --------------------------------------------------- horizontal acoustic mode
def _horizontal_mode(kappa, dx, dtau, th, pi, m, smdiv, nsub, U0, P0):
    """Complex amplitudes of u_i = Re(U e^{i kappa i dx}) (west faces) and
    pi'_i = Re(P e^{i kappa (i + 1/2) dx}) (cells) after nsub forward-backward
    substeps on a map of uniform scale m:
      U1 = U0 - dtau cp th m D P*,   P* = P0 + smdiv (P0 - Pprev),
      P1 = P0 - dtau (Rd pi / cv) m D U1,   D = 2 i sin(kappa dx / 2) / dx,
    with Pprev = P0 on the first substep (pi'prev = pi'(n) at stage start)."""
    D = 2j * math.sin(kappa * dx / 2) / dx
    U, P, Pprev = complex(U0), complex(P0), complex(P0)
    for _ in range(nsub):
        Ps = P + smdiv * (P - Pprev)
        U = U - dtau * CP * th * m * D * Ps
        Pprev, P = P, P - dtau * (RD * pi / CV) * m * D * U
    return U, P


@pytest.mark.parametrize("m,smdiv,nsub", [(1.04, 0.1, 1), (0.95, 0.1, 2), (1.05, 0.3, 2),
                                          (1.05, 0.0, 6), (1.0, 0.0, 3)])
def test_horizontal_acoustic_mode_on_scaled_map(m, smdiv, nsub):
    """An x-travelling sound wave on a grid whose map factor is a constant m:
    the pressure-gradient force and the divergence both carry m (the
    physical gradient is m d/dx), so the discrete wave must follow the
    closed-form recursion; pi* carries the divergence damping from the
    second substep on. (Beyond two substeps with smdiv != 0 the stale
    pi'prev ghosts break the mode at the edge faces; see below.)"""
    nx, dx, dtau, th, pi = 16, 2000.0, 1.5, 300.0, 0.97
    g = _Grid(nx, 3, 4, dx, dx, 500.0)
    om = _uniform_model(g, {"acoustic_smdiv": smdiv}, th=th, pi=pi, mapf=m)
    kappa = 2 * math.pi * 2 / (nx * dx)
    U0, P0 = 0.6 + 0.2j, 4e-4 - 1e-4j
    i = np.arange(g.NX) - NG
    u_row = (U0 * np.exp(1j * kappa * i * dx)).real
    p_row = (P0 * np.exp(1j * kappa * (i + 0.5) * dx)).real
    bc = lambda row: np.ascontiguousarray(np.broadcast_to(row, (g.NZ, g.NY, g.NX))).ravel()  # noqa
    om.set("u", bc(u_row))
    om.set("pip", bc(p_row))
    om.acoustic(dtau, nsub)
    U1, P1 = _horizontal_mode(kappa, dx, dtau, th, pi, m, smdiv, nsub, U0, P0)
    ii = np.arange(nx)
    u_want = (U1 * np.exp(1j * kappa * ii * dx)).real
    p_want = (P1 * np.exp(1j * kappa * (ii + 0.5) * dx)).real
    u_got = g.as3d(om.get("u"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + nx]
    p_got = g.as3d(om.get("pip"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + nx]
    np.testing.assert_allclose(u_got, np.broadcast_to(u_want, u_got.shape), rtol=0,
                               atol=1e-12 * abs(U0))
    np.testing.assert_allclose(p_got, np.broadcast_to(p_want, p_got.shape), rtol=0,
                               atol=1e-12 * abs(P0))
    w_got = g.as3d(om.get("w"))[NG:NG + g.nz + 1, NG:NG + g.ny, NG:NG + nx]
    assert np.abs(w_got).max() < 1e-14


@pytest.mark.parametrize("ns", [4, 6])
def test_horizontal_acoustic_mode_over_a_large_step(ns):
    """A whole RK3 step of a small-amplitude sound wave: the slow tendencies
    are quadratic in the amplitude (advection only), so each stage is the
    fast system restarted from time n, and the step result is the third
    stage's ns substeps of dt/ns (integrator.cpp:71-111). smdiv = 0 keeps the
    stale pi'prev ghosts out of the closed form."""
    nx, dx, dt, th, pi, m, smdiv = 16, 2000.0, 9.0, 300.0, 0.97, 1.03, 0.0
    g = _Grid(nx, 3, 4, dx, dx, 500.0)
    om = _uniform_model(g, {"acoustic_smdiv": smdiv, "acoustic_ns": ns}, th=th, pi=pi,
                        mapf=m)
    kappa = 2 * math.pi / (nx * dx)
    U0, P0 = 1e-7, 2e-11j
    i = np.arange(g.NX) - NG
    bc = lambda row: np.ascontiguousarray(np.broadcast_to(row, (g.NZ, g.NY, g.NX))).ravel()  # noqa
    om.set("u", bc((U0 * np.exp(1j * kappa * i * dx)).real))
    om.set("pip", bc((P0 * np.exp(1j * kappa * (i + 0.5) * dx)).real))
    om.step(dt)
    U1, P1 = _horizontal_mode(kappa, dx, dt / ns, th, pi, m, smdiv, ns, U0, P0)
    ii = np.arange(nx)
    u_got = g.as3d(om.get("u"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + nx]
    p_got = g.as3d(om.get("pip"))[NG:NG + g.nz, NG:NG + g.ny, NG:NG + nx]
    u_want = (U1 * np.exp(1j * kappa * ii * dx)).real
    p_want = (P1 * np.exp(1j * kappa * (ii + 0.5) * dx)).real
    np.testing.assert_allclose(u_got, np.broadcast_to(u_want, u_got.shape), rtol=0,
                               atol=1e-8 * abs(U0))
    np.testing.assert_allclose(p_got, np.broadcast_to(p_want, p_got.shape), rtol=0,
                               atol=1e-8 * abs(P0))


# ------------------------------------------------------------------- Kessler
def _column(nz=8, dz=400.0):
    """A 4 x 1 column set with a hand-made stratified base state: T falls
    6.5 K/km from 290 K, pi from 1, rho from 1.2 kg/m3 (not hydrostatic;
    Kessler only reads the local values)."""
    g = _Grid(4, 1, nz, 1000.0, 1000.0, dz)
    k = np.arange(g.NZ) - NG
    z = (k + 0.5) * dz
    pib = 1.0 - 3.2e-5 * z
    T = 290.0 - 6.5e-3 * z
    thb = T / pib
    rhob = 1.2 * np.exp(-z / 9000.0)
    base = {"thb": thb, "thbw": thb, "pib": pib, "rhob": rhob, "rhobw": rhob}
    om = _uniform_model(g, {"moisture": 1, "microphysics": 1}, base_cols=base)
    return g, om, dict(pib=pib, thb=thb, rhob=rhob, T=T)


def _fall_speed(rho, q, rho_sfc, T):
    """Klemp and Wilhelmson (1978) terminal velocity of rain in SI form,
    V = 36.34 (rho q_r)^0.1364 (rho_0 / rho)^(1/2), times 0.35 below
    freezing (kessler.cu:47-52)."""
    return 36.34 * (rho * q) ** 0.1364 * math.sqrt(rho_sfc / rho) * (0.35 if T < 273.15 else 1.0)


def _set_col(g, om, name, values):
    a = np.zeros(g.npts)
    a3 = g.as3d(a)
    for k, v in values.items():
        a3[NG + k, NG, NG:NG + g.nx] = v
    om.set(name, a)


def _get_col(g, om, name):
    return g.as3d(om.get(name))[NG:NG + g.nz, NG, NG]


@pytest.mark.parametrize("k0,qr0", [(3, 2e-3), (6, 1.5e-3), (0, 1e-3)])
def test_kessler_single_subcycle_sedimentation(k0, qr0):
    """Rain at one level of a supersaturated column (so no evaporation) and
    no cloud (so no autoconversion or accretion): one sedimentation subcycle
    moves F = rho_k q V_k out of level k into level k - 1, scaled by each
    level's rho dz, and at k = 0 into the surface rain (kg m-2)."""
    dt = 10.0
    g, om, b = _column()
    qv = {k: 0.05 for k in range(g.nz)}  # far above saturation: condensation only
    _set_col(g, om, "qv", qv)
    _set_col(g, om, "qr", {k0: qr0})
    om.kessler(dt)
    K = lambda k: k + NG  # noqa: E731
    rho = b["rhob"]
    V = _fall_speed(rho[K(k0)], qr0, rho[K(0)], b["T"][K(k0)])
    assert V * dt / (0.8 * g.dz) < 1.0  # one subcycle
    F = rho[K(k0)] * qr0 * V
    qr = _get_col(g, om, "qr")
    want = np.zeros(g.nz)
    want[k0] = qr0 - dt * F / (rho[K(k0)] * g.dz)
    if k0 > 0:
        want[k0 - 1] = dt * F / (rho[K(k0 - 1)] * g.dz)
    np.testing.assert_allclose(qr, want, rtol=1e-13, atol=1e-18)
    rain = g.as3d(np.zeros(g.npts))  # noqa: F841  (shape only)
    r = om.get("rain").reshape(g.NY, g.NX)[NG, NG]
    assert r == pytest.approx(dt * F if k0 == 0 else 0.0, rel=1e-13, abs=1e-20)


def test_kessler_fall_speed_below_freezing():
    """The 0.35 ice factor applies where T = (thb + th') (pib + pi') < 273.15."""
    dt = 10.0
    g, om, b = _column(nz=12, dz=600.0)
    K = lambda k: k + NG  # noqa: E731
    k0 = next(k for k in range(g.nz) if b["T"][K(k)] < 273.15)
    _set_col(g, om, "qv", {k: 0.05 for k in range(g.nz)})
    _set_col(g, om, "qr", {k0: 2e-3})
    om.kessler(dt)
    rho = b["rhob"]
    V = _fall_speed(rho[K(k0)], 2e-3, rho[K(0)], b["T"][K(k0)])
    qr = _get_col(g, om, "qr")
    assert qr[k0 - 1] == pytest.approx(dt * rho[K(k0)] * 2e-3 * V / (rho[K(k0 - 1)] * g.dz),
                                       rel=1e-13)


def _tetens(p, T):
    """Saturation mixing ratio over water from Tetens' formula,
    e_s = 610.78 exp(17.269 (T - 273.16) / (T - 35.86)) Pa, r = 0.622 e_s/(p - e_s)."""
    es = 610.78 * math.exp(17.269 * (T - 273.16) / (T - 35.86))
    return 0.622 * es / (p - es)


@pytest.mark.parametrize("rh,qr0", [(0.6, 1e-3), (0.9, 3e-4)])
def test_kessler_rain_evaporation_rate(rh, qr0):
    """Rain falling through subsaturated air at a warm level, no cloud: the
    evaporation is the Klemp-Wilhelmson (1978, eq. 2.15) rate
      E = (1/rho) (1 - q_v/q_s) C (rho q_r)^0.525 / (5.4e5 + 2.55e8/(p q_s)),
      C = 1.6 + 124.9 (rho q_r)^0.2046,
    taken over dt (kessler.cu:100-116), cooling by L_v dq/(c_p pi). The
    saturation adjustment that follows has no cloud to evaporate, so q_v
    only gains the rain evaporation. Sedimentation is avoided by putting the
    rain at k = 0 and measuring what stays."""
    dt = 2.0
    g, om, b = _column()
    K = lambda k: k + NG  # noqa: E731
    k0 = 2
    pi = b["pib"][K(k0)]
    T = b["T"][K(k0)]
    p = P00 * pi ** (CP / RD)
    qs = _tetens(p, T)
    qv0 = rh * qs
    _set_col(g, om, "qv", {k0: qv0})
    _set_col(g, om, "qr", {k0: qr0})
    # rain sediments out of k0 first: one subcycle, closed form as above
    rho = b["rhob"][K(k0)]
    V = _fall_speed(rho, qr0, b["rhob"][K(0)], T)
    qr_after_fall = qr0 - dt * rho * qr0 * V / (rho * g.dz)
    om.kessler(dt)
    rq = rho * qr_after_fall
    C = 1.6 + 124.9 * rq ** 0.2046
    E = (1.0 / rho) * (1 - qv0 / qs) * C * rq ** 0.525 / (5.4e5 + 2.55e8 / (p * qs))
    dq = min(dt * E, qr_after_fall, qs - qv0)
    assert 0 < dq < qs - qv0
    np.testing.assert_allclose(_get_col(g, om, "qv")[k0], qv0 + dq, rtol=1e-12)
    np.testing.assert_allclose(_get_col(g, om, "qr")[k0], qr_after_fall - dq, rtol=1e-12)
    np.testing.assert_allclose(_get_col(g, om, "thp")[k0], -LV * dq / (CP * pi), rtol=1e-12)
