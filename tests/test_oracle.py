"""Pins of the CPU oracle against closed forms, library routines and the
reference's own invariants (so that the oracle can be trusted as the checker
of the HIP path).

iface5 (kernels.cu:51-57); Thomas order (kernels.cu:589-628); Tetens
(thermo.hpp:11-26); whole-step invariants from docs/EQUATIONS.md:61-78 and
tools/run_tests.py:72-101 of the reference.
"""
import math

import numpy as np
import pytest
from scipy.linalg import solve_banded

import oracle_lib as O
from helpers import case, kitchen_sink, perturbed_state, rel_err
from oracle_lib import OracleModel


# ---------------------------------------------------------------- iface5
def _iface_poly(coef, vel, x0=0.0):
    q = [np.polyval(coef, x0 + s) for s in (-2, -1, 0, 1, 2, 3)]
    return O.iface5(*q, vel)


def test_iface5_constant_and_linear():
    assert O.iface5(*[3.25] * 6, 1.0) == pytest.approx(3.25, abs=1e-15)
    # linear q = a + b s: interface value at s = 1/2 (cell-average sense)
    for vel in (1.0, -1.0):
        v = _iface_poly([2.0, 5.0], vel)
        assert v == pytest.approx(5.0 + 2.0 * 0.5, abs=1e-13)


def test_iface5_upwind_term_is_fifth_difference():
    # cen - upw and cen + upw differ by 2 upw; upw = 5th undivided difference / 60
    for k in range(5):  # degree <= 4: no upwind correction
        coef = np.zeros(k + 1)
        coef[0] = 1.0
        assert _iface_poly(coef, 1.0) == pytest.approx(_iface_poly(coef, -1.0), abs=1e-10)
    q5 = [s ** 5 for s in (-2, -1, 0, 1, 2, 3)]  # 5th difference of s^5 = 120
    d = O.iface5(*q5, -1.0) - O.iface5(*q5, 1.0)
    assert d == pytest.approx(2 * 120 / 60, abs=1e-12)


def test_iface5_zero_velocity_takes_the_nonnegative_branch():
    q = [0.0, 0.0, 1.0, 5.0, 0.0, 0.0]
    assert O.iface5(*q, 0.0) == O.iface5(*q, 1.0)
    assert O.iface5(*q, 0.0) != O.iface5(*q, -1.0)


def test_iface5_reproduces_face_value_from_cell_averages():
    # Cell averages of f(x) = x^k over [s - 1/2, s + 1/2]: the face value at
    # x = 1/2 is reproduced exactly by the centred (6th-order) part for
    # degree <= 5 and by either upwind branch (5th order) for degree <= 4.
    def avg(k, s):
        return ((s + 0.5) ** (k + 1) - (s - 0.5) ** (k + 1)) / (k + 1)
    for k in range(7):
        q = [avg(k, s) for s in (-2, -1, 0, 1, 2, 3)]
        up, dn = O.iface5(*q, 1.0), O.iface5(*q, -1.0)
        exact = 0.5 ** k
        centred = 0.5 * (up + dn)
        if k <= 5:
            assert centred == pytest.approx(exact, abs=1e-12), k
        else:
            assert abs(centred - exact) > 1e-6
        if k <= 4:
            assert up == pytest.approx(exact, abs=1e-12), k
            assert dn == pytest.approx(exact, abs=1e-12), k
        if k == 5:
            assert abs(up - exact) > 1e-3 and abs(dn - exact) > 1e-3


# ---------------------------------------------------------------- Thomas
def test_thomas_matches_lapack_banded_solver():
    rng = np.random.default_rng(7)
    for n in (2, 5, 39, 44, 120):
        a = rng.uniform(-1, 0, n)
        c = rng.uniform(-1, 0, n)
        a[0] = 0.0
        c[-1] = 0.0
        b = 1.0 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, n)
        d = rng.standard_normal(n)
        x = O.thomas(a, b, c, d)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1] = b
        ab[2, :-1] = a[1:]
        ref = solve_banded((1, 1), ab, d)
        np.testing.assert_allclose(x, ref, rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------- Tetens
def test_qsat_tetens_values_and_floor():
    T, p = 293.15, 1.0e5
    es = 610.78 * math.exp(17.269 * (T - 273.16) / (T - 35.86))
    assert O.olib().orc_qsat_w(p, T) == pytest.approx(0.622 * es / (p - es), rel=1e-15)
    # 20 C saturation vapour pressure is ~2.34 kPa (textbook), q_s ~ 14.7 g/kg
    assert 2300 < es < 2400
    # floor: p - e_s < 1 Pa -> denominator 1
    assert O.olib().orc_qsat_w(100.0, 330.0) == pytest.approx(
        0.622 * 610.78 * math.exp(17.269 * (330 - 273.16) / (330 - 35.86)), rel=1e-14)
    Ti = 253.15
    esi = 610.78 * math.exp(21.875 * (Ti - 273.16) / (Ti - 7.66))
    assert O.olib().orc_qsat_i(p, Ti) == pytest.approx(0.622 * esi / (p - esi), rel=1e-15)
    assert O.olib().orc_qsat_i(p, Ti) < O.olib().orc_qsat_w(p, Ti)


# ------------------------------------------------------- whole-step invariants
def test_rest_over_terrain_stays_exactly_at_rest():
    c = case("schaer_rest", nx=60, ny=4, nz=30, mtn_xc=15000)
    om = OracleModel.from_case(c)
    om.run(2.5, 40)
    g = c.grid
    assert np.abs(g.interior(om.get("w"), g.nz + 1)).max() == 0.0
    assert np.abs(g.interior(om.get("thp"))).max() == 0.0
    assert np.abs(g.interior(om.get("pip"))).max() == 0.0


def test_uniform_flow_has_zero_slow_tendency():
    # flat terrain, horizontally homogeneous base state, constant u: div = 0
    # and the advection of a uniform field vanishes identically.
    c = case("bubble_outflow", nx=20, ny=12, nz=16, bubble_dtheta=0, bc_x="periodic")
    om = OracleModel.from_case(c)
    om.apply_bcs()
    om.tendencies(1.0)
    g = c.grid
    for f in ("tu", "tv", "tw", "tthp"):
        assert np.abs(g.interior(om.get(f))).max() < 1e-12, f


def test_mirror_symmetry_of_warm_bubble():
    # x <-> y symmetric set-up: u(k, j, i) == v(k, i, j) (reference reports
    # 2e-6 in FP32 at full size; in FP64 on a small grid it is ~1e-12).
    c = case("warm_bubble", nx=24, ny=24, nz=20, dx=400, dy=400, dz=400, bubble_xc=4800,
             bubble_yc=4800, bubble_zc=2000)
    om = OracleModel.from_case(c)
    om.run(1.5, 40)
    g = c.grid
    u = g.interior(om.get("u"))
    v = g.interior(om.get("v"))
    sym = np.sqrt(((u - np.transpose(v, (0, 2, 1))) ** 2).mean()) / np.sqrt((u ** 2).mean())
    assert np.sqrt((u ** 2).mean()) > 1e-3
    assert sym < 2e-6


def test_pd_moisture_stays_positive_and_fluxes_telescope():
    c = case("moist_blob", nx=60, ny=4, nz=20, qvbub_xc=3000, qvbub_zc=2000)
    om = OracleModel.from_case(c)
    g = c.grid
    om.run(1.0, 120)
    qv = g.interior(om.get("qv"))
    assert qv.min() >= -1e-12
    # The renormalised fluxes telescope: over the periodic domain with
    # Omega = 0 at the bottom and top, sum(dzeta (rho~ T_q - q div)) = 0.
    om.apply_bcs()
    om.tendencies(1.0)
    rt = g.interior(c.array("rhob")) * c.array("jac").reshape(g.NY, g.NX)[
        g.ng:g.ng + g.ny, g.ng:g.ng + g.nx][None]
    dz = c.array("dzeta_c")[g.ng:g.ng + g.nz][:, None, None]
    tq = g.interior(om.get("tqv"))
    dv = g.interior(om.get("div"))
    q = g.interior(om.get("qv"))
    flux_div = dz * (rt * tq - q * dv)
    assert np.abs(tq).max() > 1e-8
    assert abs(flux_div.sum()) < 1e-12 * np.abs(flux_div).sum()


def test_non_pd_transport_undershoots():
    # the same blob without the limiter produces negative qv (reference: 5265
    # negative cells at full size) -- the PD switch is what fixes it.
    c = case("moist_blob", nx=60, ny=4, nz=20, qvbub_xc=3000, qvbub_zc=2000, pd_moist="off")
    om = OracleModel.from_case(c)
    om.run(1.0, 120)
    assert om.get("qv").min() < -1e-6


def test_kessler_dry_column_unchanged():
    c = kitchen_sink(nx=8, ny=6, nz=12, bubble_dtheta=0)
    om = OracleModel.from_case(c)
    g = c.grid
    qv0 = om.get("qv") * 0.3  # well below saturation
    om.set("qv", qv0)
    th0 = om.get("thp")
    om.kessler(6.0)
    np.testing.assert_array_equal(g.interior(om.get("qv")), g.interior(qv0))
    np.testing.assert_array_equal(g.interior(om.get("thp")), g.interior(th0))
    assert np.all(om.get("rain") == 0)


def test_kessler_conserves_water_and_enthalpy_like_quantity():
    # total water + rain that reached the ground is conserved; latent heating
    # matches the condensed mass: d(theta') = L dq / (cp pi).
    c = kitchen_sink(nx=8, ny=6, nz=14)
    om = OracleModel.from_case(c)
    st = perturbed_state(c, 3, {"u": 0, "v": 0, "w": 0, "thp": 0, "pip": 0,
                                "qv": 4e-3, "qc": 2e-3, "qr": 2e-3})
    for f, a in st.items():
        om.set(f, a)
    g = c.grid
    rho = g.interior(c.array("rhob"))
    jac = c.array("jac").reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
    dzl = c.array("dzeta_c")[g.ng:g.ng + g.nz][:, None, None] * jac[None]
    def col_water():
        q = sum(g.interior(om.get(f)) for f in ("qv", "qc", "qr"))
        return (rho * dzl * q).sum(axis=0)
    w0 = col_water()
    om.kessler(6.0)
    rain = om.get("rain").reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
    w1 = col_water()
    assert rain.max() > 0
    np.testing.assert_allclose(w1 + rain, w0, rtol=1e-10)


def test_kessler_autoconversion_closed_form():
    # qc above qc0, no rain, saturated enough that nothing evaporates:
    # dqr = dt (k1 (qc - qc0) + 2.2 qc (1e-9)^0.875) exactly.
    c = kitchen_sink(nx=8, ny=6, nz=12, bubble_dtheta=0)
    om = OracleModel.from_case(c)
    g = c.grid
    qc = np.zeros(g.npts)
    g.as3d(qc)[g.ng + 3, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = 3e-3
    om.set("qc", qc)
    qv = om.get("qv")
    g.as3d(qv)[g.ng + 3] *= 3.0  # supersaturated: no rain evaporation at that level
    om.set("qv", qv)
    dt = 2.0
    om.kessler(dt)
    qr = g.interior(om.get("qr"))[3]
    expect = dt * (1e-3 * (3e-3 - 1e-3) + 2.2 * 3e-3 * (1e-9) ** 0.875)
    assert qr.max() == pytest.approx(expect, rel=1e-12)


def test_step_is_deterministic_across_thread_counts():
    c = kitchen_sink()
    st = perturbed_state(c, 11)
    res = []
    for nt in (1, 4):
        O.set_threads(nt)
        om = OracleModel.from_case(c)
        for f, a in st.items():
            om.set(f, a)
        om.run(6.0, 2)
        res.append(om.get("u"))
    O.set_threads(8)
    assert rel_err(res[0], res[1]) == 0.0
