"""Host builders (Config, Metric, BaseState, ICs) against closed forms.

Config semantics follow src/core/config.cpp:17-60; Metric src/core/metric.cpp:
12-168; BaseState src/core/base_state.cpp:21-193; ICs src/main.cpp:28-195.
"""
import math

import numpy as np
import pytest

from helpers import CASES, case
from wfe_amd import Case, WfeError

G, CP, RD = 9.81, 1004.5, 287.04
CV, P00 = CP - RD, 1.0e5


def test_config_parser_semantics(tmp_path):
    ini = tmp_path / "c.ini"
    ini.write_text(
        "nx = 12   # trailing comment\n"
        "  ny=5 ; other comment\n"
        "nz = 8\n"
        "no equals sign here\n"
        "dx = 250abc\n"     # atof stops at the first bad char
        "dx2 = \n"
        "nz = 9\n"          # later duplicate wins
        "typo_key = 3\n"
        "proj_lat1 = 35\n"  # prep-only keys are never reported unused
        "terrain_source = tiles\n")
    c = Case(ini, dy="300")
    g = c.grid
    assert (g.nx, g.ny, g.nz) == (12, 5, 9)
    assert g.dx == 250.0 and g.dy == 300.0
    unused = c.unused_keys()
    assert "typo_key" in unused and "dx2" in unused
    assert not any(k.startswith("proj_") or k == "terrain_source" for k in unused)


def test_missing_config_is_an_error(tmp_path):
    with pytest.raises(WfeError):
        Case(tmp_path / "nope.ini")


def test_unknown_profile_is_an_error():
    with pytest.raises(WfeError, match="unknown profile"):
        case("warm_bubble", profile="sounding_xyz", nx=8, ny=8, nz=8)


def test_uniform_and_geometric_levels():
    c = case("warm_bubble", nx=8, ny=8, nz=10, dz=250)
    g = c.grid
    zw = c.array("zeta_w")
    k = np.arange(g.NZ) - g.ng
    np.testing.assert_allclose(zw, k * 250.0, rtol=0, atol=1e-9)
    assert c.scalar("zt") == 2500.0
    # geometric: zw[k] = dz0 (r^k - 1)/(r - 1) until capped at dz_max
    c2 = case("warm_bubble", nx=8, ny=8, nz=12, stretch="geometric", dz0=100, dz_ratio=1.2,
              dz_max=300)
    zw2 = c2.array("zeta_w")[3:3 + 13]
    d = [min(100 * 1.2 ** k, 300.0) for k in range(12)]
    np.testing.assert_allclose(zw2, np.concatenate([[0], np.cumsum(d)]), rtol=1e-14)
    dzc = c2.array("dzeta_c")[3:3 + 12]
    np.testing.assert_allclose(dzc, d, rtol=1e-13)
    zc = c2.array("zeta_c")
    np.testing.assert_allclose(c2.array("dzeta_w")[4:15], np.diff(zc)[3:14], rtol=1e-13)


def test_isentropic_exner_closed_form():
    # Isentropic: pi(z) = 1 - g z / (cp theta0) exactly (the 1 m trapezoid
    # integration is exact for constant theta; linear interpolation too).
    c = case("warm_bubble", nx=6, ny=4, nz=10, dz=500, theta0=300)
    g = c.grid
    zc = c.array("zeta_c")
    pib = g.as3d(c.array("pib"))[:, g.ng, g.ng]
    expect = 1.0 - G * zc / (CP * 300.0)
    np.testing.assert_allclose(pib[g.ng:g.ng + g.nz], expect[g.ng:g.ng + g.nz], rtol=0,
                               atol=1e-12)
    rho = g.as3d(c.array("rhob"))[:, g.ng, g.ng]
    rho_cf = P00 * expect ** (CV / RD) / (RD * 300.0)
    np.testing.assert_allclose(rho[g.ng:g.ng + g.nz], rho_cf[g.ng:g.ng + g.nz], rtol=1e-12)
    assert np.all(g.as3d(c.array("dthbdz")) == 0.0)


def test_constant_N_exner_closed_form():
    # theta = theta0 exp(N^2 z / g):
    # pi(z) = 1 + g^2/(cp theta0 N^2) (exp(-N^2 z/g) - 1) (hydrostatic, exact);
    # the 1 m trapezoid rule is second order, so agreement is ~1e-9.
    N, th0 = 0.01, 288.0
    c = case("schaer_rest", nx=6, ny=4, nz=20, terrain="none")
    g = c.grid
    zc = c.array("zeta_c")[g.ng:g.ng + g.nz]
    pib = g.as3d(c.array("pib"))[g.ng:g.ng + g.nz, g.ng, g.ng]
    cf = 1.0 + G ** 2 / (CP * th0 * N ** 2) * (np.exp(-N ** 2 * zc / G) - 1.0)
    np.testing.assert_allclose(pib, cf, rtol=0, atol=2e-9)
    thb = g.as3d(c.array("thb"))[g.ng:g.ng + g.nz, g.ng, g.ng]
    np.testing.assert_allclose(thb, th0 * np.exp(N ** 2 * zc / G), rtol=1e-14)
    dth = g.as3d(c.array("dthbdz"))[g.ng:g.ng + g.nz, g.ng, g.ng]
    np.testing.assert_allclose(dth, th0 * N ** 2 / G * np.exp(N ** 2 * zc / G), rtol=1e-6)


def test_gal_chen_jacobian_and_schaer_terrain():
    c = case("schaer_rest", nx=40, ny=4, nz=10)
    g = c.grid
    zt = c.scalar("zt")
    h = c.array("h").reshape(g.NY, g.NX)
    jac = c.array("jac").reshape(g.NY, g.NX)
    np.testing.assert_allclose(jac, (zt - h) / zt, rtol=1e-15)
    x = (np.arange(-3, g.nx + 3) + 0.5) * 500.0
    xr = x - 50000.0
    hs = 250.0 * np.exp(-(xr / 5000.0) ** 2) * np.cos(math.pi * xr / 4000.0) ** 2
    np.testing.assert_allclose(h[g.ng], hs, rtol=1e-13, atol=1e-12)
    hxu = c.array("hx_u").reshape(g.NY, g.NX)[g.ng]
    hprev = 250.0 * np.exp(-((xr - 500.0) / 5000.0) ** 2) * np.cos(
        math.pi * (xr - 500.0) / 4000.0) ** 2
    np.testing.assert_allclose(hxu, (hs - hprev) / 500.0, rtol=1e-10, atol=1e-12)


def test_wk82_sounding_properties():
    c = case("wk82_supercell", nx=6, ny=6, nz=40)
    g = c.grid
    sl = (slice(g.ng, g.ng + g.nz), g.ng, g.ng)
    zc = c.array("zeta_c")[g.ng:g.ng + g.nz]
    thb = g.as3d(c.array("thb"))[sl]
    qvb = g.as3d(c.array("qvb"))[sl]
    thv = g.as3d(c.array("thvb"))[sl]
    # theta: 300 + 43 (z/12km)^1.25 below the tropopause
    below = zc <= 12000
    np.testing.assert_allclose(thb[below], 300 + 43 * (zc[below] / 12000) ** 1.25, rtol=1e-14)
    assert qvb.max() <= 0.014 + 1e-15 and qvb.min() >= 0
    np.testing.assert_allclose(thv, thb * (1 + 0.61 * qvb), rtol=1e-14)
    # hydrostatic: d pi / dz = -g / (cp theta_v) (trapezoid, 1 m steps)
    pib = g.as3d(c.array("pib"))[sl]
    dpi = np.diff(pib) / np.diff(zc)
    mid = 0.5 * (thv[1:] + thv[:-1])
    np.testing.assert_allclose(dpi, -G / (CP * mid), rtol=3e-3)
    # initial qv = qvbar in the interior
    ic = g.interior(c.array("qv"))
    np.testing.assert_array_equal(ic[:, 0, 0], qvb)


def test_rh_cap_extension():
    a = case("wk82_supercell", nx=6, ny=6, nz=40)
    b = case("wk82_supercell", nx=6, ny=6, nz=40, rh_max=0.95)
    qa, qb = a.array("qvb"), b.array("qvb")
    assert np.all(qb <= qa + 1e-18) and qb.max() < qa.max() + 1e-18


def test_quarter_circle_hodograph():
    c = case("cfg2_supercell", nx=8, ny=8)
    zc = c.array("zeta_c")
    ub, vb = c.array("ub"), c.array("vb")
    z = np.maximum(zc, 0)
    a = 0.5 * math.pi * np.minimum(z / 5000.0, 1.0)
    np.testing.assert_allclose(ub, 30 * (1 - np.cos(a)) - 15, rtol=1e-14, atol=1e-13)
    np.testing.assert_allclose(vb, 30 * np.sin(a) - 10, rtol=1e-14, atol=1e-13)
    # speed relative to the offset is U_s on the whole circle: the hodograph
    # is a quarter circle of radius U_s centred at (U_s + offset, offset_v).
    np.testing.assert_allclose((ub + 15 - 30) ** 2 + (vb + 10) ** 2, 900.0, rtol=1e-12)


def test_bubble_ic_closed_form():
    c = case("cfg1_dry_bubble")
    g = c.grid
    th = g.interior(c.array("thp"))
    x = (np.arange(g.nx) + 0.5) * 250.0
    z = c.array("zeta_c")[g.ng:g.ng + g.nz]
    X, Z = np.meshgrid(x, z)
    r = np.sqrt(((X - 10000) / 2000) ** 2 + ((Z - 2000) / 2000) ** 2)
    ref = np.where(r < 1, 2 * np.cos(math.pi * r / 2) ** 2, 0.0)
    for j in range(g.ny):
        np.testing.assert_allclose(th[:, j, :], ref, rtol=1e-12, atol=1e-11)
    assert th.max() > 1.9


def test_reference_idealized_cases_load():
    for ini in sorted(CASES.glob("*.ini")):
        c = Case(ini, nx=8, ny=8, nz=10)
        assert c.grid.nx == 8
