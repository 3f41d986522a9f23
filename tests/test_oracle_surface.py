"""Pins of the oracle's surface-layer / PBL column physics
(src/physics/surface.cu) that do not come from restating it (CPU only):
the Businger-Dyer stability functions against their defining integrals, and
the implicit column diffusion against a banded library solve, its
conservation of the column integral and its steady states."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate, linalg

from oracle_lib import _p, olib


def psi(heat, z):
    return olib().orc_sfc_psi(int(heat), float(z))


@pytest.mark.parametrize("heat,power", [(0, -0.25), (1, -0.5)])
@pytest.mark.parametrize("zeta", [-0.01, -0.3, -1.0, -4.0, -8.0])
def test_unstable_psi_is_the_integral_of_its_flux_profile(heat, power, zeta):
    """psi(zeta) = int_0^zeta (1 - phi(s)) / s ds with the Businger-Dyer
    phi_m = (1 - 16 s)^(-1/4), phi_h = (1 - 16 s)^(-1/2)."""
    val, _ = integrate.quad(lambda s: (1 - (1 - 16 * s) ** power) / s, zeta, 0.0,
                            epsabs=1e-13, epsrel=1e-12)
    assert psi(heat, zeta) == pytest.approx(-val, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("heat", [0, 1])
def test_stable_psi_is_linear_and_capped(heat):
    assert psi(heat, 0.0) == 0.0
    for z in (0.1, 0.7, 1.9):
        assert psi(heat, z) == pytest.approx(-5 * z, rel=1e-15)
    assert psi(heat, 3.0) == psi(heat, 2.0) == -10.0
    # continuous through neutral from the unstable side
    assert abs(psi(heat, -1e-9)) < 1e-7


def _column(rng, nz):
    rho = rng.uniform(0.3, 1.2, nz)
    rhow = rng.uniform(0.3, 1.2, nz)
    dzc = rng.uniform(50, 400, nz)
    dzw = rng.uniform(50, 400, nz)
    Kw = np.concatenate([[0.0], rng.uniform(0.1, 300, nz - 1)])
    return rho, rhow, dzc, dzw, Kw


def _diffuse(nz, dt, rho, rhow, dzc, dzw, Kw, src0, phi, fcg=None):
    out = np.ascontiguousarray(phi, dtype=np.float64).copy()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, rhow, dzc, dzw, Kw)]
    f = None if fcg is None else np.ascontiguousarray(fcg, dtype=np.float64)
    olib().orc_sfc_diffuse(nz, dt, *[_p(a) for a in arrs], float(src0), _p(out),
                           _p(f) if f is not None else None)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_diffusion_solves_its_tridiagonal_system(seed):
    """The backward-Euler flux form: (rho dzc / dt)(phi' - phi) = F_{k+1} - F_k
    + src0 [k = 0] + fcg_k - fcg_{k+1}, F_k = rhow_k K_k (phi'_k - phi'_{k-1})
    / dzw_k, solved by LAPACK's banded solver."""
    rng = np.random.default_rng(seed)
    nz, dt = 30, 60.0
    rho, rhow, dzc, dzw, Kw = _column(rng, nz)
    phi = rng.standard_normal(nz)
    fcg = np.concatenate([[0.0], rng.standard_normal(nz - 1) * 1e-2, [0.0]])
    src0 = 0.37
    got = _diffuse(nz, dt, rho, rhow, dzc, dzw, Kw, src0, phi, fcg)
    kw = np.zeros(nz + 1)
    kw[1:nz] = rhow[1:] * Kw[1:] / dzw[1:]
    m = rho * dzc / dt
    ab = np.zeros((3, nz))
    ab[1] = m + kw[:nz] + kw[1:]
    ab[0, 1:] = -kw[1:nz]
    ab[2, :-1] = -kw[1:nz]
    rhs = m * phi + fcg[:nz] - fcg[1:]
    rhs[0] += src0
    want = linalg.solve_banded((1, 1), ab, rhs)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)


def test_diffusion_conserves_the_column_and_keeps_uniform_columns():
    rng = np.random.default_rng(5)
    nz, dt = 40, 30.0
    rho, rhow, dzc, dzw, Kw = _column(rng, nz)
    phi = rng.uniform(280, 300, nz)
    src0 = 2.5
    out = _diffuse(nz, dt, rho, rhow, dzc, dzw, Kw, src0, phi)
    before, after = np.sum(rho * dzc * phi), np.sum(rho * dzc * out)
    assert after - before == pytest.approx(dt * src0, rel=1e-9)
    flat = np.full(nz, 287.5)
    assert np.allclose(_diffuse(nz, dt, rho, rhow, dzc, dzw, Kw, 0.0, flat), 287.5,
                       rtol=0, atol=1e-10)


# ---------------------------------------------------------------- radiation
SIGMA = 5.670e-8  # Stefan-Boltzmann constant [W m-2 K-4]
S0 = 1361.0       # solar constant [W m-2]


def cosz(doy, utc, lat, lon):
    return olib().orc_sfc_cosz(int(doy), float(utc), float(lat), float(lon))


@pytest.mark.parametrize("doy,decl_deg", [(80, 0.0), (265, 0.0), (172, 23.44), (355, -23.44)])
@pytest.mark.parametrize("lat,lon", [(40.0, 30.0), (-15.0, -60.0), (60.0, 0.0)])
def test_noon_zenith_follows_the_seasons(doy, decl_deg, lat, lon):
    """At local solar noon (UTC 12 - lon/15) the zenith angle is |lat - decl|
    with the declination 0 at the equinoxes (21 March, 22 September) and
    +-23.44 deg at the solstices, within the degree or so that a
    one-harmonic declination formula is off by in September."""
    z = np.degrees(np.arccos(cosz(doy, 12.0 - lon / 15.0, lat, lon)))
    assert z == pytest.approx(abs(lat - decl_deg), abs=1.2)


@pytest.mark.parametrize("dh", [-3.0, -1.5, 2.0, 4.0])
def test_hour_angle_turns_fifteen_degrees_an_hour(dh):
    """Equinox (declination ~ 0): cos z = cos(lat) cos(15 deg x hours from
    local noon), by spherical trigonometry."""
    lat, lon = 35.0, 45.0
    want = np.cos(np.radians(lat)) * np.cos(np.radians(15.0 * dh))
    assert cosz(80, 12.0 - lon / 15.0 + dh, lat, lon) == pytest.approx(want, abs=0.01)


def test_no_sun_at_night():
    assert cosz(172, 0.0 - 30.0 / 15.0, 40.0, 30.0) == 0.0  # local midnight


def lw(T, eps, Ts):
    nz = len(T)
    Fup, Fdn = np.zeros(nz + 1), np.zeros(nz + 1)
    T, eps = np.ascontiguousarray(T, float), np.ascontiguousarray(eps, float)
    olib().orc_sfc_lw(nz, _p(T), _p(eps), float(Ts), _p(Fup), _p(Fdn))
    return Fup, Fdn


def test_long_wave_transparent_and_black_columns():
    """Transparent layers pass the surface emission (emissivity 0.97) up
    unchanged and send nothing down; black layers each emit sigma T^4, the
    surface reflecting 3% of what reaches it (Kirchhoff)."""
    rng = np.random.default_rng(3)
    T = rng.uniform(200, 300, 12)
    Ts = 295.0
    Fup, Fdn = lw(T, np.zeros(12), Ts)
    assert np.all(Fdn == 0)
    assert np.allclose(Fup, 0.97 * SIGMA * Ts ** 4, rtol=1e-12)
    Fup, Fdn = lw(T, np.ones(12), Ts)
    assert np.allclose(Fdn[:-1], SIGMA * T ** 4, rtol=1e-12) and Fdn[-1] == 0
    assert np.allclose(Fup[1:], SIGMA * T ** 4, rtol=1e-12)
    assert Fup[0] == pytest.approx(0.97 * SIGMA * Ts ** 4 + 0.03 * SIGMA * T[0] ** 4, rel=1e-12)


def test_long_wave_isothermal_grey_column():
    """An isothermal column at T of layer emissivities e_k sends down
    sigma T^4 (1 - prod(1 - e_k)) to the surface."""
    e = np.array([0.1, 0.3, 0.05, 0.5, 0.2])
    T = np.full(5, 250.0)
    _, Fdn = lw(T, e, 280.0)
    assert Fdn[0] == pytest.approx(SIGMA * 250.0 ** 4 * (1 - np.prod(1 - e)), rel=1e-12)


def sw(cz, lwp, du, dl):
    nz = len(du)
    ab = np.zeros(nz)
    du, dl = np.ascontiguousarray(du, float), np.ascontiguousarray(dl, float)
    return olib().orc_sfc_sw(nz, float(cz), float(lwp), _p(du), _p(dl), _p(ab)), ab


def test_short_wave_conserves_the_beam_and_splits_paths():
    """What the layers absorb plus what reaches the ground is the beam at the
    top; Beer's law makes the surface flux depend only on the total path, so
    halving every layer into two gives the same surface flux; no sun, no
    flux."""
    rng = np.random.default_rng(8)
    du, dl = rng.uniform(0, 8, 10), rng.uniform(0, 0.3, 10)
    for cz, lwp in [(0.8, 0.0), (0.3, 25.0), (0.02, 5.0)]:
        sfc, ab = sw(cz, lwp, du, dl)
        top, _ = sw(cz, lwp, np.zeros(10), np.zeros(10))
        assert sfc + ab.sum() == pytest.approx(top, rel=1e-12)
        sfc2, _ = sw(cz, lwp, np.repeat(du / 2, 2), np.repeat(dl / 2, 2))
        assert sfc2 == pytest.approx(sfc, rel=1e-12)
    # slant path: the transmittance under a sun at cos z equals that of an
    # overhead sun through a path 1 / cos z times as long
    for cz in (0.9, 0.5, 0.2):
        sfc, _ = sw(cz, 0.0, du, dl)
        top, _ = sw(cz, 0.0, np.zeros(10), np.zeros(10))
        sfc1, _ = sw(1.0, 0.0, du / cz, dl / cz)
        top1, _ = sw(1.0, 0.0, np.zeros(10), np.zeros(10))
        assert sfc / top == pytest.approx(sfc1 / top1, rel=1e-12)
    top, ab = sw(0.9, 0.0, np.zeros(3), np.zeros(3))
    assert top == pytest.approx(S0 * 0.9 * 0.95, rel=1e-12) and np.all(ab == 0)
    assert sw(0.0, 0.0, du, dl)[0] == 0.0


# ---------------------------------------------------------------- surface layer
KAPPA = 0.4


def louis(rib, z1, z0):
    out = np.zeros(2)
    olib().orc_sfc_louis(float(rib), float(z1), float(z0), _p(out))
    return out


@pytest.mark.parametrize("z1,z0", [(25.0, 0.1), (10.0, 2e-4), (60.0, 0.1)])
def test_louis_neutral_limit_and_stability_ordering(z1, z0):
    """Neutral: Cd = Ch = (kappa / ln(z1/z0))^2; stable air damps exchange
    and unstable air enhances it, heat more than momentum."""
    a2 = (KAPPA / np.log(z1 / z0)) ** 2
    assert np.allclose(louis(0.0, z1, z0), a2, rtol=1e-14)
    prev = louis(0.0, z1, z0)
    for rib in (0.05, 0.2, 1.0):
        cd, ch = louis(rib, z1, z0)
        assert cd == ch and cd < prev[0]
        prev = (cd, ch)
    for rib in (-0.05, -0.5, -3.0):
        cd, ch = louis(rib, z1, z0)
        assert a2 < cd < ch


def diag(V1, rib, z1, z0, ths, th1):
    out = np.zeros(2)
    olib().orc_sfc_diag(float(V1), float(rib), float(z1), float(z0), float(ths), float(th1),
                        _p(out))
    return out


@pytest.mark.parametrize("z1,z0", [(25.0, 0.1), (40.0, 2e-4)])
def test_neutral_diagnostics_follow_the_log_law(z1, z0):
    """Neutral surface layer: wind and potential temperature vary with
    ln(z / z0) (z0h = z0 / 10 for heat), so the 10 m wind is V1 ln(10/z0) /
    ln(z1/z0) and the 2 m theta is ths + (th1 - ths) ln(2/z0h) / ln(z1/z0h)."""
    u10, th2 = diag(6.0, 0.0, z1, z0, 290.0, 288.0)
    assert u10 == pytest.approx(6.0 * np.log(10 / z0) / np.log(z1 / z0), rel=1e-12)
    z0h = z0 / 10
    assert th2 == pytest.approx(290.0 + (288.0 - 290.0) * np.log(2 / z0h) / np.log(z1 / z0h),
                                rel=1e-12)


# ---------------------------------------------------------------- soil
DZ_SOIL = np.array([0.1, 0.3, 0.6, 1.0])


def soil(dt, w, G, tdeep, T):
    t = np.ascontiguousarray(T, float).copy()
    olib().orc_sfc_soil(float(dt), float(w), float(G), float(tdeep), _p(t))
    return t


@pytest.mark.parametrize("w", [0.05, 0.2, 0.5])
def test_soil_step_conserves_heat_and_keeps_equilibrium(w):
    """The implicit step changes the layers' heat content by exactly the
    surface flux plus the conduction from the deep layer through the last
    half layer, both over dt; a soil at the deep temperature with no surface
    flux stays there."""
    C = olib().orc_sfc_soil_heat(w, 0)
    K = olib().orc_sfc_soil_heat(w, 1)
    T0 = np.array([290.0, 287.0, 285.0, 284.0])
    dt, G, tdeep = 600.0, 120.0, 282.0
    T1 = soil(dt, w, G, tdeep, T0)
    dheat = np.sum(C * DZ_SOIL * (T1 - T0))
    assert dheat == pytest.approx(dt * (G + K / DZ_SOIL[3] * (tdeep - T1[3])), rel=1e-10)
    assert np.allclose(soil(dt, w, 0.0, 283.5, np.full(4, 283.5)), 283.5, rtol=0, atol=1e-10)


def test_soil_step_solves_its_tridiagonal_system():
    """The same step as a banded system solved by LAPACK: layer interfaces
    conduct with K over the mean of the two layer depths."""
    w, dt, G, tdeep = 0.25, 900.0, -40.0, 281.0
    C = olib().orc_sfc_soil_heat(w, 0)
    K = olib().orc_sfc_soil_heat(w, 1)
    T0 = np.array([275.0, 278.0, 280.0, 281.5])
    kint = K / (0.5 * (DZ_SOIL[:-1] + DZ_SOIL[1:]))
    cap = C * DZ_SOIL / dt
    ab = np.zeros((3, 4))
    ab[1] = cap + np.concatenate([[0], kint]) + np.concatenate([kint, [K / DZ_SOIL[3]]])
    ab[0, 1:] = -kint
    ab[2, :-1] = -kint
    rhs = cap * T0
    rhs[0] += G
    rhs[3] += K / DZ_SOIL[3] * tdeep
    assert np.allclose(soil(dt, w, G, tdeep, T0), linalg.solve_banded((1, 1), ab, rhs),
                       rtol=1e-13)
