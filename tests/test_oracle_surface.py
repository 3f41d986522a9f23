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
