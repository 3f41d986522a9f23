"""The seeded synthetic regional case written in the reference's file-mode
formats (src/io/input.cpp:19-73, tools/prep_gfs.py:321-366)."""
import numpy as np
import pytest

from wfe_amd import Case
from wfe_amd.lam_case import (lambert_grid_latlon, lambert_map_factor, us_standard,
                               write_lam_case)

SMALL = dict(nx=24, ny=18, nz=16, dx=12000.0, top=12000.0, seed=7)


@pytest.fixture(scope="module")
def small_case(tmp_path_factory):
    d = tmp_path_factory.mktemp("lam")
    return write_lam_case(d, **SMALL), d


def test_files_have_the_reference_layout(small_case):
    ini, d = small_case
    meta = dict(l.split(" = ") for l in (d / "wfe_input.ini").read_text().split("\n") if l)
    nx, ny, nz, npf = (int(meta[k]) for k in ("nx", "ny", "nz", "np_prof"))
    assert int(meta["version"]) == 3 and (nx, ny, nz) == (24, 18, 16)
    n2, n3 = nx * ny, nx * ny * nz
    init = np.fromfile(d / "wfe_init.bin", dtype=np.float32)
    assert init.size == 4 * npf + 7 * n2 + 5 * n3
    ztab = init[:npf]
    np.testing.assert_allclose(np.diff(ztab), 100.0)
    n_bdy = int(meta["n_bdy"])
    step_h = int(meta["bdy_interval"]) // 3600
    for b in range(n_bdy):
        bdy = np.fromfile(d / f"wfe_bdy_{b * step_h:03d}.bin", dtype=np.float32)
        assert bdy.size == 5 * n3
        # the boundary target is the initial state (u v theta pi qv)
        np.testing.assert_array_equal(bdy, init[4 * npf + 7 * n2:])


def test_case_loads_with_consistent_base_state(small_case):
    ini, d = small_case
    c = Case(ini)
    g = c.grid
    assert c.unused_keys() == ["precision"] or c.unused_keys() == []
    th = g.interior(c.array("thp"))
    pp = g.interior(c.array("pip"))
    # perturbations, not offsets: the base state is the generator's mean state
    assert abs(th.mean()) < 0.5 and th.std() < 2.0
    assert np.abs(pp).max() < 5e-3
    h = c.array("h").reshape(g.NY, g.NX)[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
    assert h.min() >= 0 and h.max() <= 3000.0 + 1e-6
    assert c.params.bc_x_open and c.params.pd_moist


def test_full_domain_map_factors_span_the_configured_band():
    """330 x 220 at 12 km on the generator's default Lambert grid: the map
    factors the driver forms from the lat field cover 0.95..1.05."""
    lat, lon = lambert_grid_latlon(330, 220, 12000.0, 37.0, 10.0, 31.0, 67.0)
    m = lambert_map_factor(lat, 31.0, 67.0)
    assert 0.950 <= m.min() < 0.951 and 1.049 < m.max() <= 1.0501
    # exactly 1 on the standard parallels, below 1 between them
    np.testing.assert_allclose(lambert_map_factor(np.array([31.0, 67.0]), 31.0, 67.0), 1.0,
                               rtol=1e-12)
    assert lambert_map_factor(np.array([49.0]), 31.0, 67.0)[0] < 1.0


def test_lambert_grid_is_conformal_at_the_spacing():
    """The inverse projection places neighbouring cell centres dx / m apart on
    the sphere (m: the map factor), which pins the grid against the formula."""
    R, dx = 6.371e6, 12000.0
    lat, lon = lambert_grid_latlon(40, 30, dx, 37.0, 10.0, 31.0, 67.0)
    la, lo = np.radians(lat), np.radians(lon)

    def dist(a1, o1, a2, o2):  # haversine great-circle distance
        h = np.sin((a2 - a1) / 2) ** 2 + np.cos(a1) * np.cos(a2) * np.sin((o2 - o1) / 2) ** 2
        return 2 * R * np.arcsin(np.sqrt(h))

    dxs = dist(la[:, :-1], lo[:, :-1], la[:, 1:], lo[:, 1:])
    dys = dist(la[:-1], lo[:-1], la[1:], lo[1:])
    mx = lambert_map_factor(np.degrees(0.5 * (la[:, :-1] + la[:, 1:])), 31.0, 67.0)
    my = lambert_map_factor(np.degrees(0.5 * (la[:-1] + la[1:])), 31.0, 67.0)
    np.testing.assert_allclose(dxs * mx, dx, rtol=2e-4)
    np.testing.assert_allclose(dys * my, dx, rtol=2e-4)


def test_us_standard_atmosphere_known_values():
    T, p = us_standard(np.array([0.0, 11000.0, 20000.0]))
    np.testing.assert_allclose(T, [288.15, 216.65, 216.65])
    np.testing.assert_allclose(p, [101325.0, 22632.0, 5474.9], rtol=2e-3)
