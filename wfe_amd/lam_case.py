"""Seeded synthetic regional case in the reference's file-mode formats.

Writes, into one directory, what the reference's GFS preprocessor writes for
a real-data run (tools/prep_gfs.py:321-366, read by src/io/input.cpp:19-73):

  wfe_input.ini            version 3, nx ny nz, np_prof, start, bdy_interval,
                           n_bdy, hours
  wfe_init.bin             float32: profile tables z, theta, qv, u (np_prof
                           each, 100 m spacing), then h, fcor, tsk, land, lat,
                           lon, soilw (nx*ny), then u, v, theta, pi, qv
                           (nx*ny*nz, k slowest, i fastest; cell centres)
  wfe_bdy_%03d.bin         float32 u, v, theta, pi, qv (nx*ny*nz each), one
                           file per boundary time (file hour = idx*interval/3600)

plus a case .ini that runs it (profile = file, terrain = file, Lambert map
factors from the lat field and proj_lat1/2 as src/main.cpp:323-339 does).

The state is synthetic and seeded:
  terrain   k^-2 spectral random field scaled to 0..h_max, smoothed until the
            steepest slope is <= max_slope;
  lat field the cell centres of a Lambert conformal grid (centre 37 N, standard
            parallels 31 N and 67 N by default), on which the reference's
            Lambert formula gives map factors 0.950..1.050 over the
            330 x 220 x 12 km domain; f = 2 Omega sin(lat);
  theta, qv US standard atmosphere, relative humidity 40..90 % from a seeded
            smooth field, Gaussian-filtered theta perturbation (sigma_theta)
            and u, v perturbations (sigma_u) on a westerly jet profile;
  pi        hydrostatic per column from the perturbed virtual temperature;
  members   optional white theta noise N(0, low_noise_K) below low_noise_top
            above ground (ensemble members);
  surface   land where the terrain exceeds 2 % of h_max (sea at 290 K
            elsewhere), skin temperature from a 6.5 K/km lapse, soil moisture
            0.12..0.35 west to east (used with physics = simple);
  boundary  the initial state at hour 0; at later boundary times the same
            seeded state with its perturbation fields carried east by
            bdy_shift_cells_per_h grid cells per hour (pi recomputed
            hydrostatically), so the relaxation target moves in time.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

GRAV, CP, RD, P00 = 9.81, 1004.5, 287.04, 1.0e5
OMEGA = 7.292e-5
KAPPA = RD / CP


def geometric_levels(nz: int, dz0: float, ratio: float, dz_max: float) -> np.ndarray:
    """w-level heights as the geometric stretch builds them (metric.cpp:19-56)."""
    zw = np.zeros(nz + 1)
    d = dz0
    for k in range(1, nz + 1):
        zw[k] = zw[k - 1] + d
        d = min(d * ratio, dz_max)
    return zw


def ratio_for_top(nz: int, dz0: float, dz_max: float, top: float) -> float:
    lo, hi = 1.0, 1.5
    if geometric_levels(nz, dz0, hi, dz_max)[-1] < top:
        raise ValueError(f"{nz} levels from dz0={dz0} capped at {dz_max} cannot reach {top} m")
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if geometric_levels(nz, dz0, mid, dz_max)[-1] < top:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def us_standard(z: np.ndarray):
    """Temperature [K] and pressure [Pa] of the 1976 US standard atmosphere
    below 32 km (lapse -6.5 K/km to 11 km, isothermal to 20 km, +1 K/km)."""
    z = np.asarray(z, dtype=np.float64)
    T = np.where(z < 11000, 288.15 - 0.0065 * z,
                 np.where(z < 20000, 216.65, 216.65 + 0.001 * (z - 20000)))
    g_r = GRAV / RD
    p11 = 101325.0 * (216.65 / 288.15) ** (g_r / 0.0065)
    p20 = p11 * math.exp(-g_r * 9000 / 216.65)
    p = np.where(z < 11000, 101325.0 * (T / 288.15) ** (g_r / 0.0065),
                 np.where(z < 20000, p11 * np.exp(-g_r * (z - 11000) / 216.65),
                          p20 * (T / 216.65) ** (-g_r / 0.001)))
    return T, p


def qsat(p, T):
    """Saturation mixing ratio over water (Tetens, src/core/thermo.hpp:11-26)."""
    es = 610.78 * np.exp(17.269 * (T - 273.16) / (T - 35.86))
    return 0.622 * es / np.maximum(p - es, 1.0)


def _smooth_field(rng, shape, sigma, axes_sigma=None):
    from scipy.ndimage import gaussian_filter
    a = rng.standard_normal(shape)
    a = gaussian_filter(a, axes_sigma if axes_sigma is not None else sigma, mode="wrap")
    a -= a.mean()
    return a / max(float(a.std()), 1e-30)


def spectral_terrain(rng, nx, ny, dx, h_max, max_slope):
    """k^-2 amplitude spectrum, scaled to [0, h_max], 1-2-1 smoothed until
    the steepest slope is at most max_slope."""
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    amp = k ** -2.0
    phase = np.exp(2j * np.pi * rng.random((ny, nx)))
    h = np.real(np.fft.ifft2(amp * phase))
    h = (h - h.min()) / (h.max() - h.min()) * h_max
    for _ in range(200):
        gy, gx = np.gradient(h, dx)
        if float(np.sqrt(gx * gx + gy * gy).max()) <= max_slope:
            break
        p = np.pad(h, 1, mode="edge")
        h = (4 * p[1:-1, 1:-1] + p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]) / 8
    return h


def write_lam_case(out_dir, *, nx=330, ny=220, nz=45, dx=12000.0, seed=20261202,
                   top=25000.0, dz0=250.0, dz_max=1000.0, h_max=3000.0, max_slope=0.3,
                   sigma_theta=1.0, sigma_u=3.0, lat0=37.0, proj_lat1=31.0, proj_lat2=67.0,
                   dt=60.0, acoustic_ns=6, hours=24, bdy_interval_h=24, bdy_width=5,
                   bdy_tau=600.0, low_noise_K=0.0, low_noise_top=2000.0,
                   bdy_shift_cells_per_h=2.0,
                   extra: dict | None = None) -> Path:
    """Write the input files and a case .ini into out_dir; returns the .ini."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    ratio = ratio_for_top(nz, dz0, dz_max, top)
    zw = geometric_levels(nz, dz0, ratio, dz_max)
    zt = zw[-1]
    zc = 0.5 * (zw[:-1] + zw[1:])

    # horizontal grid: a Lambert conformal grid centred at (lat0, 10E) with
    # the case's standard parallels, so the map factors the driver forms
    # from lat (src/main.cpp:323-339) are those of the projection itself
    lat, lon = lambert_grid_latlon(nx, ny, dx, lat0, 10.0, proj_lat1, proj_lat2)
    fcor = 2 * OMEGA * np.sin(np.radians(lat))
    h = spectral_terrain(rng, nx, ny, dx, h_max, max_slope)
    # land where the terrain rises above 2 % of its maximum, sea elsewhere
    # (a sea surface at 290 K); a soil-moisture gradient from west to east
    land = (h > 0.02 * h_max).astype(np.float64)
    tsk = np.where(land > 0.5, 288.15 - 0.0065 * h, 290.0)
    soilw = np.broadcast_to(np.linspace(0.12, 0.35, nx)[None, :], (ny, nx)).copy()

    # physical heights of the cell centres (Gal-Chen)
    z3 = h[None] + zc[:, None, None] * (zt - h[None]) / zt

    # base-state profile tables on a 100 m grid: the area-mean state
    npf = int(zt // 100) + 2
    ztab = np.arange(npf) * 100.0
    Tt, pt = us_standard(ztab)
    th_tab = Tt * (P00 / pt) ** KAPPA

    def rh_mean(z):
        return 0.65 - 0.25 * np.clip(z / 12000.0, 0.0, 1.0)

    qv_tab = np.where(ztab > 15000, 0.0, np.minimum(rh_mean(ztab) * qsat(pt, Tt), 0.018))
    u_tab = 5.0 + 25.0 * np.minimum(ztab / 10000.0, 1.0)

    # Exner function of that mean state, hydrostatic from pi(0) = 1 with the
    # virtual potential temperature (as the base-state builder integrates it)
    zf = np.arange(0.0, zt + 1000.0, 1.0)
    thv_f = np.interp(zf, ztab, th_tab * (1 + 0.61 * qv_tab))
    inv = 1.0 / thv_f
    pib_f = 1.0 - GRAV / CP * np.concatenate([[0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]))])

    # 3-D state: the mean state at the cell heights plus seeded smooth noise
    T3, p3 = us_standard(z3)
    sig = (2.0, 6.0, 6.0)  # grid cells (k, j, i)
    n_th = _smooth_field(rng, z3.shape, None, sig)
    n_rh = _smooth_field(rng, z3.shape, None, sig)
    n_u = _smooth_field(rng, z3.shape, None, sig)
    n_v = _smooth_field(rng, z3.shape, None, sig)
    th_mean = np.interp(z3, ztab, th_tab)
    u_mean = np.interp(z3, ztab, u_tab)

    def state(shift: int, member_noise=None):
        """(u, v, theta, pi, qv) with the noise fields rolled `shift` cells
        east (they are periodic, so the roll stays smooth)."""
        roll = (lambda a: np.roll(a, shift, axis=2)) if shift else (lambda a: a)
        th3 = th_mean + sigma_theta * roll(n_th)
        rh = np.clip(rh_mean(z3) + 0.12 * roll(n_rh), 0.4, 0.9)
        qv3 = np.where(z3 > 15000, 0.0, np.minimum(rh * qsat(p3, T3), 0.018))
        u3 = u_mean + sigma_u * roll(n_u)
        v3 = sigma_u * roll(n_v)
        if member_noise is not None:
            th3 = th3 + member_noise
        # pi per column: hydrostatic with the column's own theta_v, from the
        # mean state's value at the terrain height
        thv3 = th3 * (1 + 0.61 * qv3)
        pi3 = np.empty_like(th3)
        pi3[0] = np.interp(h, zf, pib_f) - GRAV * (z3[0] - h) / (CP * thv3[0])
        for k in range(1, nz):
            thm = 0.5 * (thv3[k] + thv3[k - 1])
            pi3[k] = pi3[k - 1] - GRAV * (z3[k] - z3[k - 1]) / (CP * thm)
        return u3, v3, th3, pi3, qv3

    # ensemble member perturbation: white theta noise N(0, low_noise_K) in
    # the lowest low_noise_top metres (initial state only)
    member = None
    if low_noise_K > 0:
        member = np.where(z3 - h[None] < low_noise_top,
                          rng.normal(0.0, low_noise_K, z3.shape), 0.0)
    u3, v3, th3, pi3, qv3 = state(0, member)

    f32 = np.float32
    with open(out / "wfe_init.bin", "wb") as f:
        for a in (ztab, th_tab, qv_tab, u_tab):
            a.astype(f32).tofile(f)
        for a in (h, fcor, tsk, land, lat, lon, soilw):
            a.astype(f32).tofile(f)
        for a in (u3, v3, th3, pi3, qv3):
            a.astype(f32).tofile(f)
    # boundary files: the initial state at hour 0, then the same state with
    # its perturbations carried east by bdy_shift_cells_per_h cells per hour
    n_bdy = hours // bdy_interval_h + 1
    for b in range(n_bdy):
        hr = b * bdy_interval_h
        fields = (u3, v3, th3, pi3, qv3) if b == 0 else \
            state(int(round(hr * bdy_shift_cells_per_h)))
        np.concatenate([a.ravel() for a in fields]).astype(f32).tofile(
            out / f"wfe_bdy_{hr:03d}.bin")
    meta = [f"version = 3", f"nx = {nx}", f"ny = {ny}", f"nz = {nz}", f"np_prof = {npf}",
            f"start = 2026120200", f"bdy_interval = {bdy_interval_h * 3600}",
            f"n_bdy = {n_bdy}", f"hours = {hours}"]
    (out / "wfe_input.ini").write_text("\n".join(meta) + "\n")

    ini = {
        "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dy": dx, "dz": dz0,
        "stretch": "geometric", "dz0": dz0, "dz_ratio": f"{ratio:.8f}", "dz_max": dz_max,
        "proj_lat0": lat0, "proj_lon0": 10.0, "proj_lat1": proj_lat1, "proj_lat2": proj_lat2,
        "dt": dt, "acoustic_ns": acoustic_ns, "t_end": hours * 3600,
        "profile": "file", "terrain": "file", "input_dir": str(out),
        "bc_x": "open", "bc_y": "open", "bdy_width": bdy_width, "bdy_tau": bdy_tau,
        "nudge_tau": 0, "w_damping": "on", "pd_moist": "on",
        "rayleigh_zd": 18000, "rayleigh_alpha": 0.02, "precision": "mixed",
        "out_dir": str(out / "run"), "out_interval": 21600, "diag_interval": 3600,
    }
    if extra:
        ini.update(extra)
    p = out / "case.ini"
    p.write_text("# seeded synthetic regional case (wfe_amd.lam_case)\n" +
                 "".join(f"{k} = {v}\n" for k, v in ini.items()))
    return p


def _cone_constant(phi1, phi2):
    q = math.pi / 4
    if abs(phi1 - phi2) < 1e-6:
        return math.sin(phi1)
    return math.log(math.cos(phi1) / math.cos(phi2)) / \
        math.log(math.tan(q + phi2 / 2) / math.tan(q + phi1 / 2))


def lambert_grid_latlon(nx, ny, dx, lat0, lon0, lat1, lat2, radius=6.371e6):
    """Latitude and longitude [deg] of the cell centres of an nx x ny grid of
    spacing dx on a spherical Lambert conformal conic projection (standard
    parallels lat1, lat2) centred at (lat0, lon0): the inverse projection of
    Snyder (1987, eqs. 14-4, 15-1 to 15-5). Shape [ny, nx]."""
    q = math.pi / 4
    phi1, phi2, phi0 = (math.radians(a) for a in (lat1, lat2, lat0))
    n = _cone_constant(phi1, phi2)
    F = math.cos(phi1) * math.tan(q + phi1 / 2) ** n / n
    rho0 = radius * F / math.tan(q + phi0 / 2) ** n
    x = (np.arange(nx) + 0.5 - nx / 2) * dx
    y = (np.arange(ny) + 0.5 - ny / 2) * dx
    X, Y = np.meshgrid(x, y)
    rho = math.copysign(1.0, n) * np.sqrt(X ** 2 + (rho0 - Y) ** 2)
    lat = 2 * np.arctan((radius * F / rho) ** (1 / n)) - math.pi / 2
    lon = math.radians(lon0) + np.arctan2(X, rho0 - Y) / n
    return np.degrees(lat), np.degrees(lon)


def lambert_map_factor(lat_deg, lat1, lat2):
    """Isotropic Lambert conformal map factor (src/main.cpp:326-339)."""
    phi1, phi2 = math.radians(lat1), math.radians(lat2)
    q = math.pi / 4
    n = _cone_constant(phi1, phi2)
    A = math.cos(phi1) * math.tan(q + phi1 / 2) ** n
    phi = np.radians(lat_deg)
    return A / (np.tan(q + phi / 2) ** n * np.cos(phi))


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="write a seeded synthetic regional case")
    ap.add_argument("out_dir")
    ap.add_argument("--nx", type=int, default=330)
    ap.add_argument("--ny", type=int, default=220)
    ap.add_argument("--nz", type=int, default=45)
    ap.add_argument("--dx", type=float, default=12000.0)
    ap.add_argument("--seed", type=int, default=20261202)
    a = ap.parse_args()
    print(write_lam_case(a.out_dir, nx=a.nx, ny=a.ny, nz=a.nz, dx=a.dx, seed=a.seed))
