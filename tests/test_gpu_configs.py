"""Every benchmark configuration under the GPU tests, at its real shape and,
where the oracle is fast enough, against the oracle built by its own case
assembly (oracle/case_builder.cpp) -- so the product's host builders, file
reader and boundary manager are inside the comparison too.

  cfg 1  dry warm bubble 81 x 4 x 41, FP64, all 600 steps vs the oracle;
  cfg 2  moist supercell 120 x 120 x 40, mixed, all 1200 steps: finite,
         |w| below the guard, moisture positive; FP64 vs mixed error;
  cfg 3  regional 330 x 220 x 45, mixed, all 1440 steps (24 h) with boundary
         files every 3 h: finite, |w| guard, PD positivity, and within 1e-4
         of the FP64 path on u, v, w, theta' after 24 h; a small regional
         case crossing two boundary files against the oracle;
  cfg 4  an uneven ensemble of seeded regional members at two grid sizes,
         advanced together, each member against its own oracle;
  cfg 5  adjoint at the 120 x 120 x 40 supercell shape, dry and with q_v
         (dot products in FP64 and about the FP32 trajectory, finite
         differences, FP32 production against the FP64 check), and reference
         gradients from finite differences of the FP64 oracle on a small grid
         (dry, moist, moist with Coriolis and diffusion).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import CASES, kitchen_sink, oracle_and_gpu, perturbed_state, pin_shapes, rel_err
from oracle_lib import OracleCase

pytestmark = pytest.mark.gpu

STATE = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")


def _errs(g, a, b, fields):
    out = {}
    for f in fields:
        nzl = g.nz + 1 if f == "w" else g.nz
        out[f] = rel_err(g.interior(a.get(f), nzl), g.interior(b.get(f), nzl))
    return out


def _guards(gi, g, w_abort, pd=True):
    d = gi.diagnose()
    print("diag", {k: (f"{v:.4g}" if isinstance(v, float) else v) for k, v in d.items()},
          flush=True)
    assert d["finite"] == 1
    assert d["wmax"] < w_abort
    if pd:
        qv = g.interior(gi.get("qv"), g.nz)
        assert qv.min() >= -1e-6, qv.min()
    return d


def test_cfg1_warm_bubble_all_600_steps_match_the_oracle():
    from wfe_amd import Case, Integrator
    ini = CASES / "cfg1_dry_bubble.ini"
    c, oc = Case(ini), OracleCase(ini)
    om = oc.model()
    gi = Integrator.from_case(c, "fp64")
    dt = c.scalar("dt")
    nsteps = int(round(c.scalar("t_end") / dt))
    assert nsteps == 600
    for s in range(nsteps):
        gi.step(dt, s * dt)
        om.step(dt, s * dt)
    e = _errs(c.grid, gi, om, ("u", "v", "w", "thp", "pip"))
    print("cfg1 600 steps", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(v <= 1e-10 for v in e.values()), e
    # the bubble has risen: w_max of the reference's gate range
    assert 5.0 < gi.diagnose()["wmax"] < 30.0


def test_cfg2_supercell_all_1200_steps_mixed_stay_bounded():
    from wfe_amd import Case, Integrator
    c = Case(CASES / "cfg2_supercell.ini")
    g = c.grid
    dt = c.scalar("dt")
    mix = Integrator.from_case(c, "mixed")
    ref = Integrator.from_case(c, "fp64")
    for s in range(1200):
        mix.step(dt, s * dt)
        ref.step(dt, s * dt)
        if s + 1 in (600, 1200):
            e = _errs(g, mix, ref, ("u", "v", "w", "thp"))
            print(f"cfg2 step {s + 1} mixed vs fp64", {k: f"{v:.2e}" for k, v in e.items()},
                  flush=True)
            d = _guards(mix, g, c.params.w_abort, pd=False)
            assert d["qrmax"] > 1e-3  # the storm rains
    qv = g.interior(mix.get("qv"), g.nz)
    assert qv.min() > -1e-4  # no PD limiter in this case: only small undershoots


def test_cfg3_regional_all_1440_steps_mixed_with_boundary_files(tmp_path):
    from wfe_amd import Case, Integrator
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", hours=24, bdy_interval_h=3)
    c = Case(ini)
    g = c.grid
    assert (g.nx, g.ny, g.nz) == (330, 220, 45)
    dt = c.scalar("dt")
    gi = Integrator.from_case(c, "mixed")
    lo0 = c.array("bdy_hi_thp").copy()
    for s in range(1440):
        gi.step(dt, s * dt)
        if (s + 1) % 360 == 0:
            print(f"cfg3 step {s + 1}", flush=True)
            _guards(gi, g, c.params.w_abort)
    # the product's boundary manager advanced through the files (the targets
    # it holds now differ from the ones it started with)
    assert not np.array_equal(c.array("bdy_hi_thp"), lo0)
    # 24 h against the FP64 path on the same case: the mixed path's error
    # bound of 1e-4 (relative L2 over the interior) on u, v, w and theta'
    gd = Integrator.from_case(c, "fp64")
    for s in range(1440):
        gd.step(dt, s * dt)
    e = _errs(g, gi, gd, ("u", "v", "w", "thp", "pip", "qv"))
    print("cfg3 mixed vs fp64 after 24 h", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(e[f] <= 1e-4 for f in ("u", "v", "w", "thp")), e


def test_regional_boundary_files_and_time_interpolation_match_the_oracle(tmp_path):
    """150 steps of 60 s cross the 1 h and 2 h boundary files: the product's
    file reader, target conversion, lo/hi swap and time fraction against the
    oracle's own restatement of BdyManager (boundary.cpp:9-97)."""
    from wfe_amd import Case, Integrator
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", nx=40, ny=30, nz=20, dx=12000.0, top=15000.0,
                         seed=5, hours=3, bdy_interval_h=1, bdy_shift_cells_per_h=3.0)
    c, oc = Case(ini), OracleCase(ini)
    g = c.grid
    dt = c.scalar("dt")
    om = oc.model()
    gd = Integrator.from_case(c, "fp64")
    gm = Integrator.from_case(c, "mixed")
    for s in range(150):
        for m in (om, gd, gm):
            m.step(dt, s * dt)
    e = _errs(g, gd, om, STATE)
    print("regional fp64 vs oracle", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(v <= 1e-10 for v in e.values()), e
    em = _errs(g, gm, om, ("u", "v", "w", "thp", "pip", "qv"))
    print("regional mixed vs oracle", {k: f"{v:.2e}" for k, v in em.items()}, flush=True)
    assert all(v <= 1e-3 for v in em.values()), em
    for f in ("u", "thp"):
        assert np.array_equal(c.array(f"bdy_hi_{f}"), oc.array(f"bdy_hi_{f}"))


def test_context_created_after_others_crossed_boundary_times_starts_from_the_first_files(
        tmp_path, monkeypatch):
    """Contexts of one case run one after another (as the benchmark runs its
    mixed and FP64 passes): a context created once another has crossed
    boundary times still starts from the first two boundary files, so it
    steps exactly like a context created before any stepping."""
    pin_shapes(monkeypatch)
    from wfe_amd import Case, Integrator
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", nx=32, ny=24, nz=16, dx=12000.0, top=12000.0,
                         seed=6, hours=3, bdy_interval_h=1, bdy_shift_cells_per_h=3.0)
    c = Case(ini)
    dt = c.scalar("dt")
    early = Integrator.from_case(c, "fp64")
    first = Integrator.from_case(c, "fp64")
    for s in range(130):  # past the 1 h and 2 h files
        first.step(dt, s * dt)
    late = Integrator.from_case(c, "fp64")
    for s in range(70):
        early.step(dt, s * dt)
        late.step(dt, s * dt)
    for f in STATE:
        assert np.array_equal(early.get(f), late.get(f)), f


@pytest.mark.parametrize("extra", [{"bdy_width": 7}, {"bdy_width": 4, "nudge_tau": 5400}],
                         ids=["rim", "nudged"])
def test_relaxation_rim_and_interior_nudging_match_the_oracle(tmp_path, extra):
    """Davies relaxation over the rim only (launched over the rim columns) and
    with interior nudging (weights everywhere, launched over the grid), FP64
    against the oracle over boundary-file crossings."""
    from wfe_amd import Case, Integrator
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", nx=36, ny=26, nz=16, dx=12000.0, top=12000.0,
                         seed=12, hours=2, bdy_interval_h=1, extra=extra)
    c, oc = Case(ini), OracleCase(ini)
    dt = c.scalar("dt")
    om = oc.model()
    gd = Integrator.from_case(c, "fp64")
    for s in range(70):
        for m in (om, gd):
            m.step(dt, s * dt)
    e = _errs(c.grid, gd, om, STATE)
    print(extra, {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(v <= 1e-10 for v in e.values()), e


def test_cfg4_uneven_ensemble_members_match_their_oracles(tmp_path):
    """Seeded regional members at two grid sizes and time steps (12 km, 60 s
    and 3 km, 15 s) with member theta' noise, advanced together by one graph
    launch per ensemble step; every member against its own oracle."""
    from wfe_amd import Case, Ensemble, Integrator
    from wfe_amd.lam_case import write_lam_case
    specs = [dict(nx=40, ny=30, nz=20, dx=12000.0, top=15000.0, dt=60.0)] * 2 + \
            [dict(nx=32, ny=32, nz=16, dx=3000.0, top=12000.0, dt=15.0, h_max=800.0)] * 2
    cases, oracles, members, dts = [], [], [], []
    for q, sp in enumerate(specs):
        ini = write_lam_case(tmp_path / f"m{q}", seed=20261202 + q, hours=0, low_noise_K=0.5,
                             **sp)
        c = Case(ini)
        cases.append(c)
        oracles.append(OracleCase(ini).model())
        members.append(Integrator.from_case(c, "fp64"))
        dts.append(c.scalar("dt"))
    ens = Ensemble(members)
    nsteps = 8
    for s in range(nsteps):
        ens.step(dts)
        for om, dt in zip(oracles, dts):
            om.step(dt, s * dt)
    ens.sync()
    for q, (c, gi, om) in enumerate(zip(cases, members, oracles)):
        e = _errs(c.grid, gi, om, STATE)
        print(f"member {q}", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
        assert all(v <= 1e-10 for v in e.values()), (q, e)
    # members differ from each other (own seeds and noise)
    a = cases[0].grid.interior(members[0].get("thp"))
    b = cases[1].grid.interior(members[1].get("thp"))
    assert rel_err(a, b) > 1e-3


def _box(g, rng, size_cells):
    from wfe_amd.adjoint import box_weights
    i0, j0 = (int(x) for x in rng.integers(0, min(g.nx, g.ny) - size_cells, size=2))
    return box_weights(g, i0, i0 + size_cells, j0, j0 + size_cells)


def test_cfg5_adjoint_at_the_supercell_shape():
    """Dot-product test and finite differences of the gradient at 120 x 120 x
    40 with the supercell's levels: 20 steps with a checkpoint every 5; then
    the FP32 production gradient against the FP64 one."""
    from wfe_amd import Case, Integrator
    from wfe_amd.adjoint import FIELDS5, ad_step, ke_gradient, tl_step
    c = Case(CASES / "cfg5_adjoint_dry.ini")
    g = c.grid
    assert (g.nx, g.ny, g.nz) == (120, 120, 40)
    dt = c.scalar("dt")
    gi = Integrator.from_case(c, "fp64")
    rng = np.random.default_rng(20261202)
    mask = np.zeros((g.NZ, g.NY, g.NX), dtype=bool)
    mask[g.ng:g.ng + g.nz, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = True
    mask = mask.ravel()
    scale = {"u": 1.0, "v": 1.0, "w": 0.3, "thp": 0.5, "pip": 1e-4}

    def rand():
        return {f: np.where(mask, rng.standard_normal(g.npts) * scale[f], 0.0) for f in FIELDS5}

    dx, yb = rand(), rand()
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs = sum(float(np.dot(dy[f][mask], yb[f][mask])) for f in FIELDS5)
    rhs = sum(float(np.dot(dx[f][mask], xb[f][mask])) for f in FIELDS5)
    print("cfg5 dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)

    w = _box(g, rng, 20)
    nsteps = 20
    J, grad = ke_gradient(gi, dt, nsteps, 5, w)
    x0 = {f: gi.get(f) for f in FIELDS5}
    d = rand()
    gd = sum(float(np.dot(grad[f][mask], d[f][mask])) for f in FIELDS5)

    def J_of(xs):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS5:
            gp.set(f, xs[f])
        tot, ww = 0.0, w.reshape(g.NY, g.NX)
        for _ in range(nsteps):
            gp.step(dt)
            u = gp.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = gp.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_of(x0), rel=1e-12)
    eps = 1e-6
    fd = (J_of({f: x0[f] + eps * d[f] for f in FIELDS5}) -
          J_of({f: x0[f] - eps * d[f] for f in FIELDS5})) / (2 * eps)
    print("cfg5 gradient vs FD", gd, fd, abs(fd - gd) / abs(gd), flush=True)
    assert abs(fd - gd) <= 1e-5 * abs(gd)
    # FP32 production (mixed context: FP32 forward and trajectory, FP64
    # adjoint) against this FP64 check from the same initial state
    g32 = Integrator.from_case(c, "mixed")
    for f in FIELDS5:
        g32.set(f, x0[f])
    # dot product about the FP32 trajectory: the stretched levels' dzeta is
    # not representable in FP32, which TL and AD must treat alike
    dy = tl_step(g32, dt, dx)
    xb = ad_step(g32, dt, yb)
    lhs = sum(float(np.dot(dy[f][mask], yb[f][mask])) for f in FIELDS5)
    rhs = sum(float(np.dot(dx[f][mask], xb[f][mask])) for f in FIELDS5)
    print("cfg5 FP32-trajectory dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    J32, grad32 = ke_gradient(g32, dt, nsteps, 5, w)
    errs = {f: float(np.linalg.norm(grad32[f][mask] - grad[f][mask]) /
                     np.linalg.norm(grad[f][mask])) for f in FIELDS5}
    print("cfg5 FP32 production vs FP64: J", abs(J32 - J) / abs(J), "gradient",
          {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    assert abs(J32 - J) <= 1e-5 * abs(J)
    assert max(errs.values()) <= 1e-3, errs


@pytest.mark.parametrize("moist,extra", [
    (False, {}), (True, {}), (True, {"coriolis_f": 1e-3, "diff_K": 15.0}),
    (True, {"coriolis_f": 1e-3, "diff_K": 15.0, "terrain": "gauss", "mtn_h0": 400,
            "mtn_a": 2500, "mtn_xc": 4800, "mtn_yc": 4000}),
    (True, {"coriolis_f": 1e-3, "diff_K": 15.0, "terrain": "gauss", "mtn_h0": 400,
            "mtn_a": 2500, "mtn_xc": 4800, "mtn_yc": 4000, "bc_x": "open"})],
    ids=["dry", "moist", "moist-coriolis-diffusion", "moist-terrain", "moist-terrain-open-x"])
def test_adjoint_gradient_against_finite_differences_of_the_oracle(moist, extra):
    """Reference gradient: the objective evaluated by the FP64 oracle (its own
    case assembly and step) at x0 +- eps d, against the hand-written adjoint's
    gradient along d, on a small periodic grid; dry, on the WK82 moist
    sounding without microphysics with q_v in the direction (and along q_v
    alone), the moist case with Coriolis and diffusion, the same over a
    Gaussian hill, and that with open x boundaries. (Open edges are taken
    where the edge velocity stays away from zero: at zero the radiation
    condition switches between inflow and outflow, a jump that centred
    differences over six steps would straddle.)"""
    from wfe_amd import Case, Integrator
    from wfe_amd.adjoint import FIELDS, FIELDS5, box_weights, ke_gradient
    over = dict(nx=24, ny=20, nz=16, dx=400, dy=400, dz=300, bubble_xc=4800, bubble_yc=4000,
                bubble_zc=1500, bubble_xr=1500, bubble_yr=1500, bubble_zr=1200, wind_u=6,
                rayleigh_zd=3000, rayleigh_alpha=0.02, dt=2.0, acoustic_ns=4)
    if moist:
        over.update(profile="wk82", microphysics="off")
    over.update(extra)
    fields = FIELDS if moist else FIELDS5
    ini = CASES / "warm_bubble.ini"
    c = Case(ini, **over)
    assert bool(c.params.moisture) == moist
    g = c.grid
    dt = c.scalar("dt")
    st = perturbed_state(c, 3)
    gi = Integrator.from_case(c, "fp64")
    for f in fields:
        gi.set(f, st[f])
    w = box_weights(g, 6, 14, 5, 13)
    nsteps = 6
    J, grad = ke_gradient(gi, dt, nsteps, 2, w)
    rng = np.random.default_rng(7)
    mask = np.zeros((g.NZ, g.NY, g.NX), dtype=bool)
    mask[g.ng:g.ng + g.nz, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = True
    mask = mask.ravel()
    scale = {"u": 1.0, "v": 1.0, "w": 0.3, "thp": 0.5, "pip": 1e-4, "qv": 1e-3}
    d = {f: np.where(mask, rng.standard_normal(g.npts) * scale[f], 0.0) for f in fields}

    def J_oracle(eps, dd):
        om = OracleCase(ini, **over).model()
        for f in fields:
            om.set(f, st[f] + eps * dd[f])
        tot, ww = 0.0, w.reshape(g.NY, g.NX)
        for s in range(nsteps):
            om.step(dt, s * dt)
            u = om.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = om.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_oracle(0.0, d), rel=1e-12)
    dirs = [("all", d)]
    if moist:
        dirs.append(("qv", {f: (d[f] if f == "qv" else np.zeros(g.npts)) for f in fields}))
    for name, dd in dirs:
        gd = sum(float(np.dot(grad[f][mask], dd[f][mask])) for f in fields)
        assert gd != 0.0
        for eps, tol in ((1e-4, 1e-4), (1e-6, 1e-6)):
            fd = (J_oracle(eps, dd) - J_oracle(-eps, dd)) / (2 * eps)
            print(f"oracle FD {name} eps={eps:g}: {fd!r} adjoint {gd!r} "
                  f"rel {abs(fd - gd) / abs(gd):.2e}", flush=True)
            assert abs(fd - gd) <= tol * abs(gd)


@pytest.mark.parametrize("pd", ["off", "on"], ids=["flux-form-q", "pd-q"])
def test_regional_adjoint_against_finite_differences_of_the_oracle(tmp_path, pd):
    """A small regional case in file mode (Lambert map factors, terrain, open
    boundaries, 5-point Davies relaxation toward one boundary file, w damping,
    moisture from the file), microphysics off, q_v by flux-form or
    positive-definite transport: the adjoint's gradient along a full
    direction and along q_v alone against centred differences of the FP64
    oracle's objective over 4 steps (2 with the PD limiter acting)."""
    from wfe_amd import Case, Integrator
    from wfe_amd.adjoint import FIELDS, box_weights, ke_gradient
    from wfe_amd.lam_case import write_lam_case
    ini = write_lam_case(tmp_path / "lam", nx=24, ny=18, nz=16, dx=12000.0, top=12000.0,
                         seed=7, hours=0,
                         extra={"pd_moist": pd, "microphysics": "off"})
    c = Case(ini)
    g = c.grid
    assert c.params.bc_x_open and c.params.bc_y_open and c.params.moisture
    assert np.abs(c.array("hx")).max() > 0 and np.ptp(c.array("mapf")) > 1e-3
    dt = c.scalar("dt")
    gi = Integrator.from_case(c, "fp64")
    if pd == "on":  # q_v near zero in the top levels, so the limiter acts there
        qv = gi.get("qv")
        g.as3d(qv)[g.ng + g.nz - 4:g.ng + g.nz] *= 1e-3
        gi.set("qv", qv)
    x0 = {f: gi.get(f) for f in FIELDS}
    w = box_weights(g, 6, 18, 5, 13)
    # with the limiter acting, the differences cross its R = 1 kinks more
    # often the longer the window: 2 steps there, and the smaller eps
    nsteps = 2 if pd == "on" else 4
    epss = (1e-6, 1e-7, 1e-8) if pd == "on" else (1e-5, 1e-6, 1e-7)
    J, grad = ke_gradient(gi, dt, nsteps, 2, w)
    mask = np.zeros((g.NZ, g.NY, g.NX), dtype=bool)
    mask[g.ng:g.ng + g.nz, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = True
    mask = mask.ravel()
    rng = np.random.default_rng(17)
    scale = {"u": 1.0, "v": 1.0, "w": 0.3, "thp": 0.5, "pip": 1e-4, "qv": 1e-3}
    d = {f: np.where(mask, rng.standard_normal(g.npts) * scale[f], 0.0) for f in FIELDS}

    def J_oracle(eps, dd):
        om = OracleCase(ini).model()
        for f in FIELDS:
            om.set(f, x0[f] + eps * dd[f])
        tot, ww = 0.0, w.reshape(g.NY, g.NX)
        for s in range(nsteps):
            om.step(dt, s * dt)
            u = om.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = om.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_oracle(0.0, d), rel=1e-12)
    dq = {f: (d[f] if f == "qv" else np.zeros(g.npts)) for f in FIELDS}
    for name, dd in (("all", d), ("qv", dq)):
        gd = sum(float(np.dot(grad[f][mask], dd[f][mask])) for f in FIELDS)
        errs = []
        for eps in epss:
            fd = (J_oracle(eps, dd) - J_oracle(-eps, dd)) / (2 * eps)
            errs.append(abs(fd - gd) / abs(gd))
        print(f"regional oracle FD {name} ({nsteps} steps): adjoint {gd!r}, rel err by eps",
              epss, [f"{e:.2e}" for e in errs], flush=True)
        assert gd != 0.0 and min(errs) < 1e-6, (name, errs)


def test_cfg5_moist_adjoint_at_the_supercell_shape():
    """The moist cfg 5 case (WK82 sounding, no microphysics) at 120 x 120 x
    40: six-field dot products in FP64 and about the FP32 trajectory; the
    gradient over 20 steps (checkpoint every 5) along a q_v-only direction
    against finite differences; the FP32 production gradient against it."""
    from wfe_amd import Case, Integrator
    from wfe_amd.adjoint import FIELDS, ad_step, ke_gradient, tl_step
    c = Case(CASES / "cfg5_adjoint_moist.ini")
    g = c.grid
    assert (g.nx, g.ny, g.nz) == (120, 120, 40) and c.params.moisture
    dt = c.scalar("dt")
    rng = np.random.default_rng(20261203)
    mask = np.zeros((g.NZ, g.NY, g.NX), dtype=bool)
    mask[g.ng:g.ng + g.nz, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = True
    mask = mask.ravel()
    scale = {"u": 1.0, "v": 1.0, "w": 0.3, "thp": 0.5, "pip": 1e-4, "qv": 1e-3}

    def rand():
        return {f: np.where(mask, rng.standard_normal(g.npts) * scale[f], 0.0) for f in FIELDS}

    dx, yb = rand(), rand()
    ctx = {p: Integrator.from_case(c, p) for p in ("fp64", "mixed")}
    for p, gi in ctx.items():
        dy = tl_step(gi, dt, dx)
        xb = ad_step(gi, dt, yb)
        lhs = sum(float(np.dot(dy[f][mask], yb[f][mask])) for f in FIELDS)
        rhs = sum(float(np.dot(dx[f][mask], xb[f][mask])) for f in FIELDS)
        print(f"cfg5 moist {p} dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs), flush=True)
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    w = _box(g, rng, 20)
    nsteps = 20
    gi = ctx["fp64"]
    x0 = {f: gi.get(f) for f in FIELDS}
    J, grad = ke_gradient(gi, dt, nsteps, 5, w)
    d = rand()
    dq = {f: (d[f] if f == "qv" else np.zeros(g.npts)) for f in FIELDS}
    gq = float(np.dot(grad["qv"][mask], d["qv"][mask]))

    def J_of(xs):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS:
            gp.set(f, xs[f])
        tot, ww = 0.0, w.reshape(g.NY, g.NX)
        for _ in range(nsteps):
            gp.step(dt)
            u = gp.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = gp.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_of(x0), rel=1e-12)
    eps = 1e-5
    fd = (J_of({f: x0[f] + eps * dq[f] for f in FIELDS}) -
          J_of({f: x0[f] - eps * dq[f] for f in FIELDS})) / (2 * eps)
    print("cfg5 moist q_v gradient vs FD", gq, fd, abs(fd - gq) / abs(gq), flush=True)
    assert gq != 0.0 and abs(fd - gq) <= 1e-5 * abs(gq)
    J32, grad32 = ke_gradient(ctx["mixed"], dt, nsteps, 5, w)
    errs = {f: float(np.linalg.norm(grad32[f][mask] - grad[f][mask]) /
                     np.linalg.norm(grad[f][mask])) for f in FIELDS}
    print("cfg5 moist FP32 production vs FP64: J", abs(J32 - J) / abs(J), "gradient",
          {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    assert abs(J32 - J) <= 1e-5 * abs(J)
    assert max(errs.values()) <= 1e-3, errs


SLOT_LIMITS = [("fused-64x14", 42), ("fused-64x10x4", 40), ("fused-64x16", 48),
               ("fused-64x8", 48), ("fused-32x16", 48), ("fused-32x8", 48),
               ("fused-32x16w4", 48), ("fused-32x16x4", 64), ("split-64x16", 48),
               ("split-32x16x4", 64), ("split-32x16w4", 48),
               # FP64 LDS of the 2-D tiles: 16 x 4 takes at most 47 levels, 8 x 8 all 48
               ("tile2d-16x4", 47), ("tile2d-8x8", 48)]


@pytest.mark.parametrize("shape,nz", SLOT_LIMITS, ids=[f"{s}-nz{n}" for s, n in SLOT_LIMITS])
def test_acoustic_shapes_at_their_exact_level_limit(monkeypatch, shape, nz):
    """Each substep tile at the deepest column it takes (every level slot
    holds a level), pinned through WFE_ACOUSTIC_KERNEL, FP64 vs the oracle."""
    monkeypatch.setenv("WFE_ACOUSTIC_KERNEL", shape)
    # whole-metre dz: the WK82 theta of a negative height is NaN (a power of
    # a negative base), so the base state's 1 m table must have a level at
    # z = 0 exactly, which 16 dz being a whole number of metres ensures
    c = kitchen_sink(nx=21, ny=13, nz=nz, dz=float(round(10800.0 / nz)), terrain="none")
    st = perturbed_state(c, 41)
    om, gi = oracle_and_gpu(c, "fp64", st)
    dt = c.scalar("dt")
    for s in range(2):
        om.step(dt)
        gi.step(dt)
    e = _errs(c.grid, gi, om, STATE)
    print(shape, nz, {k: f"{v:.1e}" for k, v in e.items()}, flush=True)
    assert all(v <= 1e-10 for v in e.values()), e
    assert f"tuned:{shape}" in gi.kernel_report()
