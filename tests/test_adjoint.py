"""Tangent-linear and adjoint of the dry large step: dot-product tests,
tangent-linear against finite differences of the forward step, and the
kinetic-energy gradient against finite differences of the objective (FP64
contexts); dot products and the gradient about the FP32 trajectory of a mixed
context against the FP64 check."""
import numpy as np
import pytest

from helpers import case, perturbed_state, pin_shapes

pytestmark = pytest.mark.gpu

FIELDS5 = ("u", "v", "w", "thp", "pip")


def dry_case(**over):
    kw = dict(nx=24, ny=20, nz=16, dx=400, dy=400, dz=300, bubble_xc=4800, bubble_yc=4000,
              bubble_zc=1500, bubble_xr=1500, bubble_yr=1500, bubble_zr=1200, wind_u=6,
              rayleigh_zd=3000, rayleigh_alpha=0.02, dt=2.0, acoustic_ns=4)
    kw.update(over)
    return case("warm_bubble", **kw)


def interior_mask(g):
    m = np.zeros((g.NZ, g.NY, g.NX), dtype=bool)
    m[g.ng:g.ng + g.nz, g.ng:g.ng + g.ny, g.ng:g.ng + g.nx] = True
    return m.ravel()


def random_state(g, rng, scale):
    mask = interior_mask(g)
    return {f: np.where(mask, rng.standard_normal(g.npts) * scale[f], 0.0) for f in FIELDS5}


SCALE = {"u": 1.0, "v": 1.0, "w": 0.3, "thp": 0.5, "pip": 1e-4}


def make(c, seed, precision="fp64"):
    from wfe_amd import Integrator
    gi = Integrator.from_case(c, precision)
    for f, a in perturbed_state(c, seed).items():
        gi.set(f, a)
    return gi


def dot(a, b, mask):
    return sum(float(np.dot(a[f][mask], b[f][mask])) for f in FIELDS5)


@pytest.mark.parametrize("seed,precision", [(1, "fp64"), (2, "fp64"), (1, "mixed")])
def test_dot_product_one_step(seed, precision):
    """<M dx, yb> = <dx, M^T yb>. In a mixed context the trajectory is FP32 and
    the TL/AD arithmetic FP64, so the identity holds to FP64 rounding too."""
    from wfe_amd.adjoint import ad_step, tl_step
    c = dry_case()
    g = c.grid
    gi = make(c, seed, precision)
    rng = np.random.default_rng(100 + seed)
    mask = interior_mask(g)
    dx = random_state(g, rng, SCALE)
    yb = random_state(g, rng, SCALE)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs, rhs = dot(dy, yb, mask), dot(dx, xb, mask)
    print("dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tangent_linear_after_an_adjoint_call_is_unchanged(precision):
    """The TL and AD passes share their perturbation buffers; cells the TL
    reads but does not write (rigid-lid w levels, substep ghost rings) must
    not carry what the adjoint pass left there."""
    from wfe_amd.adjoint import ad_step, tl_step
    c = dry_case()
    g = c.grid
    gi = make(c, 12, precision)
    rng = np.random.default_rng(112)
    dx, yb = random_state(g, rng, SCALE), random_state(g, rng, SCALE)
    dt = c.scalar("dt")
    before = tl_step(gi, dt, dx)
    ad_step(gi, dt, yb)
    after = tl_step(gi, dt, dx)
    for f in FIELDS5:
        assert np.array_equal(before[f], after[f]), f


def test_tangent_linear_matches_finite_differences():
    from wfe_amd import Integrator
    from wfe_amd.adjoint import tl_step
    c = dry_case()
    g = c.grid
    gi = make(c, 3)
    x = {f: gi.get(f) for f in FIELDS5}
    rng = np.random.default_rng(7)
    dx = random_state(g, rng, SCALE)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    mask = interior_mask(g)
    eps = 1e-6
    outs = []
    for sgn in (1, -1):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS5:
            gp.set(f, x[f] + sgn * eps * dx[f])
        for f in ("qv", "qc", "qr"):
            gp.set(f, gi.get(f))
        gp.step(dt)
        outs.append({f: gp.get(f) for f in FIELDS5})
    for f in FIELDS5:
        fd = (outs[0][f] - outs[1][f])[mask] / (2 * eps)
        tl = dy[f][mask]
        err = np.linalg.norm(fd - tl) / np.linalg.norm(tl)
        print(f, err)
        assert err < 1e-5, (f, err)


def test_kinetic_energy_gradient_matches_finite_differences():
    from wfe_amd import Integrator
    from wfe_amd.adjoint import box_weights, ke_gradient
    c = dry_case()
    g = c.grid
    gi = make(c, 4)
    x = {f: gi.get(f) for f in FIELDS5}
    dt = c.scalar("dt")
    w = box_weights(g, 6, 14, 5, 13)
    nsteps, ckpt = 6, 4
    J, grad = ke_gradient(gi, dt, nsteps, ckpt, w)
    # the context is left at the initial state
    for f in FIELDS5:
        np.testing.assert_array_equal(gi.get(f), x[f])
    rng = np.random.default_rng(9)
    d = random_state(g, rng, SCALE)
    mask = interior_mask(g)
    gd = sum(float(np.dot(grad[f][mask], d[f][mask])) for f in FIELDS5)

    def J_of(xs):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS5:
            gp.set(f, xs[f])
        tot = 0.0
        for s in range(nsteps):
            gp.step(dt)
            u = gp.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = gp.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            ww = w.reshape(g.NY, g.NX)
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_of(x), rel=1e-12)
    errs = []
    for eps in (1e-3, 1e-4, 1e-5, 1e-6, 1e-7):
        jp = J_of({f: x[f] + eps * d[f] for f in FIELDS5})
        jm = J_of({f: x[f] - eps * d[f] for f in FIELDS5})
        fd = (jp - jm) / (2 * eps)
        errs.append(abs(fd - gd) / abs(gd))
    print("gradient vs FD", gd, errs)
    assert min(errs) < 1e-6


def test_torch_autograd_function():
    import torch
    from wfe_amd.adjoint import box_weights, kinetic_energy_function
    c = dry_case()
    g = c.grid
    gi = make(c, 5)
    KE = kinetic_energy_function()
    ts = [torch.tensor(gi.get(f), dtype=torch.float64, requires_grad=True) for f in FIELDS5]
    J = KE.apply(gi, c.scalar("dt"), 3, 2, box_weights(g, 4, 20, 4, 16), *ts)
    (2.0 * J).backward()
    from wfe_amd.adjoint import ke_gradient
    gi2 = make(c, 5)
    J2, grad = ke_gradient(gi2, c.scalar("dt"), 3, 2, box_weights(g, 4, 20, 4, 16))
    assert float(J.detach()) == pytest.approx(J2, rel=1e-13)
    # two adjoint runs agree to the summation order of their FP64 atomics
    for t, f in zip(ts, FIELDS5):
        scale = np.abs(grad[f]).max() + 1e-300
        np.testing.assert_allclose(t.grad.numpy(), 2.0 * grad[f], rtol=1e-9, atol=1e-12 * scale)


def test_mixed_trajectory_tangent_linear_tracks_the_fp64_one():
    """The TL about a mixed context's FP32 trajectory against the TL about the
    FP64 trajectory from the same initial state: they differ by the
    trajectory's FP32 storage error only."""
    from wfe_amd.adjoint import tl_step
    c = dry_case()
    g = c.grid
    g64, g32 = make(c, 3), make(c, 3, "mixed")
    rng = np.random.default_rng(17)
    dx = random_state(g, rng, SCALE)
    dt = c.scalar("dt")
    d64, d32 = tl_step(g64, dt, dx), tl_step(g32, dt, dx)
    mask = interior_mask(g)
    errs = {f: float(np.linalg.norm(d32[f][mask] - d64[f][mask]) /
                     np.linalg.norm(d64[f][mask])) for f in FIELDS5}
    print("TL mixed vs fp64", {f: f"{e:.2e}" for f, e in errs.items()})
    assert max(errs.values()) < 1e-4, errs


def test_mixed_trajectory_gradient_tracks_the_fp64_gradient():
    """FP32 production gradient (mixed context: FP32 forward and trajectory,
    FP64 adjoint) against the FP64 check over 6 steps, checkpoint every 4."""
    from wfe_amd.adjoint import box_weights, ke_gradient
    c = dry_case()
    g = c.grid
    dt = c.scalar("dt")
    w = box_weights(g, 6, 14, 5, 13)
    g64, g32 = make(c, 4), make(c, 4, "mixed")
    x32 = {f: g32.get(f) for f in FIELDS5}
    J64, gr64 = ke_gradient(g64, dt, 6, 4, w)
    J32, gr32 = ke_gradient(g32, dt, 6, 4, w)
    for f in FIELDS5:  # the mixed context is left at its initial state too
        np.testing.assert_array_equal(g32.get(f), x32[f])
    mask = interior_mask(g)
    errs = {f: float(np.linalg.norm(gr32[f][mask] - gr64[f][mask]) /
                     np.linalg.norm(gr64[f][mask])) for f in FIELDS5}
    print("J", J64, J32, abs(J32 - J64) / J64, "gradient mixed vs fp64",
          {f: f"{e:.2e}" for f, e in errs.items()})
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    assert max(errs.values()) < 1e-4, errs


def test_scope_is_enforced():
    from wfe_amd import Integrator, WfeError
    from wfe_amd.adjoint import tl_step
    for precision in ("fp64", "mixed"):
        c = dry_case(acoustic_ns=17)
        gi = Integrator.from_case(c, precision)
        with pytest.raises(WfeError, match="more than 16 acoustic substeps"):
            tl_step(gi, c.scalar("dt"), {})


def _relax(gi, c, seed, rim=4, w0=0.02, interior=0.0):
    """Davies weights (w0 at the edge, falling linearly over rim cells, plus
    a uniform interior nudging weight) and seeded targets for u, v, theta',
    pi' and q_v, time fraction 0.3."""
    import ctypes as C
    from wfe_amd._lib import lib
    g = c.grid
    w = np.zeros((g.NY, g.NX))
    for j in range(g.ny):
        for i in range(g.nx):
            d = min(i, j, g.nx - 1 - i, g.ny - 1 - j)
            w[g.ng + j, g.ng + i] = (w0 * (rim - d) / rim if d < rim else 0.0) + interior
    w = w.ravel()
    DP = C.POINTER(C.c_double)
    L = lib()
    assert L.wfe_set_bdy_weights(gi._h, w.ctypes.data_as(DP), w.size) == 0
    rng = np.random.default_rng(seed)
    for slot in (0, 1):
        for var, f in enumerate(("u", "v", "thp", "pip", "qv")):
            t = gi.get(f) + rng.standard_normal(g.npts) * SCALE6[f]
            assert L.wfe_set_bdy_target(gi._h, slot, var, t.ctypes.data_as(DP), t.size) == 0
    assert L.wfe_set_bdy_tfrac(gi._h, 0.3) == 0
    return w


@pytest.mark.parametrize("moist,interior", [(False, 0.0), (True, 0.0), (True, 1e-3)],
                         ids=["dry-rim", "moist-rim", "moist-nudged"])
def test_relaxation_enters_the_linearisation(moist, interior):
    """Relaxation switched on after the adjoint was set up (wfe_set_bdy_weights)
    is part of the next linearised step: rim weights over open boundaries,
    and with interior nudging everywhere."""
    from wfe_amd import Integrator
    from wfe_amd.adjoint import tl_step
    kw = dict(TERRAIN, bc_x="open", bc_y="open")
    c = moist_case(**kw) if moist else dry_case(**kw)
    gi = make_moist(c, 61) if moist else make(c, 61)
    flds = FIELDS6 if moist else FIELDS5
    dt = c.scalar("dt")
    rng = np.random.default_rng(161)
    dx = random6(c.grid, rng)
    before = tl_step(gi, dt, dx)
    _relax(gi, c, 7, interior=interior)
    after = tl_step(gi, dt, dx)
    assert any(not np.array_equal(before[f], after[f]) for f in flds)

    def fresh():
        gp = Integrator.from_case(c, "fp64")
        # same state-independent setup: weights, targets from the same seed
        for f in flds:
            gp.set(f, gi.get(f))
        _relax(gp, c, 7, interior=interior)
        return gp

    _dot_and_fd(c, gi, flds, fresh, 162)


def test_adjoint_entries_leave_state_set_and_ensembles_intact(monkeypatch):
    """tl_step / ad_step / the gradient run forward steps internally; the
    context must come back in the same state set with the same values, so an
    ensemble built with a context that was never differentiated still steps
    in lock-step with it, and both match a fresh context stepped alone."""
    pin_shapes(monkeypatch)
    from wfe_amd import Ensemble, Integrator
    from wfe_amd.adjoint import ad_step, box_weights, ke_gradient, tl_step
    c = dry_case()
    g = c.grid
    dt = c.scalar("dt")
    ga, gb, ref = make(c, 8), make(c, 8), make(c, 8)
    x = {f: ga.get(f) for f in FIELDS5}
    tl_step(ga, dt, {})
    ad_step(ga, dt, {})
    ke_gradient(ga, dt, 3, 2, box_weights(g, 4, 20, 4, 16))
    for f in FIELDS5:
        np.testing.assert_array_equal(ga.get(f), x[f])
    ens = Ensemble([ga, gb])  # raises if the members are out of step
    for _ in range(2):
        ens.step([dt, dt])
        ref.step(dt)
    ens.sync()
    for f in FIELDS5:
        np.testing.assert_array_equal(ga.get(f), ref.get(f))
        np.testing.assert_array_equal(gb.get(f), ref.get(f))
    del ens
    assert isinstance(ga, Integrator)


# ---------------------------------------------------------------- moist (q_v)
FIELDS6 = FIELDS5 + ("qv",)
SCALE6 = dict(SCALE, qv=1e-3)


def moist_case(**over):
    """The small periodic case on the WK82 moist sounding, microphysics off:
    q_v is transported and acts through buoyancy and theta_v."""
    kw = dict(profile="wk82", microphysics="off")
    kw.update(over)
    return dry_case(**kw)


def make_moist(c, seed, precision="fp64"):
    from wfe_amd import Integrator
    gi = Integrator.from_case(c, precision)
    for f, a in perturbed_state(c, seed).items():
        if f not in ("qc", "qr"):  # q_c = q_r = 0: the linearised scope
            gi.set(f, a)
    return gi


def random6(g, rng):
    mask = interior_mask(g)
    return {f: np.where(mask, rng.standard_normal(g.npts) * SCALE6[f], 0.0) for f in FIELDS6}


def dot6(a, b, mask):
    return sum(float(np.dot(a[f][mask], b[f][mask])) for f in FIELDS6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_moist_dot_product_one_step(precision):
    from wfe_amd.adjoint import ad_step, tl_step
    c = moist_case()
    assert c.params.moisture and not c.params.microphysics
    g = c.grid
    gi = make_moist(c, 21, precision)
    rng = np.random.default_rng(121)
    mask = interior_mask(g)
    dx, yb = random6(g, rng), random6(g, rng)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs, rhs = dot6(dy, yb, mask), dot6(dx, xb, mask)
    # the q_v block on its own: q_v in, q_v out
    qq = tl_step(gi, dt, {"qv": dx["qv"]})["qv"]
    qb = ad_step(gi, dt, {"qv": yb["qv"]})["qv"]
    lq, rq = float(np.dot(qq[mask], yb["qv"][mask])), float(np.dot(dx["qv"][mask], qb[mask]))
    print("moist dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs), "qv block", lq, rq)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    assert abs(lq - rq) <= 1e-10 * abs(lq)
    assert np.abs(dy["qv"][mask]).max() > 0 and np.abs(xb["qv"][mask]).max() > 0


def test_moist_tangent_linear_matches_finite_differences():
    from wfe_amd import Integrator
    from wfe_amd.adjoint import tl_step
    c = moist_case()
    g = c.grid
    gi = make_moist(c, 22)
    x = {f: gi.get(f) for f in FIELDS6}
    rng = np.random.default_rng(122)
    dx = random6(g, rng)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    mask = interior_mask(g)
    eps = 1e-6
    outs = []
    for sgn in (1, -1):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS6:
            gp.set(f, x[f] + sgn * eps * dx[f])
        gp.step(dt)
        outs.append({f: gp.get(f) for f in FIELDS6})
    for f in FIELDS6:
        fd = (outs[0][f] - outs[1][f])[mask] / (2 * eps)
        err = np.linalg.norm(fd - dy[f][mask]) / np.linalg.norm(dy[f][mask])
        print(f, err)
        assert err < 1e-5, (f, err)


def test_moist_gradient_matches_finite_differences():
    """dJ/d(u, v, w, theta', pi', q_v) along a full direction and along a
    q_v-only one, against centred differences of the forward objective."""
    from wfe_amd import Integrator
    from wfe_amd.adjoint import box_weights, ke_gradient
    c = moist_case()
    g = c.grid
    gi = make_moist(c, 23)
    x = {f: gi.get(f) for f in FIELDS6}
    dt = c.scalar("dt")
    w = box_weights(g, 6, 14, 5, 13)
    nsteps, ckpt = 6, 4
    J, grad = ke_gradient(gi, dt, nsteps, ckpt, w)
    for f in FIELDS6:
        np.testing.assert_array_equal(gi.get(f), x[f])
    mask = interior_mask(g)
    rng = np.random.default_rng(123)
    d_all = random6(g, rng)
    d_q = {f: (d_all[f] if f == "qv" else np.zeros(g.npts)) for f in FIELDS6}

    def J_of(xs):
        gp = Integrator.from_case(c, "fp64")
        for f in FIELDS6:
            gp.set(f, xs[f])
        tot, ww = 0.0, w.reshape(g.NY, g.NX)
        for _ in range(nsteps):
            gp.step(dt)
            u = gp.get("u").reshape(g.NZ, g.NY, g.NX)[g.ng]
            v = gp.get("v").reshape(g.NZ, g.NY, g.NX)[g.ng]
            tot += float((ww * 0.5 * (u * u + v * v)).sum()) / nsteps
        return tot

    assert J == pytest.approx(J_of(x), rel=1e-12)
    for name, d in (("all", d_all), ("qv", d_q)):
        gd = sum(float(np.dot(grad[f][mask], d[f][mask])) for f in FIELDS6)
        errs = []
        for eps in (1e-3, 1e-4, 1e-5, 1e-6):
            fd = (J_of({f: x[f] + eps * d[f] for f in FIELDS6}) -
                  J_of({f: x[f] - eps * d[f] for f in FIELDS6})) / (2 * eps)
            errs.append(abs(fd - gd) / abs(gd))
        print("moist gradient vs FD", name, gd, errs)
        assert gd != 0.0
        assert min(errs) < 1e-6, (name, errs)


def test_moist_scope_is_enforced():
    from wfe_amd import Integrator, WfeError
    from wfe_amd.adjoint import tl_step
    for over, msg in (({"microphysics": "on"}, "microphysics"),
                      ({"acoustic_ns": 17}, "more than 16 acoustic substeps")):
        c = moist_case(**over)
        gi = Integrator.from_case(c, "fp64")
        with pytest.raises(WfeError, match=msg):
            tl_step(gi, c.scalar("dt"), {})
    c = moist_case()
    gi = make_moist(c, 24)
    tl_step(gi, c.scalar("dt"), {})
    qc = gi.get("qc")
    qc[len(qc) // 2] = 1e-4
    gi.set("qc", qc)
    with pytest.raises(WfeError, match="q_c or q_r"):
        tl_step(gi, c.scalar("dt"), {})


def test_dry_contexts_give_a_zero_qv_output():
    from wfe_amd.adjoint import ad_step, box_weights, ke_gradient, tl_step
    c = dry_case()
    g = c.grid
    gi = make(c, 25)
    rng = np.random.default_rng(125)
    d = random6(g, rng)
    assert not np.any(tl_step(gi, c.scalar("dt"), d)["qv"])
    assert not np.any(ad_step(gi, c.scalar("dt"), d)["qv"])
    assert not np.any(ke_gradient(gi, c.scalar("dt"), 2, 1, box_weights(g, 4, 20, 4, 16))[1]["qv"])


# ------------------------------------------------- Coriolis and diffusion
@pytest.mark.parametrize("moist", [False, True], ids=["dry", "moist"])
def test_coriolis_and_diffusion_dot_product_and_finite_differences(moist):
    """Constant f (coriolis_f) and constant-K diffusion of u, v, w, theta':
    dot product, and the tangent linear against centred differences of the
    forward step."""
    from wfe_amd import Integrator
    from wfe_amd.adjoint import ad_step, tl_step
    c = (moist_case if moist else dry_case)(coriolis_f=1e-3, diff_K=15.0)
    assert c.param_dict()["coriolis_f"] == 1e-3 and c.param_dict()["diff_K"] == 15.0
    g = c.grid
    gi = make_moist(c, 31) if moist else make(c, 31)
    flds = FIELDS6 if moist else FIELDS5
    rng = np.random.default_rng(131)
    mask = interior_mask(g)
    dx, yb = random6(g, rng), random6(g, rng)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs = sum(float(np.dot(dy[f][mask], yb[f][mask])) for f in flds)
    rhs = sum(float(np.dot(dx[f][mask], xb[f][mask])) for f in flds)
    print("Coriolis + diffusion dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    x = {f: gi.get(f) for f in flds}
    eps = 1e-5
    outs = []
    for sgn in (1, -1):
        gp = Integrator.from_case(c, "fp64")
        for f in flds:
            gp.set(f, x[f] + sgn * eps * (dx[f] if f in dx else 0.0))
        gp.step(dt)
        outs.append({f: gp.get(f) for f in flds})
    for f in flds:
        fd = (outs[0][f] - outs[1][f])[mask] / (2 * eps)
        err = np.linalg.norm(fd - dy[f][mask]) / np.linalg.norm(dy[f][mask])
        print(f, err)
        assert err < 1e-6, (f, err)


# ------------------------------------------------------ terrain, map factors
TERRAIN = dict(terrain="gauss", mtn_h0=400, mtn_a=2500, mtn_xc=4800, mtn_yc=4000)


def _dot_and_fd(c, gi, flds, fresh, seed, tol_fd=1e-7):
    """Dot product of the TL and AD around gi's state, and the TL against
    centred differences of the forward step (fresh() makes a new context of
    the same model)."""
    from wfe_amd.adjoint import ad_step, tl_step
    g = c.grid
    rng = np.random.default_rng(seed)
    mask = interior_mask(g)
    dx, yb = random6(g, rng), random6(g, rng)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs = sum(float(np.dot(dy[f][mask], yb[f][mask])) for f in flds)
    rhs = sum(float(np.dot(dx[f][mask], xb[f][mask])) for f in flds)
    print("dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    x = {f: gi.get(f) for f in flds}
    # Centred differences cross the kinks of the upwind selection (where a
    # trajectory flux is within eps of zero) at the larger eps and lose
    # digits to round-off at the smaller: the best of three is kept per field.
    errs = {f: [] for f in flds}
    for eps in (1e-5, 1e-6, 1e-7):
        outs = []
        for sgn in (1, -1):
            gp = fresh()
            for f in flds:
                gp.set(f, x[f] + sgn * eps * dx[f])
            gp.step(dt)
            outs.append({f: gp.get(f) for f in flds})
        for f in flds:
            fd = (outs[0][f] - outs[1][f])[mask] / (2 * eps)
            errs[f].append(np.linalg.norm(fd - dy[f][mask]) / np.linalg.norm(dy[f][mask]))
    for f in flds:
        print(f, [f"{e:.1e}" for e in errs[f]])
        assert min(errs[f]) < tol_fd, (f, errs[f])


@pytest.mark.parametrize("moist", [False, True], ids=["dry", "moist"])
def test_terrain_dot_product_and_finite_differences(moist):
    """A Gaussian hill (400 m, 2.5 km half-width) under periodic boundaries,
    with Coriolis and diffusion: Jacobian, cross-metric mass flux, the
    terrain pi* gradient, S_v and the surface w in the linearised step."""
    from wfe_amd import Integrator
    kw = dict(TERRAIN, coriolis_f=1e-3, diff_K=15.0)
    c = moist_case(**kw) if moist else dry_case(**kw)
    hx = c.array("hx")
    assert np.abs(hx).max() > 0.05  # slopes of several percent
    gi = make_moist(c, 41) if moist else make(c, 41)
    flds = FIELDS6 if moist else FIELDS5
    _dot_and_fd(c, gi, flds, lambda: Integrator.from_case(c, "fp64"), 141)


def test_map_factors_dot_product_and_finite_differences():
    """Non-uniform map factors (0.97-1.03) on the periodic hill: the map
    factor scales the horizontal divergence D_h of the column's pi' equation
    and the pressure gradient of the horizontal step, and nothing else."""
    from wfe_amd import Integrator
    from wfe_amd._lib import BASE_NAMES, METRIC_NAMES
    c = dry_case(**TERRAIN)
    g = c.grid
    arrays = {}
    for n in BASE_NAMES + METRIC_NAMES:
        try:
            arrays[n] = c.array(n)
        except KeyError:
            pass
    jj, ii = np.meshgrid(np.arange(g.NY), np.arange(g.NX), indexing="ij")
    arrays["mapf"] = (1.0 + 0.03 * np.sin(2 * np.pi * ii / g.nx) *
                      np.cos(2 * np.pi * jj / g.ny)).ravel()

    def fresh():
        return Integrator.from_arrays(g, c.params, arrays, c.scalar("zt"), "fp64")

    gi = fresh()
    for f, a in perturbed_state(c, 42).items():
        if f in FIELDS5:
            gi.set(f, a)
    _dot_and_fd(c, gi, FIELDS5, fresh, 142)


@pytest.mark.parametrize("moist", [False, True], ids=["dry", "moist"])
def test_open_boundaries_dot_product_and_finite_differences(moist):
    """Open x and y (radiation on the edge faces: base wind on inflow,
    Klemp-Wilhelmson on outflow, the branch frozen at the trajectory's edge
    velocity; zero-gradient ghosts) over the hill, with Coriolis and
    diffusion."""
    from wfe_amd import Integrator
    kw = dict(TERRAIN, coriolis_f=1e-3, diff_K=15.0, bc_x="open", bc_y="open")
    c = moist_case(**kw) if moist else dry_case(**kw)
    assert c.params.bc_x_open and c.params.bc_y_open
    gi = make_moist(c, 51) if moist else make(c, 51)
    flds = FIELDS6 if moist else FIELDS5
    _dot_and_fd(c, gi, flds, lambda: Integrator.from_case(c, "fp64"), 151)


@pytest.mark.parametrize("setup", ["periodic", "open-relaxed"])
def test_pd_transport_dot_product_and_finite_differences(setup):
    """Positive-definite q_v transport (the limiter's branches frozen at the
    trajectory): six-field dot product and the TL against centred
    differences, periodic, and with open boundaries, terrain and
    relaxation."""
    from wfe_amd import Integrator
    kw = dict(pd_moist="on")
    if setup != "periodic":
        kw.update(TERRAIN, bc_x="open", bc_y="open")
    c = moist_case(**kw)
    assert c.params.pd_moist
    gi = make_moist(c, 71)
    # q_v near zero somewhere, so the limiter acts (R < 1) in some cells
    qv = gi.get("qv")
    g = c.grid
    q3 = g.as3d(qv)
    q3[g.ng + g.nz - 3:g.ng + g.nz, :, :] *= 1e-3
    gi.set("qv", qv)
    if setup != "periodic":  # targets from the state the forward runs start from
        _relax(gi, c, 9)

    def fresh():
        gp = Integrator.from_case(c, "fp64")
        if setup != "periodic":
            for f in FIELDS6:
                gp.set(f, gi.get(f))
            _relax(gp, c, 9)
        return gp

    _dot_and_fd(c, gi, FIELDS6, fresh, 171)


def test_w_damping_dot_product_and_finite_differences():
    """Vertical-Courant damping of w on its active branch: a thin-layer grid
    (dz = 60 m) with a few strong updrafts, so that C = |w| dt / dz exceeds 1
    in some cells."""
    from wfe_amd import Integrator
    c = moist_case(w_damping="on", dz=60, rayleigh_zd=700, bubble_zc=400, bubble_zr=300)
    g = c.grid
    gi = make_moist(c, 81)
    w = gi.get("w")
    w3 = g.as3d(w)
    w3[g.ng + 3:g.ng + 8, g.ng + 8:g.ng + 11, g.ng + 10:g.ng + 13] += 45.0
    gi.set("w", w)
    dt = c.scalar("dt")
    assert (np.abs(w3[g.ng + 1:g.ng + g.nz]) * dt / 60.0 > 1.0).sum() >= 20

    def fresh():
        gp = Integrator.from_case(c, "fp64")
        return gp

    _dot_and_fd(c, gi, FIELDS6, fresh, 181)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_everything_in_scope_dot_product(precision):
    """All linearised terms at once -- terrain, open edges, relaxation, PD
    q_v, w damping, Coriolis and diffusion -- in FP64 and about a mixed
    context's FP32 trajectory (widened constants for every term)."""
    from wfe_amd.adjoint import ad_step, tl_step
    c = moist_case(**TERRAIN, bc_x="open", bc_y="open", pd_moist="on", w_damping="on",
                   coriolis_f=1e-3, diff_K=15.0)
    g = c.grid
    gi = make_moist(c, 91, precision)
    _relax(gi, c, 11, interior=1e-4)
    rng = np.random.default_rng(191)
    mask = interior_mask(g)
    dx, yb = random6(g, rng), random6(g, rng)
    dt = c.scalar("dt")
    dy = tl_step(gi, dt, dx)
    xb = ad_step(gi, dt, yb)
    lhs, rhs = dot6(dy, yb, mask), dot6(dx, xb, mask)
    print(precision, "dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
