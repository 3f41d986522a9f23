"""Tangent-linear and adjoint of the large step through the C ABI, and a
torch.autograd.Function for the lowest-level kinetic-energy objective.

Scope (enforced by the library): FP64 or mixed contexts without
microphysics or surface physics. Terrain, map factors, w damping, open boundaries,
Davies relaxation (targets held at the context's current ones), Coriolis and
constant-K diffusion are included. With moisture on, q_v is a linearised
field, by flux-form or positive-definite transport, with q_c = q_r = 0.
Fields are (u, v, w, thp, pip, qv) in the padded layout of the case's grid;
qv is zero on output for a dry case. In a mixed context the linearisation is
about the FP32 trajectory the context itself computes; perturbations,
adjoint variables and the objective are FP64 in both kinds of context.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

FIELDS5 = ("u", "v", "w", "thp", "pip")
FIELDS = FIELDS5 + ("qv",)
_DP = C.POINTER(C.c_double)


def _ptrs(arrays):
    return (_DP * len(arrays))(*[a.ctypes.data_as(_DP) for a in arrays])


def _prep(gi, d: dict):
    n = gi.grid.npts
    out = []
    for f in FIELDS:
        a = np.ascontiguousarray(d.get(f, np.zeros(n)), dtype=np.float64)
        if a.size != n:
            raise ValueError(f"{f}: expected {n} elements, got {a.size}")
        out.append(a)
    return out


def tl_step(gi, dt: float, dx: dict) -> dict:
    """dy = M'(x) dx for one large step around the integrator's state."""
    xin = _prep(gi, dx)
    out = [np.zeros(gi.grid.npts) for _ in FIELDS]
    check(lib().wfe_tl_step_q(gi._h, dt, _ptrs(xin), _ptrs(out), gi.grid.npts), "wfe_tl_step_q")
    return dict(zip(FIELDS, out))


def ad_step(gi, dt: float, yb: dict) -> dict:
    """xb = M'(x)^T yb for one large step around the integrator's state."""
    yin = _prep(gi, yb)
    out = [np.zeros(gi.grid.npts) for _ in FIELDS]
    check(lib().wfe_ad_step_q(gi._h, dt, _ptrs(yin), _ptrs(out), gi.grid.npts), "wfe_ad_step_q")
    return dict(zip(FIELDS, out))


def ke_gradient(gi, dt: float, nsteps: int, ckpt_every: int, wgt2d: np.ndarray):
    """J = (1/N) sum_n sum_ij wgt (u^2 + v^2)/2 at the lowest level over N
    steps from the current state, and dJ/d(initial state) as a dict."""
    g = gi.grid
    w = np.ascontiguousarray(wgt2d, dtype=np.float64)
    if w.size != g.NX * g.NY:
        raise ValueError("wgt2d must be the padded NX*NY plane")
    grads = [np.zeros(g.npts) for _ in FIELDS]
    J = C.c_double()
    check(lib().wfe_adj_gradient_q(gi._h, dt, int(nsteps), int(ckpt_every),
                                   w.ctypes.data_as(_DP), w.size, C.byref(J), _ptrs(grads),
                                   g.npts), "wfe_adj_gradient_q")
    return J.value, dict(zip(FIELDS, grads))


def box_weights(grid, i0: int, i1: int, j0: int, j1: int) -> np.ndarray:
    """Padded NX*NY plane equal to 1 on interior columns [i0, i1) x [j0, j1)."""
    w = np.zeros((grid.NY, grid.NX))
    w[grid.ng + j0:grid.ng + j1, grid.ng + i0:grid.ng + i1] = 1.0
    return w.ravel()


def kinetic_energy_function():
    """torch.autograd.Function mapping the initial (u, v, w, thp, pip) tensors,
    and qv as a sixth when given (padded layout, float64), to the mean
    lowest-level kinetic energy J over `nsteps` steps of an integrator;
    backward returns dJ/dx0 from the hand-written adjoint with
    checkpointing."""
    import torch

    class KineticEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, gi, dt, nsteps, ckpt_every, wgt2d, *xs):
            if len(xs) not in (5, 6):
                raise ValueError("expected (u, v, w, thp, pip) or (u, v, w, thp, pip, qv)")
            names = FIELDS[:len(xs)]
            for name, t in zip(names, xs):
                gi.set(name, t.detach().cpu().numpy().astype(np.float64))
            J, grads = ke_gradient(gi, dt, nsteps, ckpt_every, wgt2d)
            ctx.grads = [torch.from_numpy(grads[f]).to(xs[0].device) for f in names]
            return torch.tensor(J, dtype=torch.float64, device=xs[0].device)

        @staticmethod
        def backward(ctx, gJ):
            return (None, None, None, None, None) + tuple(gJ * g for g in ctx.grads)

    return KineticEnergy
