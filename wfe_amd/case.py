"""Python mirror of the reference's driver-level interface over the C ABI.

``Case`` is the case file (Config + Metric + BaseState + initial state, built
natively by ``wfe_case_load``); ``Integrator`` mirrors the reference's
``Integrator`` (src/dynamics/integrator.hpp:18-42): ``step(dt, t)``,
``state()`` field access, ``rain()``; every call goes to the gfx950 context.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import (wfe_diag, FIELDS, STATE_FIELDS, WFE_PREC_FP64, WFE_PREC_MIXED, as_dptr, check, lib,
                   wfe_base, wfe_grid, wfe_metric, wfe_params, BASE_NAMES, METRIC_NAMES)

PRECISIONS = {"fp64": WFE_PREC_FP64, "mixed": WFE_PREC_MIXED}


class Grid:
    """Padded reference layout (src/core/grid.hpp:22-36)."""

    def __init__(self, nx, ny, nz, ng, dx, dy, dz):
        self.nx, self.ny, self.nz, self.ng = nx, ny, nz, ng
        self.dx, self.dy, self.dz = dx, dy, dz
        self.NX, self.NY, self.NZ = nx + 2 * ng, ny + 2 * ng, nz + 1 + 2 * ng

    @property
    def npts(self):
        return self.NX * self.NY * self.NZ

    @property
    def n_interior(self):
        return self.nx * self.ny * self.nz

    def as3d(self, a: np.ndarray) -> np.ndarray:
        """View a padded flat array as [NZ, NY, NX]."""
        return a.reshape(self.NZ, self.NY, self.NX)

    def interior(self, a: np.ndarray, nzlev: int | None = None) -> np.ndarray:
        g = self.ng
        nzl = self.nz if nzlev is None else nzlev
        return self.as3d(a)[g:g + nzl, g:g + self.ny, g:g + self.nx]


class Case:
    """An .ini case assembled by the native host builder."""

    def __init__(self, ini: str | Path, **overrides):
        L = lib()
        self.ini = str(ini)
        over = [f"{k}={v}".encode() for k, v in overrides.items()]
        arr = (C.c_char_p * max(1, len(over)))(*over)
        h = C.c_void_p()
        check(L.wfe_case_load(self.ini.encode(), len(over), arr, C.byref(h)),
              f"wfe_case_load({self.ini})")
        self._h = h
        gr = wfe_grid()
        check(L.wfe_case_grid(h, C.byref(gr)), "wfe_case_grid")
        self.grid = Grid(gr.nx, gr.ny, gr.nz, gr.ng, gr.dx, gr.dy, gr.dz)
        self._cgrid = gr
        p = wfe_params()
        check(L.wfe_case_params(h, C.byref(p)), "wfe_case_params")
        self.params = p

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().wfe_case_destroy(h)
            self._h = None

    def scalar(self, name: str) -> float:
        out = C.c_double()
        check(lib().wfe_case_scalar(self._h, name.encode(), C.byref(out)), f"scalar {name}")
        return out.value

    def array(self, name: str) -> np.ndarray:
        L = lib()
        n = L.wfe_case_array_size(self._h, name.encode())
        if n == 0:
            raise KeyError(name)
        a = np.empty(n, dtype=np.float64)
        check(L.wfe_case_array(self._h, name.encode(), as_dptr(a), n), f"array {name}")
        return a

    def param_dict(self) -> dict:
        return {f: getattr(self.params, f) for f, _ in wfe_params._fields_}

    def unused_keys(self) -> list[str]:
        buf = C.create_string_buffer(8192)
        check(lib().wfe_case_unused_keys(self._h, buf, len(buf)), "unused keys")
        return [k for k in buf.value.decode().split("\n") if k]

    def integrator(self, precision: str = "mixed") -> "Integrator":
        return Integrator.from_case(self, precision)


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Achieved HBM read+write bandwidth of a streaming copy [GB/s]."""
    out = C.c_double()
    check(lib().wfe_copy_probe(nbytes, reps, C.byref(out)), "wfe_copy_probe")
    return out.value


class Integrator:
    """gfx950 integrator context (src/dynamics/integrator.hpp:18-42)."""

    def __init__(self, handle: C.c_void_p, grid: Grid, case: Case | None, precision: str):
        self._h = handle
        self.grid = grid
        self.case = case
        self.precision = precision

    @classmethod
    def from_case(cls, case: Case, precision: str = "mixed") -> "Integrator":
        h = C.c_void_p()
        check(lib().wfe_case_create_ctx(case._h, PRECISIONS[precision], C.byref(h)),
              "wfe_case_create_ctx")
        return cls(h, case.grid, case, precision)

    @classmethod
    def from_arrays(cls, grid: Grid, params: wfe_params, arrays: dict, zt: float,
                    precision: str = "mixed") -> "Integrator":
        """Integrator::init from explicit FP64 base-state / metric arrays."""
        keep = {k: np.ascontiguousarray(arrays[k], dtype=np.float64)
                for k in BASE_NAMES + METRIC_NAMES if k in arrays}
        b = wfe_base(*[as_dptr(keep[n]) if n in keep else None for n in BASE_NAMES])
        m = wfe_metric(*[as_dptr(keep[n]) for n in METRIC_NAMES], zt)
        gr = wfe_grid(grid.nx, grid.ny, grid.nz, grid.ng, grid.dx, grid.dy, grid.dz)
        h = C.c_void_p()
        check(lib().wfe_create(C.byref(gr), C.byref(params), C.byref(b), C.byref(m),
                               PRECISIONS[precision], C.byref(h)), "wfe_create")
        return cls(h, grid, None, precision)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().wfe_destroy(h)
            self._h = None

    # -- stepping
    def step(self, dt: float, t: float = 0.0) -> None:
        if self.case is not None:
            check(lib().wfe_case_update_bdy(self.case._h, self._h, t), "boundary update")
        check(lib().wfe_step(self._h, dt, t), "wfe_step")

    def run(self, dt: float, nsteps: int, t0: float = 0.0) -> None:
        for s in range(nsteps):
            self.step(dt, t0 + s * dt)

    def sync(self) -> None:
        check(lib().wfe_sync(self._h), "wfe_sync")

    # -- fields
    def get(self, name: str) -> np.ndarray:
        two_d = name in ("rain", "tsk", "pblh", "t2m", "u10")
        n = self.grid.NX * self.grid.NY if two_d else self.grid.npts
        a = np.empty(n, dtype=np.float64)
        check(lib().wfe_download(self._h, FIELDS[name], as_dptr(a), n), f"download {name}")
        return a

    def set(self, name: str, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().wfe_upload(self._h, FIELDS[name], as_dptr(a), a.size), f"upload {name}")

    def state(self) -> dict:
        return {f: self.get(f) for f in STATE_FIELDS}

    def absmax(self, name: str) -> float:
        out = C.c_float()
        check(lib().wfe_absmax(self._h, FIELDS[name], C.byref(out)), "wfe_absmax")
        return out.value

    def diagnose(self) -> dict:
        """Interior diagnostics of the time-n state as a device reduction
        (wfe_diagnose): wmax, thmin, thmax, pipmax, qcmax, qrmax, finite."""
        d = wfe_diag()
        check(lib().wfe_diagnose(self._h, C.byref(d)), "wfe_diagnose")
        return {f: getattr(d, f) for f, _ in wfe_diag._fields_}

    def packed(self, name: str) -> np.ndarray:
        """Interior of a state field (or the rain) as float32 in output order
        through wfe_pack_field_async + wfe_output_mark/wait."""
        g = self.grid
        two_d = name in ("rain", "tsk", "pblh", "t2m", "u10")
        nlev = g.nz + 1 if name == "w" else (1 if two_d else g.nz)
        a = np.empty(g.nx * g.ny * nlev, dtype=np.float32)
        check(lib().wfe_pack_field_async(self._h, FIELDS[name],
                                         a.ctypes.data_as(C.POINTER(C.c_float)), a.size),
              f"pack {name}")
        t = C.c_longlong()
        check(lib().wfe_output_mark(self._h, C.byref(t)), "wfe_output_mark")
        check(lib().wfe_output_wait(self._h, t.value), "wfe_output_wait")
        return a

    # -- per-kernel HIP-event timing on the context's stream
    def kernel_timing(self, mode: int) -> None:
        check(lib().wfe_kernel_timing(self._h, int(mode)), "wfe_kernel_timing")

    def kernel_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        check(lib().wfe_kernel_timing_report(self._h, buf, len(buf)), "timing report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def kernel_times(self) -> dict:
        """Launch count and total ms per kernel (tuning entries removed)."""
        return {k: v for k, v in self.kernel_report().items()
                if v[0] > 0 and not k.startswith(("tuned:", "cand:", "uniform_columns:", "ttuned:",
                                                  "tcand:", "ktuned:", "kcand:"))}

    def kernel_bench(self, kernel: str, reps: int = 200) -> float:
        """Average launch time [us] of one kernel of the step, launched reps
        times back to back between two HIP events on the context's stream."""
        out = C.c_double()
        check(lib().wfe_kernel_bench(self._h, kernel.encode(), int(reps), C.byref(out)),
              "wfe_kernel_bench")
        return out.value

    def tuning(self) -> dict:
        """The acoustic-substep shape the context chose and every candidate's
        average time [ms] from its tuning pass; likewise the tendency kernel's
        (levels per thread)x(tile width) and the Kessler tile."""
        rep = self.kernel_report()
        chosen = next((k[6:] for k in rep if k.startswith("tuned:")), None)
        return {"chosen": chosen,
                "candidates_ms": {k[5:]: v[1] for k, v in rep.items() if k.startswith("cand:")},
                "uniform_columns": any(k == "uniform_columns:1" for k in rep),
                "tendencies": next((k[7:] for k in rep if k.startswith("ttuned:")), None),
                "tendency_candidates_ms": {k[6:]: v[1] for k, v in rep.items()
                                           if k.startswith("tcand:")},
                "kessler": next((k[7:] for k in rep if k.startswith("ktuned:")), None),
                "kessler_candidates_ms": {k[6:]: v[1] for k, v in rep.items()
                                          if k.startswith("kcand:")}}

    # -- phase-level entry points (kernels.hpp free functions)
    def apply_bcs(self) -> None:
        check(lib().wfe_apply_bcs(self._h), "wfe_apply_bcs")

    def compute_tendencies(self, dt: float) -> None:
        check(lib().wfe_compute_tendencies(self._h, dt), "wfe_compute_tendencies")

    def acoustic_substeps(self, dtau: float, nsub: int) -> None:
        check(lib().wfe_acoustic_substeps(self._h, dtau, nsub), "wfe_acoustic_substeps")

    def kessler(self, dt: float) -> None:
        check(lib().wfe_kessler(self._h, dt), "wfe_kessler")


class Ensemble:
    """Members of any grid size advancing together, one hipGraph launch per
    step (wfe_ensemble_*); compatible members share one launch per acoustic
    substep kernel. Keeps references to its members."""

    def __init__(self, members: list[Integrator]):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = C.c_void_p()
        check(lib().wfe_ensemble_create(arr, len(self.members), C.byref(h)), "wfe_ensemble_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().wfe_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def step(self, dts, times=None) -> None:
        """One step of every member, member q by dts[q]; times[q] is the time
        at the start of its step (needed by members with surface physics)."""
        n = len(self.members)
        arr = (C.c_double * n)(*[float(d) for d in dts])
        if times is None:
            check(lib().wfe_ensemble_step(self._h, arr), "wfe_ensemble_step")
        else:
            ts = (C.c_double * n)(*[float(t) for t in times])
            check(lib().wfe_ensemble_step_at(self._h, arr, ts), "wfe_ensemble_step_at")

    def sync(self) -> None:
        check(lib().wfe_ensemble_sync(self._h), "wfe_ensemble_sync")

    @property
    def batched_members(self) -> int:
        """Members advanced in batched substep launches by the last captured step."""
        n = C.c_int()
        check(lib().wfe_ensemble_batched(self._h, C.byref(n)), "wfe_ensemble_batched")
        return n.value
