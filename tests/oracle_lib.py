"""ctypes view of the CPU FP64 oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: it checks the HIP path; it is never the product.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# WFE_ORACLE_LIB points the tests at another build of the oracle (the
# mutation test loads deliberately broken copies through it).
ORACLE_PATH = Path(os.environ.get("WFE_ORACLE_LIB") or
                   Path(__file__).resolve().parent.parent / "oracle" / "liboracle.so")
_DP = C.POINTER(C.c_double)
_o = None

STATE = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")
BASE = ("thb", "pib", "rhob", "dthbdz", "thbw", "rhobw", "thvb", "thvbw", "qvb", "ub", "vb")
METRIC = ("zeta_w", "zeta_c", "dzeta_c", "dzeta_w", "h", "hx", "hy", "hx_u", "hy_v", "jac",
          "fcor", "mapf")
PARAMS = ("diff_K", "coriolis_use_ub", "rayleigh_zd", "rayleigh_alpha", "acoustic_ns",
          "acoustic_beta", "acoustic_smdiv", "bc_x_open", "bc_y_open", "cstar", "moisture",
          "microphysics", "pd_moist", "w_damping")


def olib() -> C.CDLL:
    global _o
    if _o is None:
        if not ORACLE_PATH.exists():
            raise ImportError(f"{ORACLE_PATH} missing: run `python -m wfe_amd.build_ext`")
        L = C.CDLL(str(ORACLE_PATH))
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 4 + [C.c_double] * 4
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_param.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_set_array.argtypes = [C.c_void_p, C.c_char_p, _DP, C.c_long]
        L.orc_get_array.argtypes = [C.c_void_p, C.c_char_p, _DP, C.c_long]
        for n in ("orc_init", "orc_apply_bcs"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_set_bdy.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_set_physics.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_sfc_psi.restype = C.c_double
        L.orc_sfc_psi.argtypes = [C.c_int, C.c_double]
        L.orc_sfc_diffuse.argtypes = [C.c_int, C.c_double] + [_DP] * 5 + [C.c_double, _DP, _DP]
        L.orc_sfc_cosz.restype = C.c_double
        L.orc_sfc_cosz.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_sfc_louis.argtypes = [C.c_double] * 3 + [_DP]
        L.orc_sfc_diag.argtypes = [C.c_double] * 6 + [_DP]
        L.orc_sfc_lw.argtypes = [C.c_int, _DP, _DP, C.c_double, _DP, _DP]
        L.orc_sfc_sw.restype = C.c_double
        L.orc_sfc_sw.argtypes = [C.c_int, C.c_double, C.c_double, _DP, _DP, _DP]
        L.orc_sfc_soil.argtypes = [C.c_double] * 4 + [_DP]
        L.orc_sfc_soil_heat.restype = C.c_double
        L.orc_sfc_soil_heat.argtypes = [C.c_double, C.c_int]
        L.orc_tendencies.argtypes = [C.c_void_p, C.c_double]
        L.orc_acoustic.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_kessler.argtypes = [C.c_void_p, C.c_double]
        L.orc_absmax.restype = C.c_double
        L.orc_absmax.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_iface5.restype = C.c_double
        L.orc_iface5.argtypes = [C.c_double] * 7
        L.orc_thomas.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _DP]
        L.orc_qsat_w.restype = C.c_double
        L.orc_qsat_w.argtypes = [C.c_double, C.c_double]
        L.orc_qsat_i.restype = C.c_double
        L.orc_qsat_i.argtypes = [C.c_double, C.c_double]
        L.orc_case_load.restype = C.c_void_p
        L.orc_case_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p)]
        L.orc_case_destroy.argtypes = [C.c_void_p]
        L.orc_case_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int), _DP]
        L.orc_case_value.argtypes = [C.c_void_p, C.c_char_p, _DP]
        L.orc_case_array_size.restype = C.c_long
        L.orc_case_array_size.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_case_array.argtypes = [C.c_void_p, C.c_char_p, _DP, C.c_long]
        L.orc_case_update_bdy.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        _o = L
    return _o


def _p(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class OracleCase:
    """The oracle's own case assembly (oracle/case_builder.cpp) from an .ini
    file plus key=value overrides: grid, run values, DynParams and every
    host array, built independently of the product's builders."""

    def __init__(self, ini, **over):
        L = olib()
        ov = [f"{k}={v}".encode() for k, v in over.items()]
        arr = (C.c_char_p * max(1, len(ov)))(*ov)
        self._h = C.c_void_p(L.orc_case_load(str(ini).encode(), len(ov), arr))
        if not self._h.value:
            raise ValueError(f"oracle case builder rejected {ini} {over}")
        dims = (C.c_int * 3)()
        d = (C.c_double * 3)()
        L.orc_case_dims(self._h, dims, d)
        from wfe_amd import Grid
        self.grid = Grid(dims[0], dims[1], dims[2], 3, d[0], d[1], d[2])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            olib().orc_case_destroy(h)
            self._h = None

    def value(self, name: str) -> float:
        out = C.c_double()
        if olib().orc_case_value(self._h, name.encode(), C.byref(out)) != 0:
            raise KeyError(name)
        return out.value

    def array(self, name: str) -> np.ndarray:
        n = olib().orc_case_array_size(self._h, name.encode())
        if n <= 0:
            raise KeyError(name)
        a = np.empty(n, dtype=np.float64)
        assert olib().orc_case_array(self._h, name.encode(), _p(a), n) == 0, name
        return a

    def params(self) -> dict:
        return {k: self.value(k) for k in PARAMS}

    def model(self) -> "OracleModel":
        """An oracle model of this case: base state, metric, initial state
        and (file mode) the boundary weights and first targets, all from the
        oracle's own builders; step it with step_case()."""
        arrays = {k: self.array(k) for k in BASE + METRIC}
        state = {k: self.array(k) for k in STATE}
        om = OracleModel(self.grid, self.value("zt"), self.params(), arrays, state)
        if self.value("file_mode"):
            om.set("bdy_wgt", self.array("bdy_wgt"))
            for v, n in enumerate(("u", "v", "thp", "pip", "qv")):
                om.set(f"bdy_lo_{v}", self.array(f"bdy_lo_{n}"))
                om.set(f"bdy_hi_{v}", self.array(f"bdy_hi_{n}"))
            om.set_bdy(True, 0.0)
        if self.value("sfc_on"):
            for n in ("tsk", "land", "lat", "lon", "soilw"):
                om.set(n, self.array(f"sfc_{n}"))
            olib().orc_set_physics(om._h, int(self.value("sfc_nonlocal")),
                                   self.value("sfc_start_hour"), int(self.value("sfc_doy")))
        om.case = self
        return om


class OracleModel:
    """Oracle instance fed the same FP64 arrays as the GPU context."""

    def __init__(self, grid, zt: float, params: dict, arrays: dict, state: dict | None = None):
        L = olib()
        self.grid = grid
        self._h = C.c_void_p(L.orc_create(grid.nx, grid.ny, grid.nz, grid.ng, grid.dx, grid.dy,
                                          grid.dz, zt))
        if not self._h.value:
            raise ValueError("orc_create rejected the grid")
        for k in PARAMS:
            if k in params:
                assert L.orc_set_param(self._h, k.encode(), float(params[k])) == 0, k
        for k in BASE + METRIC:
            if k in arrays:
                self.set(k, arrays[k])
        if state:
            for k, v in state.items():
                self.set(k, v)
        L.orc_init(self._h)

    @classmethod
    def from_case(cls, case) -> "OracleModel":
        arrays = {k: case.array(k) for k in BASE + METRIC}
        state = {k: case.array(k) for k in STATE}
        om = cls(case.grid, case.scalar("zt"), case.param_dict(), arrays, state)
        try:
            wgt = case.array("bdy_wgt")
        except KeyError:
            wgt = None
        if wgt is not None and wgt.size:
            om.set("bdy_wgt", wgt)
            for v, n in enumerate(("u", "v", "thp", "pip", "qv")):
                om.set(f"bdy_lo_{v}", case.array(f"bdy_lo_{n}"))
                om.set(f"bdy_hi_{v}", case.array(f"bdy_hi_{n}"))
            om.set_bdy(True, 0.0)
        return om

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            olib().orc_destroy(h)
            self._h = None

    def _size(self, name: str) -> int:
        g = self.grid
        if name in ("ub", "vb", "zeta_w", "zeta_c", "dzeta_c", "dzeta_w"):
            return g.NZ
        if name in ("h", "hx", "hy", "hx_u", "hy_v", "jac", "fcor", "mapf", "rain", "bdy_wgt",
                    "tsk", "land", "lat", "lon", "soilw", "cdv", "pblh", "t2m", "u10", "tdeep"):
            return g.NX * g.NY
        if name == "soilt":
            return 4 * g.NX * g.NY
        return g.npts

    def set(self, name: str, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert olib().orc_set_array(self._h, name.encode(), _p(a), a.size) == 0, name

    def get(self, name: str) -> np.ndarray:
        a = np.empty(self._size(name), dtype=np.float64)
        assert olib().orc_get_array(self._h, name.encode(), _p(a), a.size) == 0, name
        return a

    def set_bdy(self, active: bool, tfrac: float) -> None:
        olib().orc_set_bdy(self._h, int(active), float(tfrac))

    case = None

    def step(self, dt: float, t: float = 0.0) -> None:
        if self.case is not None:
            rc = olib().orc_case_update_bdy(self.case._h, self._h, t)
            assert rc == 0, f"oracle boundary update failed ({rc}) at t={t}"
        olib().orc_step(self._h, dt, t)

    def run(self, dt: float, n: int) -> None:
        for s in range(n):
            self.step(dt, s * dt)

    def apply_bcs(self):
        olib().orc_apply_bcs(self._h)

    def tendencies(self, dt: float):
        olib().orc_tendencies(self._h, dt)

    def acoustic(self, dtau: float, nsub: int):
        olib().orc_acoustic(self._h, dtau, nsub)

    def kessler(self, dt: float):
        olib().orc_kessler(self._h, dt)

    def absmax(self, name: str) -> float:
        return olib().orc_absmax(self._h, name.encode())


def set_threads(n: int) -> None:
    olib().orc_set_threads(int(n))


def iface5(*a) -> float:
    return olib().orc_iface5(*[float(x) for x in a])


def thomas(a, b, c, d) -> np.ndarray:
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, d)]
    x = np.empty_like(arrs[3])
    olib().orc_thomas(len(x), *[_p(t) for t in arrs], _p(x))
    return x
