"""ctypes binding of the native library ``wfe_amd/lib/libwfe.so``.

The library is the product: the gfx950 HIP kernels behind the C ABI of
``include/wfe_hip.h`` plus the host case builder. There is no fallback: if the
library is missing, importing anything that needs it raises ``ImportError``
naming the build command.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libwfe.so"

WFE_OK = 0
WFE_PREC_FP64 = 0
WFE_PREC_MIXED = 1

FIELDS = {
    "u": 0, "v": 1, "w": 2, "thp": 3, "pip": 4, "qv": 5, "qc": 6, "qr": 7,
    "tu": 8, "tv": 9, "tw": 10, "tthp": 11, "tpip": 12, "tqv": 13, "tqc": 14, "tqr": 15,
    "rain": 16, "tsk": 17, "pblh": 18, "t2m": 19, "u10": 20,
}
STATE_FIELDS = ("u", "v", "w", "thp", "pip", "qv", "qc", "qr")


class wfe_grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class wfe_params(C.Structure):
    _fields_ = [("diff_K", C.c_double), ("coriolis_f", C.c_double),
                ("coriolis_use_ub", C.c_int), ("rayleigh_zd", C.c_double),
                ("rayleigh_alpha", C.c_double), ("acoustic_ns", C.c_int),
                ("acoustic_beta", C.c_double), ("acoustic_smdiv", C.c_double),
                ("bc_x_open", C.c_int), ("bc_y_open", C.c_int), ("cstar", C.c_double),
                ("moisture", C.c_int), ("microphysics", C.c_int), ("pd_moist", C.c_int),
                ("w_damping", C.c_int), ("w_abort", C.c_double)]


_DP = C.POINTER(C.c_double)
BASE_NAMES = ("thb", "pib", "rhob", "dthbdz", "thbw", "rhobw", "thvb", "thvbw", "qvb", "ub", "vb")
METRIC_NAMES = ("zeta_w", "zeta_c", "dzeta_c", "dzeta_w", "h", "hx", "hy", "hx_u", "hy_v",
                "jac", "fcor", "mapf")


class wfe_base(C.Structure):
    _fields_ = [(n, _DP) for n in BASE_NAMES]


class wfe_metric(C.Structure):
    _fields_ = [(n, _DP) for n in METRIC_NAMES] + [("zt", C.c_double)]


class wfe_sfc(C.Structure):
    _fields_ = [("tsk", _DP), ("land", _DP), ("lat", _DP), ("lon", _DP), ("soilw", _DP),
                ("start_hour_utc", C.c_double), ("doy", C.c_int), ("nonlocal", C.c_int)]


class wfe_diag(C.Structure):
    _fields_ = [("wmax", C.c_double), ("thmin", C.c_double), ("thmax", C.c_double),
                ("pipmax", C.c_double), ("qcmax", C.c_double), ("qrmax", C.c_double),
                ("finite", C.c_int)]


_lib = None

# (name, restype, argtypes) of every entry point in include/wfe_hip.h.
SIGNATURES = [
    ("wfe_create", C.c_int, [C.POINTER(wfe_grid), C.POINTER(wfe_params), C.POINTER(wfe_base),
                             C.POINTER(wfe_metric), C.c_int, C.POINTER(C.c_void_p)]),
    ("wfe_destroy", C.c_int, [C.c_void_p]),
    ("wfe_step", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("wfe_set_physics", C.c_int, [C.c_void_p, C.POINTER(wfe_sfc), C.c_size_t]),
    ("wfe_upload", C.c_int, [C.c_void_p, C.c_int, _DP, C.c_size_t]),
    ("wfe_download", C.c_int, [C.c_void_p, C.c_int, _DP, C.c_size_t]),
    ("wfe_absmax", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    ("wfe_sync", C.c_int, [C.c_void_p]),
    ("wfe_diagnose", C.c_int, [C.c_void_p, C.POINTER(wfe_diag)]),
    ("wfe_diagnose_fields", C.c_int, [C.POINTER(wfe_grid), C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(wfe_diag)]),
    ("wfe_diagnose_fields_noalloc", C.c_int, [C.POINTER(wfe_grid), C.c_int, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.POINTER(wfe_diag)]),
    ("wfe_pack_field_async", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
    ("wfe_output_mark", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    ("wfe_output_wait", C.c_int, [C.c_void_p, C.c_longlong]),
    ("wfe_device_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("wfe_device_free", C.c_int, [C.c_void_p]),
    ("wfe_device_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("wfe_device_download", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("wfe_host_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("wfe_host_free", C.c_int, [C.c_void_p]),
    ("wfe_set_bdy_weights", C.c_int, [C.c_void_p, _DP, C.c_size_t]),
    ("wfe_set_bdy_target", C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP, C.c_size_t]),
    ("wfe_swap_bdy_targets", C.c_int, [C.c_void_p]),
    ("wfe_set_bdy_tfrac", C.c_int, [C.c_void_p, C.c_double]),
    ("wfe_apply_bcs", C.c_int, [C.c_void_p]),
    ("wfe_compute_tendencies", C.c_int, [C.c_void_p, C.c_double]),
    ("wfe_acoustic_substeps", C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    ("wfe_kessler", C.c_int, [C.c_void_p, C.c_double]),
    ("wfe_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("wfe_kernel_timing_report", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("wfe_kernel_bench", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    ("wfe_tl_step", C.c_int, [C.c_void_p, C.c_double, C.POINTER(_DP), C.POINTER(_DP), C.c_size_t]),
    ("wfe_ad_step", C.c_int, [C.c_void_p, C.c_double, C.POINTER(_DP), C.POINTER(_DP), C.c_size_t]),
    ("wfe_adj_gradient", C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, _DP, C.c_size_t,
                                   C.POINTER(C.c_double), C.POINTER(_DP), C.c_size_t]),
    ("wfe_tl_step_q", C.c_int, [C.c_void_p, C.c_double, C.POINTER(_DP), C.POINTER(_DP),
                                C.c_size_t]),
    ("wfe_ad_step_q", C.c_int, [C.c_void_p, C.c_double, C.POINTER(_DP), C.POINTER(_DP),
                                C.c_size_t]),
    ("wfe_adj_gradient_q", C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, _DP, C.c_size_t,
                                     C.POINTER(C.c_double), C.POINTER(_DP), C.c_size_t]),
    ("wfe_ensemble_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    ("wfe_ensemble_destroy", C.c_int, [C.c_void_p]),
    ("wfe_ensemble_step", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("wfe_ensemble_step_at", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("wfe_ensemble_sync", C.c_int, [C.c_void_p]),
    ("wfe_ensemble_batched", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wfe_copy_probe", C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    ("wfe_last_error", C.c_char_p, []),
    ("wfe_abi_version", C.c_int, []),
    ("wfe_device_name", C.c_int, [C.c_char_p, C.c_size_t]),
    ("wfe_case_load", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p),
                                C.POINTER(C.c_void_p)]),
    ("wfe_case_destroy", C.c_int, [C.c_void_p]),
    ("wfe_case_grid", C.c_int, [C.c_void_p, C.POINTER(wfe_grid)]),
    ("wfe_case_params", C.c_int, [C.c_void_p, C.POINTER(wfe_params)]),
    ("wfe_case_scalar", C.c_int, [C.c_void_p, C.c_char_p, _DP]),
    ("wfe_case_array_size", C.c_size_t, [C.c_void_p, C.c_char_p]),
    ("wfe_case_array", C.c_int, [C.c_void_p, C.c_char_p, _DP, C.c_size_t]),
    ("wfe_case_create_ctx", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("wfe_case_update_bdy", C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    ("wfe_case_unused_keys", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
]


class WfeError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libwfe.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m wfe_amd.build_ext` (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != WFE_OK:
        msg = lib().wfe_last_error().decode(errors="replace")
        raise WfeError(f"{what} failed ({rc}): {msg}")


def as_dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


class DeviceArray:
    """A device copy of a numpy array in memory this library allocates
    (wfe_device_alloc), freed with the object."""

    def __init__(self, a: np.ndarray):
        self.dtype = a.dtype
        self.shape = a.shape
        self.nbytes = a.nbytes
        self.ptr = C.c_void_p()
        check(lib().wfe_device_alloc(max(1, self.nbytes), C.byref(self.ptr)), "wfe_device_alloc")
        a = np.ascontiguousarray(a)
        check(lib().wfe_device_upload(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes),
              "wfe_device_upload")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().wfe_device_download(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes),
              "wfe_device_download")
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p is not None and p.value and _lib is not None:
            lib().wfe_device_free(p)
            self.ptr = None
