"""MI355X (gfx950) implementation of the WFE large-step dynamics hot path.

The compute path is ``wfe_amd/lib/libwfe.so`` (hand-written HIP kernels behind
the C ABI in ``include/wfe_hip.h``); this package is the Python view of it.
"""
from ._lib import FIELDS, STATE_FIELDS, WfeError, lib  # noqa: F401
from .case import Case, Ensemble, Grid, Integrator  # noqa: F401

__all__ = ["Case", "Ensemble", "Grid", "Integrator", "WfeError", "lib", "FIELDS", "STATE_FIELDS"]
