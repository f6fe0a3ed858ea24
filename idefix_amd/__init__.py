"""idefix_amd: MI355X-native per-stage finite-volume update of Idefix.

The hot path (reconstruction, Riemann fluxes, flux divergence, constrained transport,
conservative/primitive conversion, CFL reduction) runs in hand-written gfx950 HIP
kernels exported through the C ABI of include/idefix_amd.h (libidefix_amd.so).
"""
from . import capi  # noqa: F401
from . import problems  # noqa: F401

__all__ = ["capi", "problems", "DataBlock"]


def __getattr__(name):
    if name == "DataBlock":
        import importlib
        return importlib.import_module(".datablock", __name__).DataBlock
    raise AttributeError(name)
