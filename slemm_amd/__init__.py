"""slemm_amd -- MI355X (gfx950) engine for SLEMM's stochastic-Lanczos REML path.

The compute path is the HIP library ``libslemm_gpu.so`` (C ABI in ``include/slemm_gpu.h``),
reached through :mod:`slemm_amd.gpu`.  There is no CPU fallback.
"""
from . import _lib  # noqa: F401

__all__ = ["gpu", "_lib"]
__version__ = "0.1.0"
