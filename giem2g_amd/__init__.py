"""giem2g_amd -- MI355X-native integral-equation solve of GIEM2G.

The compute path is the HIP library giem2g_amd/_lib/libgiem2g_amd.so behind
the C ABI of include/giem2g_amd.h; this package holds its ctypes binding
(lib), the Python mirror of the reference operator interface (operator) and
the synthetic benchmark models (models).
"""
from . import models  # noqa: F401

__all__ = ["models", "lib", "operator"]
