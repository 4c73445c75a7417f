"""MI355X-native Stable Cosserat Rods step (KamiSimulator drop-in).

The product path is the HIP library libkscr.so (C ABI in include/kscr.h);
this package holds its Python binding, the solver mirror of the reference's
Solver plugin interface, and the synthetic workload generators.
"""
from . import scenes  # noqa: F401
from ._lib import KscrError, KscrParams, Sim, default_params, load, mesh_sdf, selftest_rounding  # noqa: F401

__all__ = ["Sim", "KscrParams", "KscrError", "default_params", "load", "scenes"]
