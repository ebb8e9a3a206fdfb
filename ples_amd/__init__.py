"""MI355X-native pressure-Poisson multigrid for the PLES projection step.

The compute path is libples_mg.so (HIP, gfx950) behind the C ABI declared in
include/ples_mg.h; ``ples_amd.mg`` is its Python host binding.
"""
from . import mg  # noqa: F401

__all__ = ["mg"]
