"""MI355X-native NUFFT backend with the jax_finufft API (nufft1 / nufft2 / nufft3).

Mirrors src/jax_finufft/__init__.py of the reference: the public names are
``nufft1``, ``nufft2``, ``nufft3`` and the ``options`` module (``Opts``,
``NestedOpts``).  Arrays are torch tensors on a HIP device; every transform
runs through the C-ABI library ``libjfa.so`` (hand-written gfx950 kernels
plus rocFFT).
"""

__all__ = ["__version__", "nufft1", "nufft2", "nufft3", "nufft1_ragged", "options", "Plan",
           "RaggedPlan"]

__version__ = "0.1.0"

from . import options
from ._native import Plan
from .ops import nufft1, nufft2, nufft3
from .ragged import RaggedPlan, nufft1_ragged
