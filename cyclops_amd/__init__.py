"""MI355X-native cyclic-coordinate-descent engine for the Cyclops regression
models (logistic, Poisson, Breslow Cox, conditional logistic).

The compute path is libcyclops_amd.so (hand-written HIP kernels for gfx950
behind the C ABI in include/cyclops_amd.h); this package is the host-side
mirror of the reference's fitting interface over that ABI.
"""
from . import _abi  # noqa: F401

__version__ = "0.1.0"
