"""cellweave_amd: the cellweave diffusion field solve on MI355X (gfx950).

The solver is libcwfield.so (HIP kernels behind the C ABI in
include/cwfield.h); cellweave_amd.field is a thin Python handle on it and
cellweave_amd.cases builds seeded synthetic workloads.
"""
from . import _abi  # noqa: F401

__version__ = "0.1.0"
