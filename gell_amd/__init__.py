"""MI355X (gfx950) implementation of the Gell per-step hot path.

Mechanics (Morton cell list, radix sort, adhesion/repulsion forces, AB2
movement), LOD diffusion (x/y/z Thomas sweeps) and proliferation/death
(Philox-driven division, stream compaction) run as hand-written HIP kernels
behind the C ABI of ``include/gell_abi.h``; ``gell_amd.runtime`` binds that
ABI from Python.
"""
from .runtime import (  # noqa: F401 (re-exported)
    CELL_DTYPE, Context, DeviceBuffer, Ensemble, GellError, Params, Surrogate, default_params, float_loop_count,
    load_library, make_params)
