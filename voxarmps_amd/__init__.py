"""voxarmps_amd — MI355X-native MPS time-step pipeline (HIP, gfx950).

The compute lives in libvoxarmps.so (hand-written HIP kernels behind the C ABI
of include/voxarmps.h); this package only binds it and generates cases.
"""
from ._abi import F, FIELDS, LIST_L, LIST_S, VoxParams, VoxSolveInfo, VoxStatus  # noqa: F401
from .cases import (Case, batch_cases, box_of, dam_break, dam_break_2d_c0, dam_break_3d_c1,  # noqa: F401
                    ensemble_2d, ensemble_c4, jitter, jittered_dam_breaks_2d, two_phase)
from .params import (dam2d_params, dam3d_params, default_params, ensemble_params,  # noqa: F401
                     set_box, two_radius, wc_params)
from .sim import Backend, Simulation, VoxError, hip_backend, read_vtu  # noqa: F401
from .lpc import LearnedPreconditioner  # noqa: F401

__all__ = [
    "F", "FIELDS", "LIST_L", "LIST_S", "VoxParams", "VoxSolveInfo", "VoxStatus", "Case",
    "dam_break", "dam_break_2d_c0", "dam_break_3d_c1", "jitter", "dam2d_params", "dam3d_params",
    "default_params", "set_box", "two_radius", "Backend", "Simulation", "VoxError", "hip_backend",
    "LearnedPreconditioner", "read_vtu",
]
