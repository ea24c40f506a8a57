"""MI355X-native convergent Born series solver (drop-in backend for the CBS
forward/adjoint solves of BMOLKAIST/Inverse_solver).

The compute path is libcbs_hip.so (hand-written HIP kernels for gfx950) behind
the C ABI in include/cbs_hip.h; this package mirrors the reference's MATLAB
class API on top of it.
"""
from .forward_solver import FORWARD_SOLVER_CONVERGENT_BORN, get_default_parameters  # noqa: F401
