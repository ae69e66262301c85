"""MI355X (gfx950) RHMC hot path: staggered Dslash, multi-shift CG, fermion
force and SU(3) link/momentum update behind the C ABI of include/gpu.h and
include/cu_inverter.h.  This package holds the ctypes binding, synthetic
field generators and the rational-approximation generator."""
from .lib import Gpu, RhmcError, load, header_functions, multi_trajectory, traj_params  # noqa: F401
