"""vibestran_amd -- MI355X-native hip-pcg backend for vibestran's linear-static
hot path (CSR stiffness assembly + Jacobi-PCG), driving hand-written gfx950
HIP kernels through the C ABI in include/vbh.h (libvbh.so)."""
from .backend import CsrData, HipMixedPcgSolverBackend, HipPcgSolverBackend, SolverError

__all__ = ["CsrData", "HipPcgSolverBackend", "HipMixedPcgSolverBackend", "SolverError"]
