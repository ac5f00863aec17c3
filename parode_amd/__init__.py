"""ParODE for one AMD MI355X: the parallel-in-time fixed-step LTI solver
(RK4 / ABM4 step maps composed by a Blelloch scan) behind the reference's
C ABI, built as libParODEShared.so from HIP kernels for gfx950."""
from . import api, synth  # noqa: F401
from .api import ABM, FP64, MIXED, PLAIN32, RK4, ParODEError, Plan  # noqa: F401
