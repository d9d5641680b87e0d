"""sfr3d_amd -- MI355X (gfx950) linear-solve path for SFR3D-assembled systems.

Jacobi-preconditioned CG, BiCGSTAB, BiCGSTAB(2), GMRES(15) and MINRES over
SFR3D's CSR-family and ELLPACK storages, run by hand-written HIP kernels
behind a C ABI (include/sfr3d.h, include/sfr3d_setup.h) that keeps the
reference's solverC contract.
"""
from . import _lib
from ._lib import (PCG, PBICGSTAB, PBICGSTABL2, GMRES, MINRES, CSR, CSRD,
                   CSRC, ELLPACK, CSRDCOO, CSRCCOO, PREC_FP64, PREC_MIXED, FMT_AUTO, FMT_ELL, FMT_CSR, LIB_PATH)
from .solver import (Context, System, DeviceMatrix, DeviceArray, Batch,
                     SolverError)
from . import structure, systems

__all__ = ["Context", "System", "DeviceMatrix", "DeviceArray", "Batch",
           "SolverError",
           "structure", "systems", "PCG", "PBICGSTAB", "PBICGSTABL2", "GMRES", "MINRES", "CSR", "CSRD", "CSRC",
           "ELLPACK", "CSRDCOO", "CSRCCOO", "PREC_FP64", "PREC_MIXED",
           "FMT_AUTO", "FMT_ELL", "FMT_CSR", "LIB_PATH"]
