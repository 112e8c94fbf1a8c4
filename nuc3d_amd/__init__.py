"""nuc3d_amd -- MI355X (gfx950) implementation of NUC3d's per-block RHS and
TVD-RK3 update.

The product is the C-ABI library `libnuc3d_hip.so` (include/nuc3d_hip.h),
built from nuc3d_amd/csrc/ by __graft_entry__.build().  This package loads it
with ctypes and offers a thin Python view of a case (`nuc3d_amd.solver.Case`)
for tests and benchmarks.  There is no CPU fallback: importing the solver
without the built library raises.  NUC3D_LIB names another build of the same
library (for example one compiled with different tile or occupancy settings).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NUC3D_LIB") or os.path.join(PKG_DIR, "libnuc3d_hip.so")

__all__ = ["PKG_DIR", "LIB_PATH"]
