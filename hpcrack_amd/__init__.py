"""MI355X (gfx950) implementation of HPCrack's matrix-free multigrid/PCG
phase-field fracture solves.  The compute path is libhpcrack_amd.so (HIP
kernels behind the C-ABI of include/hpcrack.h); this package is its host-side
Python mirror."""
from .capi import (D, DOLD, DPREV, FD, FLAG, FU, G, GC, GE, H, K, KE, P, PD, RD, RU, U, UOLD, W, WD)  # noqa: F401
from .solver import Fuser, HpcrError, Solver, device_count, estimate_bytes  # noqa: F401

__all__ = ["Solver", "Fuser", "HpcrError", "device_count", "estimate_bytes"]
