"""MI355X free-vortex-wake induced-velocity engine (gfx950).

Drop-in for WInDS's BiotSavartMex / BiotSavart.m: the arithmetic runs in
hand-written HIP kernels behind the C ABI of include/winds_bs.h.
"""
from ._lib import (CORE_CUTOFF, CORE_VATISTAS, PREC_FP32C, PREC_FP64, SEM_MATLAB, SEM_MEX,
                   WindsError, lib, opts)
from .biotsavart import BiotSavart, BiotSavartMex

__all__ = [
    "BiotSavart", "BiotSavartMex", "WindsError", "lib", "opts",
    "CORE_CUTOFF", "CORE_VATISTAS", "SEM_MEX", "SEM_MATLAB", "PREC_FP64", "PREC_FP32C",
]
