"""Host-side mirror of the reference's operator interface.

``BiotSavartMex`` keeps the MEX gateway's 11-input / 2-output contract
(cuda/source/BiotSavartMex.cu:14-28, mexFunction at :688-807) and
``BiotSavart`` keeps core/BiotSavart.m's 8-input signature (BiotSavart.m:1).
Both take MATLAB-shaped float64 arrays ([ns,3,nt,nb] endpoints and targets,
[ns,1,nt,nb] circulation and squared core radius) and evaluate on the GPU
through the C ABI (winds_bs_eval); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib


def _pad4(shape, what: str):
    """mexFunction's padding of P to 4-D (BiotSavartMex.cu:745-764)."""
    if len(shape) <= 1:
        raise ValueError(f"Need more dimensions in {what}.")
    if len(shape) > 4:
        raise ValueError(f"Wrong dimensions in {what}.")
    return tuple(shape) + (1,) * (4 - len(shape))


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a.size else None


def BiotSavartMex(F1, F2, P, gamma, rc, d, cmod, co, type, gpu="true", gpuhw=(256, 1, 0)):
    """[uind, L] = BiotSavartMex(F1, F2, P, gamma, rc, d, cmod, co, type, gpu, gpuhw).

    cmod "none": cut-off core with MEX semantics (what the CUDA MEX computes);
    "cutm": cut-off with BiotSavart.m semantics; "visc": Vatistas order n=d
    (WInDS.m:84-91 encoding).  type "full"/"fuln"/"fuls": FP64 sum;
    "fulc": FP32 compensated with FP64 near pairs; "leng": lengths only.
    Returns uind with P's dims padded to 4-D and L with F's [ns,1,nt,nb].
    """
    F1 = _f64(F1)
    F2 = _f64(F2)
    P = _f64(P)
    fd = _pad4(F1.shape, "F1")
    if F2.shape != F1.shape:
        raise ValueError("F1 and F2 must have the same dimensions")
    pd = _pad4(P.shape, "P")
    ng = fd[0] * fd[2] * fd[3]
    gamma = _f64(gamma).reshape(-1, order="F")
    rc = _f64(rc).reshape(-1, order="F")
    if gamma.size != ng or rc.size != ng:
        raise ValueError(f"gamma and rc need {ng} elements ([ns,1,nt,nb] of F)")
    uind = np.zeros(pd, dtype=np.float64, order="F")
    L = np.zeros((fd[0], 1, fd[2], fd[3]), dtype=np.float64, order="F")
    hw = (C.c_double * 3)(*[float(x) for x in gpuhw]) if gpuhw is not None else None
    L_ = _lib.lib()
    code = L_.winds_bs_eval(_ptr(F1), _ptr(F2), _lib.dims4(fd), _ptr(P), _lib.dims4(pd),
                            _ptr(gamma), _ptr(rc), float(d), cmod.encode(), float(co),
                            type.encode(), gpu.encode(), hw, _ptr(uind), _ptr(L))
    _lib.check(code)
    return uind, L


def BiotSavart(F1, F2, P, gamma, rc, d, co, type="full", precision="fp64", device=0):
    """[uind, L] = BiotSavart(F1, F2, P, gamma, rc, d, co, type)  (core/BiotSavart.m).

    d numeric: cut-off core; d = 'viscN': Vatistas core of order N
    (BiotSavart.m:49-53).  type 'full' returns L = |F2-F1|^2, 'length'
    returns L = |F2-F1| and uind = 0 (BiotSavart.m:71-76).
    precision 'fp64' (BiotSavart.m's arithmetic) or 'fp32c'.
    """
    if isinstance(d, str):
        if not d.startswith("visc"):
            raise ValueError("string d must be 'viscN'")
        n = float(d[4:])
        cmod, dd = "visc", n
    else:
        cmod, dd = "cutm", float(d)
    if type == "length":
        t = "leng"
    elif type == "full":
        t = "fulc" if precision == "fp32c" else "full"
    else:
        raise ValueError("type must be 'full' or 'length'")
    uind, L = BiotSavartMex(F1, F2, P, gamma, rc, dd, cmod, co, t, "true", (0, 1, device))
    return uind.reshape(_pad4(np.shape(P), "P")), L
