"""ctypes wrapper of the CPU oracle (oracle/bs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the engine never imports it.
Arrays are MATLAB-shaped float64 ([ns,3,nt,nb] endpoints/targets,
[ns,1,nt,nb] circulation and squared core radius), Fortran order.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None
_P = C.c_void_p
_I = C.POINTER(C.c_int64)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.oracle_bs_mex.argtypes = [_P, _P, _I, _P, _I, _P, C.c_double, _P, C.c_int]
        L.oracle_bs_mex_range.argtypes = [_P, _P, _I, _P, _I, _P, C.c_double, C.c_int64,
                                          C.c_int64, _P, C.c_int]
        L.oracle_bs_matlab.argtypes = [_P, _P, _I, _P, _I, _P, _P, C.c_double, C.c_int,
                                       C.c_double, _P, C.c_int]
        L.oracle_bs_quad.argtypes = [_P, _P, _I, _P, _I, _P, _P, C.c_double, C.c_int,
                                     C.c_double, _P]
        L.oracle_filament_length.argtypes = [_P, _P, _I, _P]
        for f in (L.oracle_bs_mex, L.oracle_bs_mex_range, L.oracle_bs_matlab, L.oracle_bs_quad,
                  L.oracle_filament_length):
            f.restype = None
        _lib = L
    return _lib


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _pad(shape):
    return tuple(shape) + (1,) * (4 - len(shape))


def _d4(s):
    return (C.c_int64 * 4)(*[int(x) for x in s])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def default_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def bs_mex(F1, F2, P, gamma, d, threads: int | None = None) -> np.ndarray:
    """BiotSavart_nogpu (BiotSavartMex.cu:268-282) on MATLAB-shaped arrays."""
    F1, F2, P, gamma = _f(F1), _f(F2), _f(P), _f(gamma)
    fd, pd = _pad(F1.shape), _pad(P.shape)
    out = np.zeros(pd, order="F")
    lib().oracle_bs_mex(_p(F1), _p(F2), _d4(fd), _p(P), _d4(pd), _p(gamma), float(d), _p(out),
                        threads or default_threads())
    return out


def bs_mex_range(F1, F2, P, gamma, d, j0: int, j1: int, threads: int | None = None) -> np.ndarray:
    """Targets [j0, j1) of bs_mex, as an [n, 3] array in packed target order."""
    F1, F2, P, gamma = _f(F1), _f(F2), _f(P), _f(gamma)
    fd, pd = _pad(F1.shape), _pad(P.shape)
    out = np.zeros((j1 - j0, 3))
    lib().oracle_bs_mex_range(_p(F1), _p(F2), _d4(fd), _p(P), _d4(pd), _p(gamma), float(d),
                              j0, j1, _p(out), threads or default_threads())
    return out


def _n_of(d):
    if isinstance(d, str):
        return int(float(d[4:])), 0.0
    return 0, float(d)


def bs_matlab(F1, F2, P, gamma, rc, d, co=math.inf, threads: int | None = None) -> np.ndarray:
    """core/BiotSavart.m 'full' (cut-off for numeric d, Vatistas for 'viscN')."""
    F1, F2, P, gamma, rc = _f(F1), _f(F2), _f(P), _f(gamma), _f(rc)
    fd, pd = _pad(F1.shape), _pad(P.shape)
    n, dd = _n_of(d)
    out = np.zeros(pd, order="F")
    lib().oracle_bs_matlab(_p(F1), _p(F2), _d4(fd), _p(P), _d4(pd), _p(gamma), _p(rc), dd, n,
                           float(co), _p(out), threads or default_threads())
    return out


def bs_quad(F1, F2, P, gamma, rc, d, co=math.inf) -> np.ndarray:
    """BiotSavart.m arithmetic in __float128, rounded to double at the end."""
    F1, F2, P, gamma, rc = _f(F1), _f(F2), _f(P), _f(gamma), _f(rc)
    fd, pd = _pad(F1.shape), _pad(P.shape)
    n, dd = _n_of(d)
    out = np.zeros(pd, order="F")
    lib().oracle_bs_quad(_p(F1), _p(F2), _d4(fd), _p(P), _d4(pd), _p(gamma), _p(rc), dd, n,
                         float(co), _p(out))
    return out


def filament_length(F1, F2) -> np.ndarray:
    """core/FilamentLength.m: |F2 - F1|, shape [ns,1,nt,nb]."""
    F1, F2 = _f(F1), _f(F2)
    fd = _pad(F1.shape)
    out = np.zeros((fd[0], 1, fd[2], fd[3]), order="F")
    lib().oracle_filament_length(_p(F1), _p(F2), _d4(fd), _p(out))
    return out
