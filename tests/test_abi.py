"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/winds_bs.h, and rejects bad arguments with the right
status code before touching a device."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available
import winds
from winds import _lib

ROOT = Path(__file__).resolve().parents[1]


def header_functions(name="winds_bs.h"):
    text = (ROOT / "include" / name).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(winds_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = header_functions()
    assert len(names) >= 11
    L = C.CDLL(str(_lib.LIB_PATH))
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == set(names)
    march = header_functions("winds_march.h")
    assert len(march) >= 10
    for n in march:
        assert hasattr(L, n), n
    mlp = header_functions("winds_mlp.h")
    assert len(mlp) == 10
    for n in mlp:
        assert hasattr(L, n), n
    sur = header_functions("winds_surrogate.h")
    assert len(sur) == 3
    for n in sur:
        assert hasattr(L, n), n
    gnd = header_functions("winds_ground.h")
    assert len(gnd) == 6
    for n in gnd:
        assert hasattr(L, n), n
    # every header in include/ is covered above
    assert {p.name for p in (ROOT / "include").glob("*.h")} == {
        "winds_bs.h", "winds_march.h", "winds_mlp.h", "winds_surrogate.h", "winds_ground.h"}


def test_version_and_defaults():
    L = _lib.lib()
    assert b"gfx950" in L.winds_version()
    o = _lib.BsOpts()
    L.winds_bs_opts_default(C.byref(o))
    assert o.core == 0 and o.semantics == 0 and o.precision == 0
    assert o.d == 1e-4 and math.isinf(o.co) and o.near_factor == 40.0 and o.near_len_frac == 0.51
    assert o.guard_kappa == 1.0


def _call(type="fuln", cmod="none", d=1e-4, fd=(2, 3, 2, 1), pd=(2, 3, 2, 1), gpuhw=(256, 1, 0)):
    F = np.zeros(fd, order="F")
    P = np.zeros(pd, order="F")
    g = np.zeros(fd[0] * fd[2] * fd[3])
    u = np.zeros(pd, order="F")
    hw = (C.c_double * 3)(*gpuhw)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    return _lib.lib().winds_bs_eval(p(F), p(F), _lib.dims4(fd), p(P), _lib.dims4(pd), p(g), p(g),
                                    d, cmod.encode(), 1000.0, type.encode(), b"true", hw, p(u),
                                    None)


def test_argument_errors_before_device():
    err = lambda: _lib.lib().winds_last_error().decode()
    assert _call(type="xxxx") == -3 and "unknown type 'xxxx'" in err()
    assert _call(cmod="bogus") == -3 and "unknown cmod 'bogus'" in err()
    assert _call(cmod="visc", d=2.5) == -3 and "integer Vatistas order" in err()
    assert _call(fd=(2, 2, 2, 1)) == -2 and "F1/F2 must be [ns,3,nt,nb]" in err()
    assert _call(pd=(2, 4, 2, 1)) == -2 and "P must be [ns,3,nt,nb]" in err()
    assert _call(gpuhw=(100, 1, 0)) == -1 and "threads must be 0 or a multiple of 64" in err()
    assert _call(gpuhw=(256, 2, 0)) == -1 and "one GPU per call" in err()
    # a failing call leaves its own message, not the previous one
    assert _call(type="yyyy") == -3 and "unknown type 'yyyy'" in err()


def test_device_opts_validation():
    L = _lib.lib()
    o = _lib.opts(core=winds.CORE_VATISTAS, semantics=winds.SEM_MEX)
    assert L.winds_bs_workspace_bytes(10, 10, C.byref(o)) == 0
    o = _lib.opts(precision=winds.PREC_FP32C, guard_kappa=float("nan"))
    assert L.winds_bs_workspace_bytes(10, 10, C.byref(o)) == 0
    # the guard's buffers are part of the FP32C workspace; off, they are not
    on = L.winds_bs_workspace_bytes(1000, 5000, C.byref(_lib.opts(precision=winds.PREC_FP32C)))
    off = L.winds_bs_workspace_bytes(1000, 5000, C.byref(_lib.opts(precision=winds.PREC_FP32C,
                                                                   guard_kappa=-1.0)))
    assert on > off > 0
    o = _lib.opts(precision=winds.PREC_FP32C)
    assert L.winds_bs_workspace_bytes(1000, 5000, C.byref(o)) >= 5024 * 80


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_is_reported_not_faked():
    assert _call() == -6
    assert b"device" in _lib.lib().winds_last_error()
    with pytest.raises(winds.WindsError):
        winds.BiotSavartMex(np.zeros((2, 3, 2, 1)), np.ones((2, 3, 2, 1)), np.zeros((2, 3, 1, 1)),
                            np.ones((2, 1, 2, 1)), np.ones((2, 1, 2, 1)), 1e-4, "none", 1000.0,
                            "fuln", "true", (256, 1, 0))


def test_python_mirror_argument_checks():
    with pytest.raises(ValueError):
        winds.BiotSavartMex(np.zeros(3), np.zeros(3), np.zeros((2, 3)), [0], [0], 1e-4, "none",
                            1.0, "fuln")
    with pytest.raises(ValueError):
        winds.BiotSavart(np.zeros((2, 3, 1, 1)), np.zeros((2, 3, 1, 1)), np.zeros((2, 3)),
                         np.zeros(2), np.zeros(2), "bogus", 1.0, "full")
