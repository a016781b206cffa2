"""The MATLAB/Octave gateway (winds-master_amd/mex/BiotSavartMex.cpp) built
against an in-process mock of the MEX runtime (mex/mock).

Argument checks follow cuda/source/BiotSavartMex.cu:688-807: 11 inputs and
2 outputs, P padded to 4-D, string arguments read with mxGetString.  The GPU
test checks the gateway's uind against the CPU oracle of the MEX arithmetic
(bit-exact in FP64) and L against |F2 - F1|^2."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available
from winds import _lib, gen

MOCK = Path(_lib.LIB_PATH).parent / "libwinds_mexmock.so"


@pytest.fixture(scope="module")
def mex():
    _lib.lib()                          # the engine first (after torch)
    M = C.CDLL(str(MOCK))
    P = C.c_void_p
    M.mock_double.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    M.mock_double.restype = P
    M.mock_string.argtypes = [C.c_char_p]
    M.mock_string.restype = P
    M.mock_ndims.argtypes = [P]
    M.mock_dims.argtypes = [P, C.POINTER(C.c_int64)]
    M.mock_data.argtypes = [P]
    M.mock_data.restype = C.POINTER(C.c_double)
    M.mock_free.argtypes = [P]
    M.mock_call.argtypes = [C.c_int, C.POINTER(P), C.c_int, C.POINTER(P), C.c_char_p, C.c_int]
    return M


class Args:
    def __init__(self, M):
        self.M = M
        self.held = []

    def num(self, a):
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        if a.ndim == 0:
            a = a.reshape(1, 1)
        dims = (C.c_int64 * a.ndim)(*a.shape)
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        p = self.M.mock_double(a.ndim, dims, flat.ctypes.data_as(C.POINTER(C.c_double)))
        self.held.append(p)
        return p

    def str(self, s):
        p = self.M.mock_string(s.encode())
        self.held.append(p)
        return p

    def call(self, ins, nlhs=2):
        arr_in = (C.c_void_p * len(ins))(*ins)
        arr_out = (C.c_void_p * max(nlhs, 1))()
        msg = C.create_string_buffer(512)
        rc = self.M.mock_call(nlhs, arr_out, len(ins), arr_in, msg, 512)
        outs = [arr_out[i] for i in range(nlhs)] if rc == 0 else []
        self.held.extend(o for o in outs if o)
        return rc, msg.value.decode(), outs

    def array(self, p):
        nd = self.M.mock_ndims(p)
        dims = (C.c_int64 * nd)()
        self.M.mock_dims(p, dims)
        shape = tuple(dims)
        n = int(np.prod(shape))
        data = np.ctypeslib.as_array(self.M.mock_data(p), shape=(n,)).copy()
        return data.reshape(shape, order="F")

    def free(self):
        for p in self.held:
            self.M.mock_free(p)
        self.held = []


def small_call(A, w, d=1e-4, cmod="none", typ="fuln", P=None):
    F1, F2, g, rc = w.shed_segments()
    P = w.domain if P is None else P
    return [A.num(F1), A.num(F2), A.num(P), A.num(g), A.num(rc), A.num(d), A.str(cmod),
            A.num(1000.0), A.str(typ), A.str("true"), A.num([256, 1, 0])]


def test_argument_count_and_strings(mex):
    A = Args(mex)
    w = gen.rotor_wake(nb=2, nst=5, ages=6)
    ins = small_call(A, w)
    rc, msg, _ = A.call(ins[:10])
    assert rc == 1 and msg == "11 inputs expected."
    rc, msg, _ = A.call(ins, nlhs=1)
    assert rc == 1 and msg == "2 outputs expected."
    bad = list(ins)
    bad[6] = A.num(3.0)
    rc, msg, _ = A.call(bad)
    assert rc == 1 and msg == "Couldn't make Core Model string."
    bad = list(ins)
    bad[5] = A.str("visc2")
    rc, msg, _ = A.call(bad)
    assert rc == 1 and "d must be numeric" in msg
    bad = list(ins)
    bad[1] = A.num(np.zeros((3, 3, 6, 2)))
    rc, msg, _ = A.call(bad)
    assert rc == 1 and "same dims" in msg
    A.free()


def test_gamma_and_rc_sizes_are_checked(mex):
    """gamma and rc must hold one value per segment (the engine reads that
    many): a short gamma or a scalar rc is an error, not a heap over-read."""
    A = Args(mex)
    w = gen.rotor_wake(nb=2, nst=5, ages=6)
    bad = small_call(A, w)
    bad[3] = A.num(np.ones(3))
    rc, msg, _ = A.call(bad)
    assert rc == 1 and msg == "gamma must hold one value per segment."
    bad = small_call(A, w)
    bad[4] = A.num(0.0)
    rc, msg, _ = A.call(bad)
    assert rc == 1 and msg == "rc must be empty or hold one value per segment."
    A.free()


def test_engine_errors_reach_matlab(mex):
    A = Args(mex)
    w = gen.rotor_wake(nb=2, nst=5, ages=6)
    rc, msg, _ = A.call(small_call(A, w, cmod="bogus"))
    assert rc == 1 and msg.startswith("BiotSavartMex: ") and "cmod" in msg
    if not gpu_available():
        rc, msg, _ = A.call(small_call(A, w))
        assert rc == 1 and msg.startswith("BiotSavartMex: ")
    A.free()


@pytest.mark.gpu
def test_gateway_matches_the_oracle(need_gpu, mex):
    from oracle import bs as O
    A = Args(mex)
    w = gen.rotor_wake(nb=3, nst=9, ages=20)
    F1, F2, g, rc = w.shed_segments()
    rc_, msg, outs = A.call(small_call(A, w))
    assert rc_ == 0, msg
    u, L = A.array(outs[0]), A.array(outs[1])
    ref = O.bs_mex(F1, F2, w.domain, g, d=1e-4)
    assert u.shape == w.domain.shape and np.array_equal(u, ref)
    l2 = np.sum((F2 - F1) ** 2, axis=1, keepdims=True)
    assert L.shape == (F1.shape[0], 1, F1.shape[2], F1.shape[3])
    assert np.allclose(L, l2, rtol=1e-15, atol=0)
    # 2-D targets are padded to [n, 3, 1, 1]
    P2 = np.asfortranarray(w.domain[:, :, 0, 0])
    rc_, msg, outs = A.call(small_call(A, w, P=P2))
    assert rc_ == 0, msg
    u2 = A.array(outs[0])
    assert u2.shape == (P2.shape[0], 3, 1, 1)
    assert np.array_equal(u2[:, :, 0, 0], ref[:, :, 0, 0])
    # an empty rc is accepted for the cut-off core (r_c = 0 everywhere)
    ins = small_call(A, w)
    ins[4] = A.num(np.zeros((0, 0)))
    rc_, msg, outs = A.call(ins)
    assert rc_ == 0, msg
    assert np.array_equal(A.array(outs[0]), ref)
    A.free()
