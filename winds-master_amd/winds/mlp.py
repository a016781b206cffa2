"""BF16 MFMA multilayer perceptron (include/winds_mlp.h) over torch tensors.

torch provides device memory and the current stream only; the forward pass,
backward pass and Adam update run in the engine's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _bind(L):
    if getattr(L, "_winds_mlp_bound", False):
        return L
    P = C.c_void_p
    F = C.POINTER(C.c_float)
    L.winds_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_uint64, C.POINTER(C.c_void_p)]
    L.winds_mlp_destroy.argtypes = [P]
    L.winds_mlp_destroy.restype = None
    L.winds_mlp_num_params.argtypes = [P]
    L.winds_mlp_num_params.restype = C.c_int64
    L.winds_mlp_get_params.argtypes = [P, F]
    L.winds_mlp_set_params.argtypes = [P, F]
    L.winds_mlp_forward.argtypes = [P, P, C.c_int32, P, P]
    L.winds_mlp_gradients.argtypes = [P, P, P, C.c_int32, F, F]
    L.winds_mlp_train_step.argtypes = [P, P, P, C.c_int32, C.c_float, P, P]
    L.winds_mlp_set_fused.argtypes = [P, C.c_int32]
    for f in ("create", "get_params", "set_params", "forward", "gradients", "train_step",
              "set_fused"):
        getattr(L, "winds_mlp_" + f).restype = C.c_int
    L._winds_mlp_bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """n_in -> n_layers x n_hidden (SiLU) -> n_out, trained with Adam on the MSE."""

    def __init__(self, n_in: int, n_hidden: int, n_layers: int, n_out: int, max_batch: int,
                 seed: int = 7):
        import torch
        torch.cuda.init()
        self.L = _bind(_lib.lib())
        self.n_in, self.n_hidden, self.n_layers, self.n_out = n_in, n_hidden, n_layers, n_out
        self.max_batch = max_batch
        self._m = C.c_void_p()
        _lib.check(self.L.winds_mlp_create(n_in, n_hidden, n_layers, n_out, max_batch, seed,
                                           C.byref(self._m)))

    def close(self):
        if self._m:
            self.L.winds_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._m

    def layer_shapes(self):
        dims = [self.n_in] + [self.n_hidden] * self.n_layers + [self.n_out]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def num_params(self) -> int:
        return int(self.L.winds_mlp_num_params(self._m))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params(), dtype=np.float32)
        _lib.check(self.L.winds_mlp_get_params(self._m, _fp(out)))
        return out

    def set_params(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.num_params():
            raise ValueError("wrong parameter count")
        _lib.check(self.L.winds_mlp_set_params(self._m, _fp(flat)))

    def unflatten(self, flat: np.ndarray):
        """[(W [n_out, n_in], b [n_out]), ...] views of a flat parameter array."""
        out, o = [], 0
        for n, k in self.layer_shapes():
            W = flat[o:o + n * k].reshape(n, k)
            o += n * k
            b = flat[o:o + n]
            o += n
            out.append((W, b))
        return out

    @staticmethod
    def _check_rows(t, cols, name):
        import torch
        if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != cols or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float32 [B, {cols}] tensor")
        if not t.is_cuda:
            raise ValueError(f"{name} must be device memory")

    def forward(self, X, out=None):
        import torch
        self._check_rows(X, self.n_in, "X")
        B = X.shape[0]
        if out is None:
            out = torch.empty((B, self.n_out), dtype=torch.float32, device=X.device)
        st = torch.cuda.current_stream(X.device).cuda_stream
        _lib.check(self.L.winds_mlp_forward(self._m, C.c_void_p(X.data_ptr()), B,
                                            C.c_void_p(out.data_ptr()), C.c_void_p(st)))
        return out

    def set_fused(self, on: bool) -> None:
        """Forward passes (inference and training) in one kernel per 32 rows
        (True, the default for hidden widths up to 256) or one GEMM per layer
        (False); the same results."""
        _lib.check(self.L.winds_mlp_set_fused(self._m, 1 if on else 0))

    def gradients(self, X, T):
        """(flat FP32 gradient of the MSE, loss) for one batch, without an update."""
        import torch
        self._check_rows(X, self.n_in, "X")
        self._check_rows(T, self.n_out, "T")
        torch.cuda.current_stream(X.device).synchronize()
        g = np.empty(self.num_params(), dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _lib.check(self.L.winds_mlp_gradients(self._m, C.c_void_p(X.data_ptr()),
                                              C.c_void_p(T.data_ptr()), X.shape[0], _fp(g),
                                              _fp(loss)))
        return g, float(loss[0])

    def train_step(self, X, T, lr: float, loss=None):
        import torch
        self._check_rows(X, self.n_in, "X")
        self._check_rows(T, self.n_out, "T")
        st = torch.cuda.current_stream(X.device).cuda_stream
        lp = C.c_void_p(loss.data_ptr()) if loss is not None else None
        _lib.check(self.L.winds_mlp_train_step(self._m, C.c_void_p(X.data_ptr()),
                                               C.c_void_p(T.data_ptr()), X.shape[0],
                                               float(lr), lp, C.c_void_p(st)))
