"""Device-resident evaluation over torch CUDA tensors.

torch only provides device memory and the current HIP stream here; all
arithmetic is the engine's (winds_bs_eval_dev).  Targets are float64 [3, n]
and segments float64 [n, 8] (ax ay az bx by bz gamma rc), both contiguous on
the same GPU.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


class Evaluator:
    """Holds a grow-only workspace tensor for repeated device evaluations."""

    def __init__(self, device: int | str | torch.device = 0):
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        self._ws = None
        _lib.lib()

    def workspace(self, nt: int, nseg: int, o: _lib.BsOpts) -> torch.Tensor:
        need = int(_lib.lib().winds_bs_workspace_bytes(nt, nseg, C.byref(o)))
        if need == 0:
            raise _lib.WindsError(-1, "invalid options")
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need + 256, dtype=torch.uint8, device=self.device)
        return self._ws

    def eval(self, targets: torch.Tensor, segs: torch.Tensor, o: _lib.BsOpts,
             out: torch.Tensor | None = None, stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        if targets.dtype != torch.float64 or segs.dtype != torch.float64:
            raise TypeError("targets and segments must be float64")
        if targets.dim() != 2 or targets.shape[0] != 3 or not targets.is_contiguous():
            raise ValueError("targets must be a contiguous [3, n] tensor")
        if segs.dim() != 2 or segs.shape[1] != 8 or not segs.is_contiguous():
            raise ValueError("segments must be a contiguous [n, 8] tensor")
        if targets.device != self.device or segs.device != self.device:
            raise ValueError("tensors must live on the evaluator's device")
        nt, ns = targets.shape[1], segs.shape[0]
        if out is None:
            out = torch.empty((3, nt), dtype=torch.float64, device=self.device)
        elif out.shape != (3, nt) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 [3, n] tensor")
        ws = self.workspace(nt, ns, o)
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        self._last = (nt, ns, o)
        code = _lib.lib().winds_bs_eval_dev(
            C.c_void_p(targets.data_ptr()), nt, C.c_void_p(segs.data_ptr()), ns, C.byref(o),
            C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel(), C.c_void_p(st))
        _lib.check(code)
        return out

    def guard_count(self, stream: torch.cuda.Stream | None = None) -> int:
        """Targets the FP32C guard re-evaluated in FP64 in the last eval()."""
        last = getattr(self, "_last", None)
        if last is None or self._ws is None:
            return 0
        nt, ns, o = last
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        n = C.c_int64()
        _lib.check(_lib.lib().winds_bs_guard_count(C.c_void_p(self._ws.data_ptr()), nt, ns,
                                                   C.byref(o), C.byref(n), C.c_void_p(st)))
        return int(n.value)
