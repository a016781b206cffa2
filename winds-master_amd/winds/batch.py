"""Many independent wakes of uneven size, evaluated in one persistent launch.

Wraps the winds_batch_* C ABI: the wakes' targets and segments are
concatenated on the device, a plan (work items, persistent grid) is built
once, and each evaluation is optionally replayed from a hipGraph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, layout


class BatchTune(C.Structure):
    _fields_ = [("block", C.c_int32), ("chunk", C.c_int32), ("wg_per_cu", C.c_int32),
                ("reserved", C.c_int32)]

    def as_dict(self):
        return {"block": self.block, "chunk": self.chunk, "wg_per_cu": self.wg_per_cu}


def _bind(L):
    if getattr(L, "_winds_batch_bound", False):
        return L
    P = C.c_void_p
    L.winds_batch_create.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(_lib.BsOpts), C.POINTER(BatchTune),
                                     C.POINTER(C.c_void_p)]
    L.winds_batch_create.restype = C.c_int
    L.winds_batch_destroy.argtypes = [P]
    L.winds_batch_destroy.restype = None
    L.winds_batch_workspace_bytes.argtypes = [P]
    L.winds_batch_workspace_bytes.restype = C.c_size_t
    L.winds_batch_eval_dev.argtypes = [P, P, P, P, P, C.c_size_t, P, C.c_int]
    L.winds_batch_eval_dev.restype = C.c_int
    L.winds_batch_autotune.argtypes = [P, P, P, P, P, C.POINTER(BatchTune),
                                       C.POINTER(C.c_double)]
    L.winds_batch_autotune.restype = C.c_int
    L.winds_batch_info.argtypes = [P, C.POINTER(C.c_int64)]
    L.winds_batch_info.restype = C.c_int
    L.winds_batch_guard_count.argtypes = [P, P, C.POINTER(C.c_int64), P]
    L.winds_batch_guard_count.restype = C.c_int
    L._winds_batch_bound = True
    return L


class Batch:
    """A set of wakes (RotorWake objects, or (targets[3,n], segments[m,8])
    numpy pairs) resident on one GPU, with an evaluation plan."""

    def __init__(self, wakes, o: _lib.BsOpts, tune: dict | None = None, device: int = 0):
        self.L = _bind(_lib.lib())
        self.device = torch.device("cuda", device)
        tgs, sgs = [], []
        for w in wakes:
            if isinstance(w, tuple):
                t, s = w
            else:
                t, s = layout.wake_targets(w), layout.wake_segments(w, chained=True)
            tgs.append(np.asarray(t, dtype=np.float64))
            sgs.append(np.asarray(s, dtype=np.float64))
        self.nt = np.array([t.shape[1] for t in tgs], dtype=np.int64)
        self.ns = np.array([s.shape[0] for s in sgs], dtype=np.int64)
        self.targets = torch.from_numpy(np.concatenate(tgs, axis=1) if tgs else
                                        np.zeros((3, 0))).to(self.device).contiguous()
        self.segments = torch.from_numpy(np.concatenate(sgs) if sgs else
                                         np.zeros((0, 8))).to(self.device).contiguous()
        self.opts = o
        self._plan = C.c_void_p()
        tp = None
        if tune is not None:
            tp = BatchTune(int(tune["block"]), int(tune["chunk"]), int(tune.get("wg_per_cu", 0)), 0)
        torch.cuda.set_device(self.device)
        code = self.L.winds_batch_create(
            len(self.nt), self.nt.ctypes.data_as(C.POINTER(C.c_int64)),
            self.ns.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(o),
            C.byref(tp) if tp is not None else None, C.byref(self._plan))
        _lib.check(code)
        self.out = torch.empty((3, int(self.nt.sum())), dtype=torch.float64, device=self.device)
        self._ws = None

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self.L.winds_batch_destroy(plan)
                plan.value = None
            except Exception:       # interpreter shutdown
                pass

    def info(self) -> dict:
        a = (C.c_int64 * 4)()
        _lib.check(self.L.winds_batch_info(self._plan, a))
        return {"targets": a[0], "segments": a[1], "interactions": a[2], "items": a[3]}

    def _workspace(self) -> torch.Tensor:
        need = int(self.L.winds_batch_workspace_bytes(self._plan))
        if self._ws is None or self._ws.numel() < need:
            self._ws = None
            self._ws = torch.empty(need + 256, dtype=torch.uint8, device=self.device)
        return self._ws

    def eval(self, use_graph: bool = True, stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        ws = self._workspace()
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        code = self.L.winds_batch_eval_dev(
            self._plan, C.c_void_p(self.targets.data_ptr()), C.c_void_p(self.segments.data_ptr()),
            C.c_void_p(self.out.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel(),
            C.c_void_p(st), 1 if use_graph else 0)
        _lib.check(code)
        return self.out

    def autotune(self, stream: torch.cuda.Stream | None = None):
        """Time every candidate launch shape on this batch; keep the fastest."""
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        best = BatchTune()
        ms = C.c_double()
        self._ws = None  # the tuner allocates its own workspace per candidate
        code = self.L.winds_batch_autotune(
            self._plan, C.c_void_p(self.targets.data_ptr()), C.c_void_p(self.segments.data_ptr()),
            C.c_void_p(self.out.data_ptr()), C.c_void_p(st), C.byref(best), C.byref(ms))
        _lib.check(code)
        return best.as_dict(), float(ms.value)

    def guard_count(self, stream: torch.cuda.Stream | None = None) -> int:
        """Targets the FP32C guard re-evaluated in FP64 in the last evaluation."""
        if self._ws is None:
            return 0
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        n = C.c_int64()
        _lib.check(self.L.winds_batch_guard_count(self._plan, C.c_void_p(self._ws.data_ptr()),
                                                  C.byref(n), C.c_void_p(st)))
        return int(n.value)

    def split(self, u: torch.Tensor | None = None):
        """Per-wake [3, nt] views of a batch result."""
        u = self.out if u is None else u
        offs = np.concatenate([[0], np.cumsum(self.nt)])
        return [u[:, offs[i]:offs[i + 1]] for i in range(len(self.nt))]
