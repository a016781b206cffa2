"""ctypes binding of libwinds_bs.so (include/winds_bs.h).

The product path has no CPU fallback: if the HIP library is missing, loading
raises.  Build it with __graft_entry__.build() or winds/build.py.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parents[1] / "lib" / "libwinds_bs.so"

CORE_CUTOFF, CORE_VATISTAS = 0, 1
SEM_MEX, SEM_MATLAB = 0, 1
PREC_FP64, PREC_FP32C = 0, 1

EXPORTS = (
    "winds_bs_opts_default", "winds_last_error", "winds_version", "winds_bs_eval",
    "winds_bs_workspace_bytes", "winds_bs_eval_dev", "winds_pack_segments_dev",
    "winds_pack_targets_dev", "winds_unpack_uind_dev", "winds_segment_length_dev",
    "winds_device_count", "winds_batch_create", "winds_batch_destroy",
    "winds_batch_workspace_bytes", "winds_batch_eval_dev", "winds_batch_autotune",
    "winds_batch_info", "winds_bs_guard_count", "winds_batch_guard_count",
    "winds_bs_eval_dev_noalloc",
)


class WindsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"winds error {code}: {msg}")
        self.code = code


class BsOpts(C.Structure):
    _fields_ = [
        ("core", C.c_int32), ("vat_n", C.c_int32), ("semantics", C.c_int32),
        ("precision", C.c_int32), ("d", C.c_double), ("co", C.c_double),
        ("near_factor", C.c_double), ("near_len_frac", C.c_double),
        ("block", C.c_int32), ("split", C.c_int32), ("seg_col", C.c_int32),
        ("seg_blk", C.c_int32), ("guard_kappa", C.c_double),
    ]


_lib = None

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64x4 = C.POINTER(C.c_int64)


def lib():
    """Load the engine library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} not found: the HIP engine is not built "
                           "(run __graft_entry__.build()); there is no CPU fallback")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loading torch first makes the engine bind to that copy, so the process
    # holds one HIP runtime and torch's streams and allocations are valid in
    # engine calls; loading the engine first would hand torch a runtime it was
    # not built against.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    L.winds_bs_opts_default.argtypes = [C.POINTER(BsOpts)]
    L.winds_bs_opts_default.restype = None
    L.winds_last_error.restype = C.c_char_p
    L.winds_version.restype = C.c_char_p
    L.winds_device_count.restype = C.c_int
    L.winds_bs_eval.argtypes = [_P, _P, _I64x4, _P, _I64x4, _P, _P, C.c_double, C.c_char_p,
                                C.c_double, C.c_char_p, C.c_char_p, _D, _P, _P]
    L.winds_bs_eval.restype = C.c_int
    L.winds_bs_workspace_bytes.argtypes = [C.c_int64, C.c_int64, C.POINTER(BsOpts)]
    L.winds_bs_workspace_bytes.restype = C.c_size_t
    L.winds_bs_eval_dev.argtypes = [_P, C.c_int64, _P, C.c_int64, C.POINTER(BsOpts), _P, _P,
                                    C.c_size_t, _P]
    L.winds_bs_eval_dev.restype = C.c_int
    L.winds_pack_segments_dev.argtypes = [_P, _P, _I64x4, _P, _P, _P, _P]
    L.winds_pack_segments_dev.restype = C.c_int
    L.winds_pack_targets_dev.argtypes = [_P, _I64x4, _P, _P]
    L.winds_pack_targets_dev.restype = C.c_int
    L.winds_unpack_uind_dev.argtypes = [_P, _I64x4, _P, _P]
    L.winds_unpack_uind_dev.restype = C.c_int
    L.winds_segment_length_dev.argtypes = [_P, C.c_int64, C.c_int, _P, _P]
    L.winds_segment_length_dev.restype = C.c_int
    L.winds_bs_guard_count.argtypes = [_P, C.c_int64, C.c_int64, C.POINTER(BsOpts),
                                       C.POINTER(C.c_int64), _P]
    L.winds_bs_guard_count.restype = C.c_int
    L.winds_bs_eval_dev_noalloc.argtypes = [_P, C.c_int64, _P, C.c_int64, C.POINTER(BsOpts), _P,
                                            _P]
    L.winds_bs_eval_dev_noalloc.restype = C.c_int
    _lib = L
    return L


def check(code: int) -> None:
    if code != 0:
        raise WindsError(code, lib().winds_last_error().decode())


def opts(core: int = CORE_CUTOFF, vat_n: int = 2, semantics: int = SEM_MEX,
         precision: int = PREC_FP64, d: float = 1e-4, co: float = math.inf,
         near_factor: float = 40.0, near_len_frac: float = 0.51, block: int = 0,
         split: int = 0, seg_col: int = 0, seg_blk: int = 0,
         guard_kappa: float = 0.0) -> BsOpts:
    o = BsOpts()
    lib().winds_bs_opts_default(C.byref(o))
    o.core, o.vat_n, o.semantics, o.precision = core, vat_n, semantics, precision
    o.d, o.co, o.near_factor, o.near_len_frac = d, co, near_factor, near_len_frac
    o.block, o.split, o.seg_col, o.seg_blk = block, split, seg_col, seg_blk
    o.guard_kappa = guard_kappa
    return o


def dims4(a) -> "C.Array":
    return (C.c_int64 * 4)(*[int(x) for x in a])
