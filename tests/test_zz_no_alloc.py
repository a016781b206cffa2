"""The no-allocation entry point winds_bs_eval_dev_noalloc against
winds_bs_eval_dev, through the C ABI: two calls in a row on the same GPU,
the second with a smaller input (the scratch array keeps the first call's
contents), in FP64 (bitwise equal) and FP32C (bitwise equal at the same
segment split; within 1e-6 per target of FP64 where the scratch array
forces a smaller split).  Sizes beyond the scratch array are refused before
any launch."""
import ctypes as C

import numpy as np
import pytest

from conftest import flush_print
from errstats import stats
import winds
from winds import _lib, gen, layout

pytestmark = pytest.mark.gpu


def _noalloc(tg, sg, o, out):
    import torch
    st = torch.cuda.current_stream().cuda_stream
    return _lib.lib().winds_bs_eval_dev_noalloc(
        C.c_void_p(tg.data_ptr()), tg.shape[1], C.c_void_p(sg.data_ptr()), sg.shape[0],
        C.byref(o), C.c_void_p(out.data_ptr()), C.c_void_p(st))


def _opts(prec, core):
    if core == "cutoff":
        return winds.opts(precision=prec, core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    return winds.opts(precision=prec, core=winds.CORE_VATISTAS, vat_n=2,
                      semantics=winds.SEM_MATLAB)


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_noalloc_matches_eval_dev(need_gpu, core):
    import torch
    from winds.device import Evaluator
    ev = Evaluator(0)
    big = gen.rotor_wake(nb=3, nst=18, ages=144)
    small = gen.rotor_wake(nb=2, nst=9, ages=30)
    for step, w in enumerate((big, small)):
        tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
        sg = torch.from_numpy(layout.wake_segments(w, chained=True)).cuda().contiguous()
        for prec in (winds.PREC_FP64, winds.PREC_FP32C):
            o = _opts(prec, core)
            flush_print(f"[noalloc {core}] call {step + 1}, precision {prec}, "
                        f"{tg.shape[1]} x {sg.shape[0]}")
            ref = ev.eval(tg, sg, o).clone()
            out = torch.full_like(ref, float("nan"))
            _lib.check(_noalloc(tg, sg, o, out))
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (step, prec)


def test_noalloc_lowers_the_split_to_fit(need_gpu):
    """The 133,920 x 263,427 single wake needs ~290 MB at its tuned split;
    in the 64 MiB scratch array it runs at a smaller split: same bound
    against FP64 as the allocating call."""
    import torch
    from winds.device import Evaluator
    w = gen.config2()
    tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
    sg = torch.from_numpy(layout.wake_segments(w, chained=True)).cuda().contiguous()
    ev = Evaluator(0)
    u64 = ev.eval(tg, sg, _opts(winds.PREC_FP64, "vatistas2")).cpu().numpy()
    out = torch.empty((3, tg.shape[1]), dtype=torch.float64, device="cuda")
    flush_print("[noalloc] config-2 size")
    _lib.check(_noalloc(tg, sg, _opts(winds.PREC_FP32C, "vatistas2"), out))
    torch.cuda.synchronize()
    s = stats(out.cpu().numpy(), u64)
    flush_print(f"   {s}")
    assert s["max_rel"] <= 1e-6 and s["normwise"] <= 1e-7


def test_noalloc_refuses_what_does_not_fit(need_gpu):
    import torch
    o = _opts(winds.PREC_FP32C, "vatistas2")
    o.split = 64                       # explicit split: must fit as it is
    tg = torch.zeros((3, 200_000), dtype=torch.float64, device="cuda")
    sg = torch.zeros((300_000, 8), dtype=torch.float64, device="cuda")
    out = torch.zeros((3, 200_000), dtype=torch.float64, device="cuda")
    assert _noalloc(tg, sg, o, out) == -5
    assert b"no-allocation" in _lib.lib().winds_last_error()
    assert torch.count_nonzero(out) == 0
