"""Batched evaluation of uneven wakes: one persistent launch, hipGraph replay,
autotuned launch shape.  Every wake of the batch must equal its own
evaluation: within 1e-6 per target of the FP64 engine, and bit for bit
between eager runs and graph replays."""
import numpy as np
import pytest

from conftest import flush_print
from errstats import stats
import winds
from winds import gen, layout

pytestmark = pytest.mark.gpu


def small_batch_wakes():
    ps = gen.config3_params(n_wakes=6, seed0=100)
    out = []
    for i, p in enumerate(ps):
        p = dict(p)
        p["ages"] = 30 + (p["ages"] % 90)          # keep the FP64 check quick
        p["nst"] = 12 + (p["nst"] % 9)
        out.append(gen.rotor_wake(**p))
    return out


def vat2(prec):
    return winds.opts(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB,
                      precision=prec)


@pytest.fixture(scope="module")
def batch(need_gpu):
    from winds.batch import Batch
    wakes = small_batch_wakes()
    return wakes, Batch(wakes, vat2(winds.PREC_FP32C), tune={"block": 128, "chunk": 1024})


def test_batch_matches_each_wake(batch):
    import torch
    from winds.device import Evaluator
    wakes, b = batch
    flush_print(f"[batch] {b.info()}")
    u = b.eval(use_graph=False).clone()
    ev = Evaluator(0)
    for w, uw in zip(wakes, b.split(u)):
        tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
        sg = torch.from_numpy(layout.wake_segments(w)).cuda().contiguous()
        ref = ev.eval(tg, sg, vat2(winds.PREC_FP64)).cpu().numpy()
        s = stats(uw.cpu().numpy(), ref)
        flush_print(f"   wake nt={w.n_targets}: {s}")
        assert s["max_rel"] <= 1e-6 and s["normwise"] <= 1e-7


def test_graph_replays_equal_eager(batch):
    import torch
    wakes, b = batch
    eager = b.eval(use_graph=False).clone()
    for i in range(3):
        g = b.eval(use_graph=True).clone()
        torch.cuda.synchronize()
        assert torch.equal(g, eager), f"replay {i}"
        # eager work between replays, on the same stream
        b.eval(use_graph=False)
        torch.cuda.synchronize()
        assert torch.equal(b.out, eager)


def test_autotune_keeps_results(need_gpu):
    from winds.batch import Batch
    wakes = small_batch_wakes()[:3]
    b = Batch(wakes, vat2(winds.PREC_FP32C))
    ref = b.eval(use_graph=False).clone()
    tune, ms = b.autotune()
    flush_print(f"[autotune] {tune} {ms:.3f} ms")
    assert tune["block"] in (64, 128, 256) and tune["chunk"] % 16 == 0 and ms > 0
    u = b.eval(use_graph=True).clone()
    # a new launch shape changes only the FP32 rounding of partial sums
    assert float((u - ref).abs().max()) <= 1e-7 * float(ref.abs().max())


@pytest.mark.parametrize("core", ["vatistas3", "vatistas1_co", "cutoff_mex"])
def test_batch_other_cores(need_gpu, core):
    """The persistent kernel's other instances (general Vatistas order,
    Vatistas n = 1 with the far cut-off, cut-off core with MEX semantics)
    against each wake's own FP64 evaluation."""
    import torch
    from winds.batch import Batch
    from winds.device import Evaluator
    if core == "vatistas3":
        kw = dict(core=winds.CORE_VATISTAS, vat_n=3, semantics=winds.SEM_MATLAB)
    elif core == "vatistas1_co":
        kw = dict(core=winds.CORE_VATISTAS, vat_n=1, semantics=winds.SEM_MATLAB, co=40.0)
    else:
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    wakes = small_batch_wakes()[:4]
    b = Batch(wakes, winds.opts(precision=winds.PREC_FP32C, **kw),
              tune={"block": 256, "chunk": 2048})
    u = b.eval(use_graph=True).clone()
    ev = Evaluator(0)
    for w, uw in zip(wakes, b.split(u)):
        tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
        sg = torch.from_numpy(layout.wake_segments(w)).cuda().contiguous()
        ref = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
        s = stats(uw.cpu().numpy(), ref)
        flush_print(f"   [{core}] wake nt={w.n_targets}: {s}")
        assert s["max_rel"] <= 1e-6 and s["normwise"] <= 1e-7


def test_empty_and_tiny_wakes(need_gpu):
    import torch
    from winds.batch import Batch
    w = gen.rotor_wake(nb=2, nst=5, ages=4)
    t = layout.wake_targets(w)
    s = layout.wake_segments(w)
    items = [(t, s), (t[:, :0], s), (t, s[:0]), (t[:, :1], s[:3])]
    b = Batch(items, winds.opts(precision=winds.PREC_FP32C, d=1e-4))
    u = b.eval(use_graph=True).clone()
    parts = b.split(u)
    assert parts[1].shape[1] == 0
    assert torch.count_nonzero(parts[2]) == 0
    assert torch.allclose(parts[0][:, :1], parts[0][:, :1]) and torch.isfinite(u).all()
