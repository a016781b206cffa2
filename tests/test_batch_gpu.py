"""Batched evaluation of uneven wakes: one persistent launch, hipGraph replay,
autotuned launch shape.  Every wake of the batch must equal its own
evaluation: within 1e-6 per target of the FP64 engine, and bit for bit
between eager runs and graph replays."""
import math

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
    """A full small wake, a wake without targets, a wake without segments and
    a 1-target x 3-segment wake in one batch: the first and the last against
    their own FP64 evaluations, the one without segments exactly 0."""
    import torch
    from winds.batch import Batch
    from winds.device import Evaluator
    w = gen.rotor_wake(nb=2, nst=5, ages=4)
    t = layout.wake_targets(w)
    s = layout.wake_segments(w)
    items = [(t, s), (t[:, :0], s), (t, s[:0]), (t[:, :1], s[:3])]
    o32 = winds.opts(precision=winds.PREC_FP32C, d=1e-4)
    b = Batch(items, o32)
    u = b.eval(use_graph=True).clone()
    parts = b.split(u)
    assert parts[1].shape[1] == 0
    assert torch.count_nonzero(parts[2]) == 0
    assert torch.isfinite(u).all()
    ev = Evaluator(0)
    for i in (0, 3):
        tg = torch.from_numpy(np.ascontiguousarray(items[i][0])).cuda()
        sg = torch.from_numpy(np.ascontiguousarray(items[i][1])).cuda()
        ref = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, d=1e-4)).cpu().numpy()
        st = stats(parts[i].cpu().numpy(), ref)
        flush_print(f"   tiny item {i}: {st}")
        assert st["max_rel"] <= 1e-6 and st["normwise"] <= 1e-7
    # FP64 batch: every item bitwise equal to its own FP64 evaluation
    b64 = Batch(items, winds.opts(precision=winds.PREC_FP64, d=1e-4))
    p64 = b64.split(b64.eval(use_graph=True).clone())
    assert torch.count_nonzero(p64[2]) == 0 and p64[1].shape[1] == 0
    for i in (0, 3):
        tg = torch.from_numpy(np.ascontiguousarray(items[i][0])).cuda()
        sg = torch.from_numpy(np.ascontiguousarray(items[i][1])).cuda()
        ref = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, d=1e-4))
        assert torch.equal(p64[i], ref), i


# --------------------------------------------------------------------------
# FP64 mode of the batch
# --------------------------------------------------------------------------

@pytest.mark.parametrize("sem", ["mex", "vatistas2", "vatistas3_co"])
def test_fp64_batch_bitexact(need_gpu, sem):
    """FP64 items sum each target over its wake's whole segment range in
    segment order, so every wake of the batch is bitwise equal to its own
    single-wake FP64 evaluation; MEX semantics are also bitwise equal to the
    CPU oracle of BiotSavart_nogpu (BiotSavartMex.cu:268-282), BiotSavart.m
    semantics within 1e-12 of the oracle of BiotSavart.m:86-115.  Graph
    replays equal the eager evaluation."""
    import torch
    from oracle import bs as O
    from winds.batch import Batch
    from winds.device import Evaluator
    if sem == "mex":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    elif sem == "vatistas2":
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=3, semantics=winds.SEM_MATLAB, co=60.0)
    wakes = small_batch_wakes()[:4]
    o64 = winds.opts(precision=winds.PREC_FP64, **kw)
    b = Batch(wakes, o64, tune={"block": 128, "chunk": 1024})
    eager = b.eval(use_graph=False).clone()
    for _ in range(2):
        assert torch.equal(b.eval(use_graph=True).clone(), eager)
        torch.cuda.synchronize()
    ev = Evaluator(0)
    for j, (w, uw) in enumerate(zip(wakes, b.split(eager))):
        recs = layout.wake_segments(w, chained=True)
        tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
        sg = torch.from_numpy(recs).cuda().contiguous()
        ref = ev.eval(tg, sg, o64)
        assert torch.equal(uw, ref), (sem, j)
        idx = np.random.default_rng(j).choice(w.n_targets, size=16, replace=False)
        P = np.asfortranarray(layout.wake_targets(w)[:, idx].T.reshape(-1, 3, 1, 1))
        F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
        F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
        got = uw.cpu().numpy()[:, idx].T
        if sem == "mex":
            oref = O.bs_mex(F1, F2, P, recs[:, 6], 1e-4)[:, :, 0, 0]
            assert np.array_equal(got, oref), (sem, j)
        else:
            n = kw["vat_n"]
            oref = O.bs_matlab(F1, F2, P, recs[:, 6], recs[:, 7], f"visc{n}",
                               co=kw.get("co", math.inf))[:, :, 0, 0]
            st = stats(got.T, oref.T)
            assert st["max_rel"] <= 1e-12, (sem, j, st)
    flush_print(f"[fp64 batch {sem}] {b.info()} bitwise equal per wake")


# --------------------------------------------------------------------------
# the 64-wake batch at full size
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def config3_full(need_gpu):
    from winds.batch import Batch  # noqa: F401
    flush_print("[config3 full] generating 64 wakes")
    return gen.config3()


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_config3_full_size_per_target(config3_full, core):
    """All 64 wakes of the uneven batch at full size (200-2,000 ages, 12-40
    stations; 1.22 T interactions) in one FP32C persistent launch, each wake
    against its own FP64 evaluation (the FP64 batch, bitwise equal to it):
    per-target relative error <= 1e-6 for every target with |u| >= 1e-6 of
    its wake's largest."""
    import torch
    from winds.batch import Batch
    wakes = config3_full
    if core == "cutoff":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    flush_print(f"[config3 full {core}] FP32C batch")
    b32 = Batch(wakes, winds.opts(precision=winds.PREC_FP32C, **kw))
    u32 = b32.eval(use_graph=True).clone()
    torch.cuda.synchronize()
    n_fix = b32.guard_count()
    info = b32.info()
    flush_print(f"   {info}, guard re-evaluated {n_fix} targets")
    del b32
    flush_print(f"[config3 full {core}] FP64 batch")
    b64 = Batch(wakes, winds.opts(precision=winds.PREC_FP64, **kw))
    u64 = b64.eval(use_graph=False).clone()
    torch.cuda.synchronize()
    worst = {"max_rel": 0.0, "rms_rel": 0.0, "normwise": 0.0}
    for i, (a, r) in enumerate(zip(b64.split(u32), b64.split(u64))):
        st = stats(a.cpu().numpy(), r.cpu().numpy())
        for k in worst:
            worst[k] = max(worst[k], st[k])
        assert st["max_rel"] <= 1e-6 and st["rms_rel"] <= 1e-7 and st["normwise"] <= 1e-7, (i, st)
    flush_print(f"   worst over 64 wakes: {worst}")
    assert info["interactions"] > 1.2e12 and n_fix <= 2e-3 * info["targets"]
