"""Error paths forced from the host.

Every device allocation of the engine goes through one wrapper; while the
environment variable WINDS_FAIL_ALLOC_BYTES holds n > 0, requests of at
least n bytes fail with hipErrorOutOfMemory (torch's own allocations are not
affected).  Each entry point must then return an error code with a message,
leave no half-built object behind, and work again, with unchanged results,
once allocations succeed: the host-array call (whose grow-only buffers fail
while growing), the march, the network, the ground plane and the batch.
No kernel is made to fail: the condition is forced on the host."""
import contextlib
import math
import os

import numpy as np
import pytest

import winds
from conftest import flush_print

pytestmark = pytest.mark.gpu

VAR = "WINDS_FAIL_ALLOC_BYTES"


@contextlib.contextmanager
def refused(min_bytes):
    os.environ[VAR] = str(int(min_bytes))
    try:
        yield
    finally:
        del os.environ[VAR]


def wake(n_fil, n_seg, n_t, seed):
    rng = np.random.default_rng(seed)
    F1 = rng.normal(0, 2, (n_seg, 3, n_fil, 1))
    F2 = F1 + rng.normal(0, 0.2, F1.shape)
    P = rng.normal(0, 2, (n_t, 3, 1, 1))
    g = rng.normal(0, 1, n_seg * n_fil)
    rc = np.full(n_seg * n_fil, 1e-3)
    return F1, F2, P, g, rc


def mex(w, type_="full"):
    F1, F2, P, g, rc = w
    return winds.BiotSavartMex(F1, F2, P, g, rc, 1e-4, "none", math.inf, type_)


@pytest.mark.parametrize("type_", ["full", "fulc"])
def test_host_array_call_recovers(need_gpu, type_):
    flush_print(f"[failures] host-array call {type_}")
    small = wake(3, 40, 50, 1)
    big = wake(20, 400, 3000, 2)
    ref_small, _ = mex(small, type_)
    with refused(1 << 16):               # the small call's buffers exist; growing fails
        with pytest.raises(winds.WindsError) as ei:
            mex(big, type_)
    assert ei.value.code == -4 and "out of memory" in str(ei.value).lower()
    u_big, _ = mex(big, type_)           # allocations succeed again
    u_small, _ = mex(small, type_)
    assert np.array_equal(u_small, ref_small)
    u_big2, _ = mex(big, type_)
    assert np.array_equal(u_big, u_big2)


def test_march_create_recovers(need_gpu):
    from winds import cases
    from winds.march import March
    flush_print("[failures] march")
    tb = cases.tables(cases.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0,
                                            dpsi_deg=10.0, n_times=20))
    with refused(1):
        with pytest.raises(winds.WindsError, match="(?i)out of memory"):
            March(tb, 9, winds.opts(d=1e-4))
    m = March(tb, 9, winds.opts(d=1e-4))
    m.step(5)
    m.sync()
    assert np.all(np.isfinite(m.perf()))


def test_network_and_ground_create_recover(need_gpu):
    from winds.ground import Ground
    from winds.mlp import MLP
    flush_print("[failures] network, ground plane")
    with refused(1):
        with pytest.raises(winds.WindsError, match="(?i)out of memory"):
            MLP(49, 256, 4, 3, 1024, 1)
        with pytest.raises(winds.WindsError, match="(?i)out of memory"):
            Ground(20, -500.0, 1500.0, -375.0, 375.0)
    net = MLP(49, 256, 4, 3, 1024, 1)
    assert net.num_params() == 49 * 256 + 256 + 3 * (256 * 256 + 256) + 256 * 3 + 3
    g = Ground(20, -500.0, 1500.0, -375.0, 375.0)
    assert g.n == 400


def test_batch_create_recovers(need_gpu):
    from winds.batch import Batch
    flush_print("[failures] batch")
    rng = np.random.default_rng(4)
    wakes = []
    for k in range(3):
        t = rng.normal(0, 2, (3, 200 + 50 * k))
        a = rng.normal(0, 2, (300, 3))
        s = np.concatenate([a, a + rng.normal(0, 0.2, (300, 3)), rng.normal(0, 1, (300, 1)),
                            np.full((300, 1), 1e-3)], axis=1)
        wakes.append((t, s))
    o = winds.opts(precision=winds.PREC_FP64, d=1e-4)
    with refused(1):
        with pytest.raises(winds.WindsError, match="(?i)out of memory"):
            Batch(wakes, o)
    b = Batch(wakes, o)
    u = b.eval(use_graph=False)
    assert bool(u.isfinite().all())


def test_noalloc_entry_point_needs_no_allocation(need_gpu):
    """winds_bs_eval_dev_noalloc runs, with the same result as
    winds_bs_eval_dev, while every device allocation of the engine fails."""
    import ctypes as C

    import torch
    from winds import _lib, gen, layout
    from winds.device import Evaluator
    flush_print("[failures] no-allocation entry point under refused allocations")
    w = gen.rotor_wake(nb=2, nst=9, ages=40)
    tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
    sg = torch.from_numpy(layout.wake_segments(w, chained=True)).cuda().contiguous()
    for prec in (winds.PREC_FP64, winds.PREC_FP32C):
        o = winds.opts(precision=prec, core=winds.CORE_VATISTAS, vat_n=2,
                       semantics=winds.SEM_MATLAB)
        ref = Evaluator(0).eval(tg, sg, o).clone()
        out = torch.full_like(ref, float("nan"))
        st = torch.cuda.current_stream().cuda_stream
        with refused(1):
            _lib.check(_lib.lib().winds_bs_eval_dev_noalloc(
                C.c_void_p(tg.data_ptr()), tg.shape[1], C.c_void_p(sg.data_ptr()), sg.shape[0],
                C.byref(o), C.c_void_p(out.data_ptr()), C.c_void_p(st)))
            torch.cuda.synchronize()
        assert torch.equal(out, ref), prec
