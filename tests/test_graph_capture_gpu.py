"""The device evaluation (winds_bs_eval_dev) recorded into a HIP graph.

A caller can record the evaluation, with its workspace allocated beforehand,
into a graph (here torch.cuda.graph, which captures the launches the engine
enqueues on the capturing stream) and replay it: the engine enqueues only
kernel launches in that call (no allocation, memset or synchronisation).
Each replay reads the current inputs: between replays the test changes the
targets and the circulation with torch kernels, evaluates eagerly for the
reference and synchronises; every replay must equal the eager evaluation of
the same inputs bit for bit (the kernels and their order are the same), for
the FP64 engine and for the FP32 compensated one with its FP64 guard."""
import numpy as np
import pytest

import winds
from conftest import flush_print

pytestmark = pytest.mark.gpu


def wake(seed, n_fil=12, n_seg=150, n_t=900):
    """Chained filaments (a random walk each) and targets around them."""
    import torch
    rng = np.random.default_rng(seed)
    segs = []
    for f in range(n_fil):
        p = rng.normal(0, 3, 3)
        for k in range(n_seg):
            q = p + rng.normal(0, 0.3, 3) + np.array([0.4, 0.0, 0.0])
            segs.append([*p, *q, rng.normal(0, 2), 0.05])
            p = q
    S = torch.tensor(np.array(segs), dtype=torch.float64, device="cuda").contiguous()
    T = torch.tensor(rng.normal(0, 4, (3, n_t)), dtype=torch.float64, device="cuda").contiguous()
    return T, S


@pytest.mark.parametrize("name,prec,core,sem", [
    ("fp64 cut-off mex", winds.PREC_FP64, winds.CORE_CUTOFF, winds.SEM_MEX),
    ("fp32c vatistas2", winds.PREC_FP32C, winds.CORE_VATISTAS, winds.SEM_MATLAB),
    ("fp32c cut-off mex", winds.PREC_FP32C, winds.CORE_CUTOFF, winds.SEM_MEX),
])
def test_graph_replays_match_eager(need_gpu, name, prec, core, sem):
    import torch
    from winds.device import Evaluator
    flush_print(f"[graph capture] {name}")
    o = winds.opts(core=core, vat_n=2, semantics=sem, precision=prec, d=1e-4)
    ev = Evaluator(0)
    T, S = wake(7)
    nt, ns = T.shape[1], S.shape[0]
    ev.workspace(nt, ns, o)                       # allocated before the capture
    out = torch.empty((3, nt), dtype=torch.float64, device="cuda")
    ev.eval(T, S, o, out=out)                     # warm-up (first-call set-up)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ev.eval(T, S, o, out=out)
    torch.cuda.synchronize()
    for r in range(5):
        T.mul_(1.0 + 0.01 * r)                    # eager torch work between replays
        S[:, 6].add_(0.1)
        ref = ev.eval(T, S, o).clone()            # eager reference on the same inputs
        torch.cuda.synchronize()
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (name, r)
    flush_print(f"[graph capture] {name}: 5 replays equal to eager")
