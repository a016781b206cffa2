"""FP32 compensated sums on seeded filament soups: random polylines (1-60
segments each, steps of 0.2-5 m turning at random, circulations of either
sign, core radii r_c^2 from 1e-4 to 0.1 m^2) in a 200 m box, with targets
far away, close to segment midpoints (0.05-5 m off) and on every node.
Unlike the rotor wakes, nothing here is smooth or ordered, so this is where
the FP32 rounding, the chained r = r1 - L, the near-pair lists, the far
cut-off's band and the cancellation guard meet their worst cases.  Each
core against the engine's FP64 evaluation of the same soup: per-target
relative error <= 1e-6 for targets with |u| >= 1e-6 of the largest, and
normwise <= 1e-7.  (Measured over the 32 seeds: 3.5e-7 to 4.1e-7 without
the far cut-off, 9.8e-7 with co = 50 m inside the box.)"""
import numpy as np
import pytest

from conftest import flush_print
from errstats import stats
import winds

pytestmark = pytest.mark.gpu


def soup(seed):
    rng = np.random.default_rng(seed)
    recs, nodes = [], []
    for _ in range(int(rng.integers(5, 41))):
        n = int(rng.integers(1, 61))
        p = rng.uniform(-100, 100, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        g = rng.choice([-1, 1]) * rng.uniform(1, 80)
        rc = 10 ** rng.uniform(-4, -1)
        for _ in range(n):
            d = d + 0.3 * rng.normal(size=3)
            d /= np.linalg.norm(d)
            q = p + rng.uniform(0.2, 5.0) * d
            recs.append([*p, *q, g * rng.uniform(0.9, 1.1), rc])
            nodes.append(p)
            p = q
        nodes.append(p)
    recs = np.array(recs)
    mids = 0.5 * (recs[:, :3] + recs[:, 3:6])
    far = rng.uniform(-120, 120, (2000, 3))
    near = mids[rng.integers(0, len(recs), 2000)] + \
        rng.normal(size=(2000, 3)) * rng.uniform(0.05, 5.0, (2000, 1))
    return recs, np.concatenate([far, near, np.array(nodes)]).T.copy()


CORES = {
    "cutoff_mex": dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4),
    "cutoff_matlab": dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MATLAB, d=1e-4),
    "vatistas1": dict(core=winds.CORE_VATISTAS, vat_n=1, semantics=winds.SEM_MATLAB),
    "vatistas2": dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB),
    "vatistas3": dict(core=winds.CORE_VATISTAS, vat_n=3, semantics=winds.SEM_MATLAB),
    "vatistas2_co50": dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB, co=50.0),
}


@pytest.mark.parametrize("core", list(CORES))
def test_fp32c_filament_soups(need_gpu, core):
    import torch
    from winds.device import Evaluator
    ev = Evaluator(0)
    kw = CORES[core]
    worst = {"max_rel": 0.0, "normwise": 0.0, "rms_rel": 0.0}
    for seed in range(32):
        recs, tg = soup(seed)
        T = torch.from_numpy(tg).cuda().contiguous()
        S = torch.from_numpy(recs).cuda().contiguous()
        ref = ev.eval(T, S, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
        u = ev.eval(T, S, winds.opts(precision=winds.PREC_FP32C, **kw)).cpu().numpy()
        st = stats(u, ref)
        for k in worst:
            worst[k] = max(worst[k], st[k])
        assert st["max_rel"] <= 1e-6 and st["normwise"] <= 1e-7, (seed, st)
    flush_print(f"[soups {core}] worst over 32 seeds: {worst}")
