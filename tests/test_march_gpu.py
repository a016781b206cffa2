"""Device time march against the CPU restatement of the WInDS loop.

The FP64 march must follow the oracle march (oracle/march.py, which restates
WInDS.m, initials.m, pcc.m/fe.m/ab4.m, vcore.m, filamentmod.m,
KuttaJoukowski.m and perform.m) to 1e-9 relative over every step, including
steps after the wake reaches its maximum length.  Sums differ only in order
(the device splits the lifting-line sum into the filaments fixed during the
Kutta-Joukowski iteration and those it changes)."""
import math

import numpy as np
import pytest

from conftest import flush_print
import winds
from winds import rotor as rmod

pytestmark = pytest.mark.gpu


def small_rotor(**kw):
    return rmod.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0, dpsi_deg=10.0, **kw)


def run_pair(rot, max_ages, steps, sem, d, integ):
    from oracle.march import OracleMarch
    from winds.march import March
    if sem == "mex":
        o = winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=winds.PREC_FP64, d=d)
    elif isinstance(d, str):
        o = winds.opts(core=winds.CORE_VATISTAS, vat_n=int(d[4:]), semantics=winds.SEM_MATLAB,
                       precision=winds.PREC_FP64)
    else:
        o = winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MATLAB, precision=winds.PREC_FP64,
                       d=d, co=1000.0)
    m = March(rot, max_ages, o, integrator=integ)
    om = OracleMarch(rot, max_ages, d=d, co=1000.0 if sem != "mex" else math.inf,
                     semantics=sem, integ=integ)
    for _ in range(steps):
        m.step(1)
        om.step()
    m.sync()
    return m, om


@pytest.mark.parametrize("sem,d,integ", [("mex", 1e-4, "pcc"), ("matlab", "visc2", "pcc"),
                                         ("mex", 1e-4, "ab4"), ("matlab", 1e-4, "fe")])
def test_march_fp64_matches_oracle(need_gpu, sem, d, integ):
    rot = small_rotor()
    steps, max_ages = 14, 10        # the wake hits max_ages after 8 steps
    m, om = run_pair(rot, max_ages, steps, sem, d, integ)
    nodes, _, _ = m.state()
    assert nodes.shape == om.dom.shape
    scale = np.max(np.abs(om.dom))
    dn = np.max(np.abs(nodes - om.dom)) / scale
    perf = m.perf()
    ref = np.array(om.history)
    assert perf.shape == ref.shape
    assert np.array_equal(perf[:, 0], ref[:, 0])                     # time indices
    assert np.array_equal(perf[:, 4], ref[:, 4])                     # KJ iterations
    dT = np.max(np.abs(perf[:, 1:4] - ref[:, 1:4]) / np.abs(ref[:, 1:4]))
    cl, cd, aoa = m.kj()
    dcl = np.max(np.abs(cl - om.last["cl"])) / np.max(np.abs(om.last["cl"]))
    gs, gt = m.state_circulation(om.gs.shape[2])
    dg = max(np.max(np.abs(gs - om.gs)), np.max(np.abs(gt - om.gt))) / np.max(np.abs(om.gs))
    flush_print(f"[march {sem} {d} {integ}] nodes {dn:.2e} perf {dT:.2e} cl {dcl:.2e} gamma {dg:.2e}"
                f" T={ref[-1, 1]:.4g}")
    assert dn <= 1e-9 and dT <= 1e-9 and dcl <= 1e-9 and dg <= 1e-9


def test_zero_lift_wake_convects_with_the_free_stream(need_gpu):
    from winds.march import March
    rot = small_rotor()
    rot.cl = np.zeros_like(rot.cl)          # no lift: no circulation, no induction
    m = March(rot, 9, winds.opts(d=1e-4))
    m.step(12)
    m.sync()
    nodes, _, _ = m.state()
    from oracle.march import OracleMarch
    om = OracleMarch(rot, 9)
    p = m.info()["p"]
    # the node of age a (0-based, a >= 1) left the trailing edge a-1 steps ago
    for a in range(1, nodes.shape[2]):
        t_shed = (p - 1 - (a - 1)) * rot.dt
        _, tr = om.quarter_trail(t_shed)
        expect = tr + rot.uinf[None, :, None] * (a - 1) * rot.dt
        assert np.max(np.abs(nodes[:, :, a, :] - expect)) <= 1e-11 * rot.R, a
