"""Device time march against the CPU restatement of the WInDS loop.

The FP64 march must follow the oracle march (oracle/march.py, which restates
WInDS.m, initials.m, pcc.m/fe.m/ab4.m, vcore.m, filamentmod.m,
KuttaJoukowski.m and perform.m) to 1e-9 relative over every step, including
steps after the wake reaches its maximum length.  Both read the same
kinematic tables (winds.cases.tables).  Sums differ only in order (the device
splits the lifting-line sum into the filaments fixed during the
Kutta-Joukowski iteration and those it changes).  Cases: a small synthetic
rotor in every integrator and core, the NREL 5-MW rotor (eight airfoil
tables, shaft tilt, precone, tower height and overhang), the same rotor on a
surging and pitching platform with yaw, and the 3 x 31-node, 5-degree
geometry of the full-length benchmark march, and the rotor read from a FAST
file set."""
import math
from functools import lru_cache

import numpy as np
import pytest

from conftest import flush_print
import winds
from winds import cases

pytestmark = pytest.mark.gpu


@lru_cache(maxsize=None)
def small_tables(n_times=40):
    return cases.tables(cases.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0,
                                              dpsi_deg=10.0, n_times=n_times))


def wake_opts(sem, d, prec=winds.PREC_FP64):
    if sem == "mex":
        return winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=prec, d=d)
    if isinstance(d, str):
        return winds.opts(core=winds.CORE_VATISTAS, vat_n=int(d[4:]), semantics=winds.SEM_MATLAB,
                          precision=prec)
    return winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MATLAB, precision=prec, d=d,
                      co=1000.0)


def run_pair(tb, max_ages, steps, sem, d, integ):
    from oracle.march import OracleMarch
    from winds.march import March
    m = March(tb, max_ages, wake_opts(sem, d), integrator=integ)
    om = OracleMarch(tb, max_ages, d=d, co=1000.0 if sem != "mex" else math.inf,
                     semantics=sem, integ=integ)
    for _ in range(steps):
        m.step(1)
        om.step()
    m.sync()
    return m, om


def compare(m, om, label):
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
    dcd = np.max(np.abs(cd - om.last["cd"])) / np.max(np.abs(om.last["cd"]))
    gs, gt = m.state_circulation(om.gs.shape[2])
    dg = max(np.max(np.abs(gs - om.gs)), np.max(np.abs(gt - om.gt))) / np.max(np.abs(om.gs))
    flush_print(f"[march {label}] nodes {dn:.2e} perf {dT:.2e} cl {dcl:.2e} cd {dcd:.2e} "
                f"gamma {dg:.2e} T={ref[-1, 1]:.4g} P={ref[-1, 3]:.4g}")
    assert dn <= 1e-9 and dT <= 1e-9 and dcl <= 1e-9 and dcd <= 1e-9 and dg <= 1e-9


@pytest.mark.parametrize("sem,d,integ", [("mex", 1e-4, "pcc"), ("matlab", "visc2", "pcc"),
                                         ("mex", 1e-4, "ab2"), ("mex", 1e-4, "ab4"),
                                         ("matlab", 1e-4, "fe")])
def test_march_fp64_matches_oracle(need_gpu, sem, d, integ):
    steps, max_ages = 14, 10        # the wake hits max_ages after 8 steps
    m, om = run_pair(small_tables(), max_ages, steps, sem, d, integ)
    compare(m, om, f"{sem} {d} {integ}")


def test_nrel_5mw_rotor_matches_oracle(need_gpu):
    """modeldata/NRELrotor.m at WInDS.m's defaults (TSR 7, 11.4 m/s, 10 Hz):
    18 stations on eight airfoils (two cylinders, five DU and the NACA64
    tables), shaft tilt 5 deg, precone -2.5 deg, hub 90 m up and 5 m over."""
    tb = cases.tables(cases.nrel_rotor(t=(0.0, 2.0, 10.0)))
    assert sorted(set(tb.foil.tolist())) == list(range(1, 9))
    m, om = run_pair(tb, 12, 16, "mex", 1e-4, "pcc")
    compare(m, om, "NREL 5-MW")


def test_nrel_5mw_on_a_moving_platform_matches_oracle(need_gpu):
    """The same rotor yawed 10 deg on a platform that surges 2 m and pitches
    3 deg at 0.2-0.3 Hz (NRELrotor.m:92-123 modes), blade pitch 2 deg,
    Vatistas core: the kinematic tables drive the march."""
    modes = (("surge", 2.0, 0.2, 0.0, 0.5, 0.3, 0.4, 0.0, 0.0),
             ("pitch", 3.0, 0.25, 0.2, 0.0, 0.0, 0.0, 1.0, 0.0))
    tb = cases.tables(cases.nrel_rotor(yaw=10.0, pitch=2.0, modes=modes, t=(0.0, 2.0, 10.0)))
    assert np.ptp(tb.hub[:, 0]) > 1.0                      # the hub moves with the platform
    m, om = run_pair(tb, 12, 16, "matlab", "visc2", "pcc")
    compare(m, om, "NREL 5-MW moving platform")


def test_full_length_geometry_matches_oracle(need_gpu):
    """3 blades x 31 nodes at 5-degree steps with PCC, the geometry of the
    1,440-age benchmark march, over 34 steps with a 30-age wake (the wake is
    truncated for the last 6)."""
    tb = cases.tables(cases.synthetic_rotor(nb=3, nst=31, n_times=40))
    m, om = run_pair(tb, 30, 34, "mex", 1e-4, "pcc")
    compare(m, om, "3 x 31, 5 deg")


def test_long_lifting_line_matches_oracle(need_gpu):
    """3 blades x 41 nodes (120 lifting-line points): the Kutta-Joukowski
    matrix in the chord frames (2 x 120 x 120 doubles) exceeds a workgroup's
    LDS, so the kernel forms its terms from device memory at every
    iteration, the path the 3 x 31 and NREL rotors do not take."""
    tb = cases.tables(cases.synthetic_rotor(nb=3, nst=41, n_times=20))
    m, om = run_pair(tb, 8, 12, "mex", 1e-4, "pcc")
    compare(m, om, "3 x 41")


def test_zero_lift_wake_convects_with_the_free_stream(need_gpu):
    """No lift: no circulation and no induction, so a node of age a >= 1
    sits where the trailing edge was a-1 steps ago, moved by the free
    stream (fe.m/pcc.m with vel.domain = wind.infty)."""
    from winds.march import March
    tb = cases.tables(cases.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0,
                                            dpsi_deg=10.0, n_times=20), cl_init=np.zeros(6))
    tb.cl = np.zeros_like(tb.cl)
    m = March(tb, 9, winds.opts(d=1e-4))
    m.step(12)
    m.sync()
    nodes, _, _ = m.state()
    p = m.info()["p"]
    U = tb.wind[0]
    for a in range(1, nodes.shape[2]):
        tr = tb.trail[:, :, p - 1 - (a - 1), :]
        expect = tr + U[None, :, None] * (a - 1) * tb.dt
        assert np.max(np.abs(nodes[:, :, a, :] - expect)) <= 1e-11 * tb.tip_radius, a


def test_tables_end_is_an_error(need_gpu):
    from winds.march import March
    m = March(small_tables(10), 9, winds.opts(d=1e-4))
    m.step(8)                     # time indices 2..9 need tables up to 10
    with pytest.raises(winds.WindsError, match="tables end"):
        m.step(1)


def test_nrel_5mw_with_tower_shadow_matches_oracle(need_gpu):
    """user.towerflag 'true': the tower's potential-flow velocity at the
    bound points (tower_shadow.m) joins the wake's induced velocity in the
    Kutta-Joukowski solve (KuttaJoukowski.m:232-238); the rotor sits 5 m
    downwind of the tower (user.toweroverhang = 5), downwind model."""
    tb = cases.tables(cases.nrel_rotor(t=(0.0, 2.0, 10.0)), tower=True, up_down=2)
    assert np.any(tb.ll_extra != 0)
    m, om = run_pair(tb, 12, 16, "mex", 1e-4, "pcc")
    compare(m, om, "NREL 5-MW tower shadow")
    # the tower changes the loads
    plain = cases.tables(cases.nrel_rotor(t=(0.0, 2.0, 10.0)))
    m2, _ = run_pair(plain, 12, 16, "mex", 1e-4, "pcc")
    assert np.max(np.abs(m.perf()[:, 1] - m2.perf()[:, 1])) > 1e-4 * np.max(np.abs(m2.perf()[:, 1]))


def test_fast_file_set_matches_oracle(need_gpu, tmp_path):
    """A FAST v7 file set (.fst, AeroDyn file, eight airfoil files, .out) of
    the NREL 5-MW rotor, imported by winds.fast (input_import.m,
    output_import.m, input_mod.m): 17 AeroDyn sections, the tip at half
    chord, the azimuth unwrapped from the output's 0-360 channel."""
    import fastfiles as FF
    from winds import fast
    tb = cases.tables(fast.fast_case(FF.nrel_file_set(tmp_path / "fast"), (0.0, 2.0, 10.0)))
    assert tb.ns == 17 and tb.nst == 18
    m, om = run_pair(tb, 12, 16, "mex", 1e-4, "pcc")
    compare(m, om, "FAST file set")


@pytest.mark.parametrize("sem,d,integ,prec", [("mex", 1e-4, "pcc", winds.PREC_FP64),
                                              ("mex", 1e-4, "ab4", winds.PREC_FP32C),
                                              ("matlab", "visc2", "fe", winds.PREC_FP64),
                                              ("mex", 1e-4, "ab2", winds.PREC_FP32C)])
def test_graph_replayed_steps_equal_eager_steps(need_gpu, sem, d, integ, prec):
    """winds_march_set_graphs: once the 9-age wake is steady, each of the
    six phases of the buffer rotation is recorded once and replayed.  A
    march with graphs and one without step side by side; between chunks of
    steps the program does eager work of its own (state and history reads,
    torch kernels); after every chunk nodes, circulation and the
    performance history are equal bit for bit."""
    import torch
    from winds.march import March
    flush_print(f"[march graphs] {sem} {d} {integ} precision {prec}")
    tb = small_tables(60)
    a = March(tb, 9, wake_opts(sem, d, prec), integrator=integ)
    b = March(tb, 9, wake_opts(sem, d, prec), integrator=integ)
    a.set_graphs(True)
    scratch = torch.zeros(1000, device="cuda")
    for chunk in (8, 3, 5, 7, 11, 13):       # 47 steps: growth, then ~7 rounds of 6 phases
        a.step(chunk)
        b.step(chunk)
        scratch.add_(1.0)
        na, gsa, gta = a.state()
        nb, gsb, gtb = b.state()
        assert np.array_equal(na, nb) and np.array_equal(gsa, gsb) and np.array_equal(gta, gtb)
        assert np.array_equal(a.perf(), b.perf())
    assert a.info()["p"] == b.info()["p"] == 49
