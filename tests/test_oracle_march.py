"""Pins of the CPU restatement of the WInDS time march (oracle/march.py).

- vcore.m:34-37 and initials.m:134-135 against hand values: r0 = 0.04558 m
  for the NREL rotor at TSR 7 and 11.4 m/s, and an r_c^2 growth of
  0.0196506 m^2 per second of Time for |Gamma| = 60 m^2/s;
- filamentmod.m:27-41: no strain leaves r_c unchanged; filaments four times
  as long (strain 1 after the reference's double square root) halve r_c^2 on
  the interior ages only;
- Kelvin bookkeeping (WInDS.m:131-140): over a consistent history the
  circulation of the vortex lattice is continuous at every node (what enters
  a node leaves it) and each age's trailing circulation sums to zero;
- advection (pcc.m): with no lift there is no circulation, and every node
  moves with the free stream exactly; in a rigid rotation field the node
  updates of fe.m, pcc.m and ab4.m are first-, second- and fourth-order
  accurate (halving dt divides the error by 2, 4 and 16).
"""
import numpy as np

from oracle.march import OracleMarch
from winds import rotor as rmod


def small_rotor(**kw):
    return rmod.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0, dpsi_deg=10.0,
                                **kw)


def test_r0_and_vcore_hand_values():
    nrel = rmod.synthetic_rotor(nb=3, nst=7, R=63.0, hub=1.5, tsr=7.0, U=11.4)
    om = OracleMarch(nrel, 6)
    assert abs(om.r0 - 0.04558) <= 5e-6
    dom = om.dom
    A = dom.shape[2]
    gs = np.full((nrel.ns, 1, A, nrel.nb), 60.0)
    gt = np.full((nrel.nst, 1, A - 1, nrel.nb), -60.0)
    rcs, rct, _ = om.vcore(1 + 0, dom, gs, gt, None)          # Time 0: r0^2 only
    assert np.allclose(rcs, om.r0 ** 2, rtol=1e-15) and np.allclose(rct, om.r0 ** 2, rtol=1e-15)
    rcs, rct, lens = om.vcore(3, dom, gs, gt, None)           # p <= 3: no strain
    T = 2 * om.dt
    rate_s = (rcs - om.r0 ** 2) / T
    rate_t = (rct - om.r0 ** 2) / T
    assert np.allclose(rate_s, 0.0196506, rtol=3e-6) and np.allclose(rate_t, 0.0196506, rtol=3e-6)


def test_filamentmod_strain_cases():
    om = OracleMarch(small_rotor(), 12)
    for _ in range(6):
        om.step()
    dom = om.dom
    A = dom.shape[2]
    r = om.r
    gs = np.full((r.ns, 1, A, r.nb), 10.0)
    gt = np.full((r.nst, 1, A - 1, r.nb), 10.0)
    p = 8
    # the cores of p = 3 (no strain yet) and this wake's lengths
    rcs0, rct0, lens = om.vcore(3, dom, gs, gt, None)
    # no strain: the previous lengths are these lengths
    rcs, rct, _ = om.vcore(p, dom, gs, gt, lens)
    Tfac = (p - 1) / 2.0
    base_s = om.r0 ** 2 + (rcs0 - om.r0 ** 2) * Tfac
    base_t = om.r0 ** 2 + (rct0 - om.r0 ** 2) * Tfac
    assert np.allclose(rcs, base_s, rtol=1e-14) and np.allclose(rct, base_t, rtol=1e-14)
    # every filament four times as long (about the origin): strain 1
    big = 4.0 * dom
    rcs2, rct2, _ = om.vcore(p, big, gs, gt, lens)
    it, ish = slice(1, A - 2), slice(1, A - 1)
    assert np.allclose(rct2[:, :, it], base_t[:, :, it] / 2, rtol=1e-13)
    assert np.allclose(rcs2[:, :, ish], base_s[:, :, ish] / 2, rtol=1e-13)
    assert np.array_equal(rct2[:, :, 0], rct[:, :, 0])
    assert np.array_equal(rct2[:, :, A - 2], rct[:, :, A - 2])
    assert np.array_equal(rcs2[:, :, 0], rcs[:, :, 0])
    assert np.array_equal(rcs2[:, :, A - 1], rcs[:, :, A - 1])


def lattice_residual(gs, gt):
    """Circulation leaving minus entering each node (m, a), a < A-1.  Shed
    filament (m, a) runs node (m, a) -> (m+1, a); trailing filament (m, a)
    runs node (m, a+1) -> (m, a)."""
    ns, _, A, nb = gs.shape
    S = np.zeros((ns + 2, A, nb))
    S[1:-1] = gs[:, 0]
    T = np.zeros((ns + 1, A + 1, nb))
    T[:, 1:A] = gt[:, 0]
    m = np.arange(ns + 1)[:, None]
    a = np.arange(A - 1)[None, :]
    return S[m + 1, a] + T[m, a] - S[m, a] - T[m, a + 1]


def test_kelvin_keeps_the_lattice_closed():
    rng = np.random.default_rng(0)
    ns, nb = 6, 2
    gt = np.zeros((ns + 1, 1, 0, nb))
    A = 1
    for _ in range(9):
        A += 1
        gs, gt = OracleMarch.kelvin(rng.normal(size=(ns, nb)), gt, A)
    assert np.max(np.abs(lattice_residual(gs, gt))) <= 1e-14
    assert np.max(np.abs(gt[:, 0].sum(axis=0))) <= 1e-14
    # a planted mistake: shed circulation of one age with the wrong sign
    bad = gs.copy()
    bad[:, :, 3] *= -1
    assert np.max(np.abs(lattice_residual(bad, gt))) > 1e-3


def test_zero_lift_nodes_move_with_the_free_stream():
    rot = small_rotor()
    rot.cl = np.zeros_like(rot.cl)
    om = OracleMarch(rot, 9)
    for _ in range(12):
        om.step()
    p = 2 + 12
    dom = om.dom
    for a in range(1, dom.shape[2]):
        t_shed = (p - 1 - (a - 1)) * rot.dt
        _, tr = om.quarter_trail(t_shed)
        expect = tr + rot.uinf[None, :, None] * (a - 1) * rot.dt
        assert np.max(np.abs(dom[:, :, a, :] - expect)) <= 1e-11 * rot.R, a


def _rotation_field(omega):
    def v(dom):
        out = np.zeros_like(dom)
        out[:, 1] = -omega * dom[:, 2]
        out[:, 2] = omega * dom[:, 1]
        return out
    return v


def _integrate(om, dom, steps, dt, field, pcc):
    om.dt = dt
    for _ in range(steps):
        v = field(dom)
        pred = om._advect(dom, v, 0.0)
        dom = om._advect(dom, v, 0.0, vel_next=field(pred)) if pcc else pred
    return dom


def test_integrators_have_their_order_of_accuracy():
    om = OracleMarch(small_rotor(), 200)
    dom0 = om.dom.copy()
    omega, T = 2.0, 0.4
    field = _rotation_field(omega)
    c, s = np.cos(omega * T), np.sin(omega * T)
    start = dom0[:, :, 1, :]                       # the trailing-edge nodes
    exact = start.copy()
    exact[:, 1] = c * start[:, 1] - s * start[:, 2]
    exact[:, 2] = s * start[:, 1] + c * start[:, 2]
    for pcc, slope in ((False, 1.0), (True, 2.0)):
        errs = []
        for k in (16, 32, 64):
            dom = _integrate(om, dom0.copy(), k, T / k, field, pcc)
            errs.append(np.max(np.abs(dom[:, :, 1 + k, :] - exact)))
        rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
        assert np.all(np.abs(rates - slope) <= 0.15), (pcc, rates)


def _rotate(x, ang):
    out = x.copy()
    c, s = np.cos(ang), np.sin(ang)
    out[:, 1] = c * x[:, 1] - s * x[:, 2]
    out[:, 2] = s * x[:, 1] + c * x[:, 2]
    return out


def test_ab4_is_fourth_order():
    """Adams-Bashforth 4 (ab4.m) on nodes old enough to have three earlier
    velocities, started from the exact history."""
    om = OracleMarch(small_rotor(), 400)
    rng = np.random.default_rng(4)
    A0, a0 = 8, 5
    dom0 = np.asfortranarray(rng.uniform(-10, 10, size=(om.nst, 3, A0, om.nb)))
    omega, T = 2.0, 0.4
    field = _rotation_field(omega)
    exact = _rotate(dom0[:, :, a0, :], omega * T)
    errs = []
    for k in (16, 32, 64):
        dt = T / k
        om.dt = dt
        # the node now at age i + j was at age i j steps ago
        hist = [field(np.asfortranarray(_rotate(dom0[:, :, j:, :], -j * omega * dt)))
                for j in (1, 2, 3)]
        dom = dom0.copy()
        for _ in range(k):
            v = field(dom)
            dom = om._advect(dom, v, 0.0, hist=(v, hist[0], hist[1], hist[2]))
            hist = [v, hist[0], hist[1]]
        errs.append(np.max(np.abs(dom[:, :, a0 + k, :] - exact)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - 4.0) <= 0.3), rates
