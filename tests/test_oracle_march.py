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
  accurate (halving dt divides the error by 2, 4 and 16);
- Kutta-Joukowski (KuttaJoukowski.m:42-49, :173-298) against Prandtl's
  lifting line for an elliptic wing (C_L = 2 pi alpha0 / (1 + 2/AR), uniform
  downwash, elliptic circulation), with first-order convergence in the
  station spacing;
- perform.m:42-62 against blade-element forces built as vectors (relative
  speed projected on the inflow direction, lift normal and drag along it).
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


def _elliptic_wing(nst, alpha0_deg, hub=5.0, R=65.0, c0=7.5, U=10.0, far=1e6):
    """A fixed blade (Omega = 0, one blade along +y) with an elliptic chord
    and a linear polar cl = 2 pi alpha, set at geometric angle of attack
    alpha0 in a uniform stream U e_x, with a flat wake of semi-infinite
    trailing lines: one wake age far downstream (A = 2), so the Kelvin update
    of KuttaJoukowski.m:285-288 moves every trailing filament with the bound
    circulation and kj() alone solves the steady lifting line.  The MEX
    cut-off is made negligible (d = 1e-12)."""
    from types import SimpleNamespace
    from winds import gen

    r = gen.cosine_stations(nst, R, hub)
    rn = r[:-1] + np.diff(r) / 2
    eta = (2 * rn - R - hub) / (R - hub)
    cn = c0 * np.sqrt(1 - eta ** 2)
    a = np.linspace(-20.0, 20.0, 81)
    # twist + pitch of 90 deg lays the chord along e_x; 90 - alpha0 pitches
    # the nose up by alpha0 (aoa of the bare stream = alpha0)
    blade = SimpleNamespace(r_node=rn, twist_node=np.full(rn.size, 90.0 - alpha0_deg),
                            pitch_deg=0.0, chord_node=cn, alpha=a,
                            cl=2 * np.pi * np.radians(a), cd=np.zeros_like(a))
    om = object.__new__(OracleMarch)
    om.r, om.nb, om.nst, om.ns, om.om = blade, 1, nst, nst - 1, 0.0
    om.U = np.array([U, 0.0, 0.0])
    om.d, om.co, om.sem, om.threads = 1e-12, np.inf, "mex", 1
    om.relax, om.tol, om.maxiter = 0.25, 1e-13, 2000
    dom = np.zeros((nst, 3, 2, 1))
    dom[:, 1, :, 0] = r[:, None]
    dom[:, 0, 1, 0] = far
    om.dom = np.asfortranarray(dom)
    g0 = np.zeros((nst - 1, 1))
    om.gs, om.gt = OracleMarch.kelvin(g0, np.zeros((nst, 1, 0, 1)), 2)
    om.gbound = g0
    iters, cl, _, aoa = om.kj(0.0, np.ones((nst - 1, 1, 2, 1)), np.ones((nst, 1, 1, 1)))
    assert iters < om.maxiter
    b = R - hub
    S = np.pi * b * c0 / 4
    CL = float(np.sum(cl[:, 0] * cn * np.diff(r)) / S)
    return CL, aoa[:, 0], eta, b * b / S, om.gbound[:, 0]


def test_kj_matches_prandtl_elliptic_wing():
    """Prandtl's lifting line for an elliptic planform of aspect ratio AR
    with lift slope 2 pi: the downwash is uniform along the span, the
    effective angle of attack is alpha0 / (1 + 2/AR) everywhere, and
    C_L = 2 pi alpha0 / (1 + 2/AR).  The discrete lattice of kj() (trailing
    lines at the stations, downwash at the nodes) converges to it at first
    order in the station spacing: 0.7% on C_L at 33 stations, half the
    deviation of 17 stations.  A sign error in the induced velocity, the
    Kelvin update or the factor 1/2 of Gamma = c |V| cl / 2 makes the fixed
    point diverge; swapped blade-frame axes put C_L 17-20% off."""
    devs = {}
    for nst in (17, 33):
        for alpha0 in (2.0, 6.0):
            CL, aoa, eta, AR, gam = _elliptic_wing(nst, alpha0)
            ae = alpha0 / (1 + 2 / AR)
            CL_p = 2 * np.pi * np.radians(ae)
            devs[nst, alpha0] = CL / CL_p - 1
            if nst == 33:
                assert abs(CL / CL_p - 1) <= 0.01, (alpha0, CL, CL_p)
                # uniform effective angle over the inner half of the span,
                # to 2% of the induced angle alpha0 - ae
                mid = np.abs(eta) < 0.5
                assert np.max(np.abs(aoa[mid] - ae)) <= 0.02 * (alpha0 - ae)
                # elliptic circulation
                shape = np.sqrt(1 - eta ** 2)
                assert np.max(np.abs(gam / gam.max() - shape / shape.max())) <= 0.03
    # linear in alpha0 (up to the arctan of the angle of attack: 1.3e-5
    # between 2 and 6 degrees), and first-order convergence in the spacing
    assert abs(devs[33, 2.0] - devs[33, 6.0]) <= 5e-5
    assert 0.3 <= devs[33, 2.0] / devs[17, 2.0] <= 0.6, devs


def test_perform_resolves_blade_element_forces():
    """perform.m:42-62 against the blade-element forces built as vectors:
    the relative speed is the projection of the kinematic velocity on the
    inflow direction n_D = sin(phi) e_x - cos(phi) e_t (the induction factors
    of perform.m make the induced velocity normal to it: v1 (1 - a) =
    W sin(phi), v2 (1 + a') = W cos(phi)), drag acts along n_D and lift along
    n_L = cos(phi) e_x + sin(phi) e_t, thrust is the e_x component of the sum
    and torque the sum of r times the e_t components."""
    rot = small_rotor()
    om = OracleMarch(rot, 6)
    rng = np.random.default_rng(11)
    ns, nb = rot.ns, rot.nb
    aoa = rng.uniform(2.0, 14.0, (ns, nb))
    cl = rng.uniform(0.3, 1.4, (ns, nb))
    cd = rng.uniform(0.005, 0.05, (ns, nb))
    t = 0.37
    T, Q, Pw = om.perform(cl, cd, aoa, t)
    psi = 2 * np.pi * np.arange(nb) / nb + rot.omega * t
    ex = np.array([1.0, 0.0, 0.0])
    Tv = Qv = 0.0
    for b in range(nb):
        et = np.array([0.0, -np.sin(psi[b]), np.cos(psi[b])])
        for i in range(ns):
            rn = rot.r_node[i]
            V = rot.uinf - rot.omega * rn * et
            phi = np.radians(aoa[i, b] - (rot.pitch_deg + rot.twist_node[i]))
            nD = np.sin(phi) * ex - np.cos(phi) * et
            nL = np.cos(phi) * ex + np.sin(phi) * et
            W = V @ nD
            F = 0.5 * 1.23 * rot.chord_node[i] * W * W * rot.dr_node[i] * (cl[i, b] * nL + cd[i, b] * nD)
            Tv += F @ ex
            Qv += rn * (F @ et)
    assert abs(T - Tv) <= 1e-12 * abs(Tv)
    assert abs(Q - Qv) <= 1e-12 * abs(Qv)
    assert abs(Pw - rot.omega * Qv) <= 1e-12 * abs(rot.omega * Qv)
    # an axial rotor's loads do not depend on its azimuth
    assert abs(om.perform(cl, cd, aoa, t + 0.9)[0] - T) <= 1e-12 * abs(T)
