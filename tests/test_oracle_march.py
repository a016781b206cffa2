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
  updates of fe.m, pcc.m, ab2.m and ab4.m are first-, second-, second- and
  fourth-order
  accurate (halving dt divides the error by 2, 4 and 16);
- Kutta-Joukowski (KuttaJoukowski.m:42-49, :173-298) against Prandtl's
  lifting line for an elliptic wing (C_L = 2 pi alpha0 / (1 + 2/AR), uniform
  downwash, elliptic circulation), with first-order convergence in the
  station spacing;
- perform.m:16-62 against blade-element forces built as vectors (relative
  speed projected on the inflow direction, lift normal and drag along it),
  with the rotor-frame inflow velX of a rigid rotor against its closed form;
- KuttaJoukowski.m:262-269: each station reads its own airfoil's table.

The oracle reads the same kinematic tables as the device march
(winds.cases.tables); the tables themselves are pinned in test_kinematics.py.
"""
import math

import numpy as np

from oracle.march import OracleMarch
from winds import cases


def small_rotor(n_times=40, **kw):
    return cases.tables(cases.synthetic_rotor(nb=2, nst=7, R=20.0, hub=1.0, tsr=6.0, U=10.0,
                                              dpsi_deg=10.0, n_times=n_times, **kw))


def test_r0_and_vcore_hand_values():
    nrel = cases.tables(cases.synthetic_rotor(nb=3, nst=7, R=63.0, hub=1.5, tsr=7.0, U=11.4,
                                              n_times=8))
    om = OracleMarch(nrel, 6)
    assert abs(om.r0(1) - 0.04558) <= 5e-6
    dom = om.dom
    A = dom.shape[2]
    gs = np.full((nrel.ns, 1, A, nrel.nb), 60.0)
    gt = np.full((nrel.nst, 1, A - 1, nrel.nb), -60.0)
    r0 = om.r0(1)
    rcs, rct, _ = om.vcore(1 + 0, dom, gs, gt, None)          # Time 0: r0^2 only
    assert np.allclose(rcs, r0 ** 2, rtol=1e-15) and np.allclose(rct, r0 ** 2, rtol=1e-15)
    rcs, rct, lens = om.vcore(3, dom, gs, gt, None)           # p <= 3: no strain
    T = 2 * om.dt
    rate_s = (rcs - r0 ** 2) / T
    rate_t = (rct - r0 ** 2) / T
    assert np.allclose(rate_s, 0.0196506, rtol=3e-6) and np.allclose(rate_t, 0.0196506, rtol=3e-6)


def test_filamentmod_strain_cases():
    om = OracleMarch(small_rotor(), 12)
    for _ in range(6):
        om.step()
    dom = om.dom
    A = dom.shape[2]
    r = om.tb
    gs = np.full((r.ns, 1, A, r.nb), 10.0)
    gt = np.full((r.nst, 1, A - 1, r.nb), 10.0)
    p = 8
    # the cores of p = 3 (no strain yet) and this wake's lengths
    rcs0, rct0, lens = om.vcore(3, dom, gs, gt, None)
    # no strain: the previous lengths are these lengths
    rcs, rct, _ = om.vcore(p, dom, gs, gt, lens)
    Tfac = (p - 1) / 2.0
    r0 = om.r0(p)
    base_s = r0 ** 2 + (rcs0 - r0 ** 2) * Tfac
    base_t = r0 ** 2 + (rct0 - r0 ** 2) * Tfac
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
    tb = small_rotor()
    tb.cl = np.zeros_like(tb.cl)
    tb.cl_init = np.zeros_like(tb.cl_init)
    om = OracleMarch(tb, 9)
    for _ in range(12):
        om.step()
    p = 2 + 12
    dom = om.dom
    for a in range(1, dom.shape[2]):
        tr = tb.trail[:, :, p - 1 - (a - 1), :]
        expect = tr + tb.wind[0][None, :, None] * (a - 1) * tb.dt
        assert np.max(np.abs(dom[:, :, a, :] - expect)) <= 1e-11 * tb.tip_radius, a


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
        pred = om._advect(dom, v, 2)
        dom = om._advect(dom, v, 2, vel_next=field(pred)) if pcc else pred
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
            dom = om._advect(dom, v, 2, hist=(v, hist[0], hist[1], hist[2]))
            hist = [v, hist[0], hist[1]]
        errs.append(np.max(np.abs(dom[:, :, a0 + k, :] - exact)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - 4.0) <= 0.3), rates


def test_ab2_is_second_order():
    """Adams-Bashforth 2 (ab2.m:38-39) on nodes old enough to have an
    earlier velocity, started from the exact history."""
    om = OracleMarch(small_rotor(), 400)
    rng = np.random.default_rng(5)
    A0, a0 = 8, 5
    dom0 = np.asfortranarray(rng.uniform(-10, 10, size=(om.nst, 3, A0, om.nb)))
    omega, T = 2.0, 0.4
    field = _rotation_field(omega)
    exact = _rotate(dom0[:, :, a0, :], omega * T)
    errs = []
    for k in (16, 32, 64):
        dt = T / k
        om.dt = dt
        prev = field(np.asfortranarray(_rotate(dom0[:, :, 1:, :], -omega * dt)))
        dom = dom0.copy()
        for _ in range(k):
            v = field(dom)
            dom = om._advect(dom, v, 2, hist=(v, prev))
            prev = v
        errs.append(np.max(np.abs(dom[:, :, a0 + k, :] - exact)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - 2.0) <= 0.2), rates


def _elliptic_wing(nst, alpha0_deg, hub=5.0, R=65.0, c0=7.5, U=10.0, far=1e6):
    """A fixed wing along +y with an elliptic chord and a linear polar
    cl = 2 pi alpha, its chord along e_x pitched nose-up by alpha0 (blade
    frame bxn = cos(a0) e_x - sin(a0) e_z, bzn = e_y, byn = bzn x bxn, so the
    bare stream U e_x meets it at alpha0), in a flat wake of semi-infinite
    trailing lines: one wake age far downstream (A = 2), so the Kelvin update
    of KuttaJoukowski.m:285-288 moves every trailing filament with the bound
    circulation and kj() alone solves the steady lifting line.  The MEX
    cut-off is made negligible (d = 1e-12)."""
    from types import SimpleNamespace
    from winds import gen

    r = gen.cosine_stations(nst, R, hub)
    rn = r[:-1] + np.diff(r) / 2
    ns = nst - 1
    eta = (2 * rn - R - hub) / (R - hub)
    cn = c0 * np.sqrt(1 - eta ** 2)
    a = np.linspace(-20.0, 20.0, 81)
    a0 = np.radians(alpha0_deg)
    bx = np.array([np.cos(a0), 0.0, -np.sin(a0)])
    bz = np.array([0.0, 1.0, 0.0])
    by = np.cross(bz, bx)
    Uv = np.array([U, 0.0, 0.0])
    tile = lambda v: np.asfortranarray(np.broadcast_to(v[None, :, None, None], (ns, 3, 1, 1)))
    bound = np.zeros((ns, 3, 1, 1), order="F")
    bound[:, 1, 0, 0] = rn
    tb = SimpleNamespace(nb=1, nst=nst, ns=ns, n_times=1, tip_radius=R, dt=1.0,
                         chord_node=cn, twist_node=np.zeros(ns), dr_node=np.diff(r), r_node=rn,
                         foil=np.ones(ns, dtype=np.int32), foil_len=np.array([a.size]),
                         alpha=a, cl=2 * np.pi * np.radians(a), cd=np.zeros_like(a),
                         bound=bound, bxn=tile(bx), byn=tile(by), bzn=tile(bz),
                         vblade=tile(np.array([bx @ Uv, by @ Uv, bz @ Uv])), ll_extra=None,
                         wind=np.array([[U, 0.0, 0.0]]))
    om = OracleMarch(tb, 2, d=1e-12, semantics="mex", relax=0.25, tol=1e-13, maxiter=2000,
                     threads=1, setup=False)
    dom = np.zeros((nst, 3, 2, 1))
    dom[:, 1, :, 0] = r[:, None]
    dom[:, 0, 1, 0] = far
    om.dom = np.asfortranarray(dom)
    g0 = np.zeros((nst - 1, 1))
    om.gs, om.gt = OracleMarch.kelvin(g0, np.zeros((nst, 1, 0, 1)), 2)
    om.gbound = g0
    iters, cl, _, aoa = om.kj(1, np.ones((nst - 1, 1, 2, 1)), np.ones((nst, 1, 1, 1)))
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
    """perform.m:16-62 against the blade-element forces built as vectors.
    For a rigid rotor about +x the rotor-frame inflow velX of perform.m:19-24
    is (axial speed, in-plane speed Omega r, 0) up to ctdiff's truncation;
    the relative speed is the projection of the kinematic velocity
    V = U e_x - Omega r e_t on the inflow direction n_D = sin(phi) e_x -
    cos(phi) e_t (the induction factors of perform.m make the induced velocity
    normal to it: v1 (1 - a) = W sin(phi), v2 (1 + a') = W cos(phi)), drag
    acts along n_D and lift along n_L = cos(phi) e_x + sin(phi) e_t, thrust is
    the e_x component of the sum and torque the sum of r times the e_t
    components."""
    tb = small_rotor()
    om = OracleMarch(tb, 6)
    rng = np.random.default_rng(11)
    ns, nb = tb.ns, tb.nb
    aoa = rng.uniform(2.0, 14.0, (ns, nb))
    cl = rng.uniform(0.3, 1.4, (ns, nb))
    cd = rng.uniform(0.005, 0.05, (ns, nb))
    p = 9
    U = tb.wind[0][0]
    omega = tb.rot_speed[0] * math.pi / 30
    vX = om.velX(p)
    assert np.allclose(vX[:, 0, :], U, rtol=2e-4)
    assert np.allclose(vX[:, 1, :], omega * tb.r_node[:, None], rtol=2e-4)
    assert np.all(np.abs(vX[:, 2, :]) <= 1e-9 * U)
    T, Q, Pw = om.perform(cl, cd, aoa, p)
    ex = np.array([1.0, 0.0])          # (axial, tangential) plane of the section
    et = np.array([0.0, 1.0])
    Tv = Qv = 0.0
    for b in range(nb):
        pitch = tb.pitch[p - 1, b]
        for i in range(ns):
            rn = tb.r_node[i]
            V = vX[i, 0, b] * ex - vX[i, 1, b] * et
            phi = np.radians(aoa[i, b] - (pitch + tb.twist_node[i]))
            nD = np.sin(phi) * ex - np.cos(phi) * et
            nL = np.cos(phi) * ex + np.sin(phi) * et
            W = V @ nD
            F = 0.5 * 1.23 * tb.chord_node[i] * W * W * tb.dr_node[i] * (cl[i, b] * nL + cd[i, b] * nD)
            Tv += F @ ex
            Qv += rn * (F @ et)
    assert abs(T - Tv) <= 1e-12 * abs(Tv)
    assert abs(Q - Qv) <= 1e-12 * abs(Qv)
    assert abs(Pw - omega * Qv) <= 1e-12 * abs(omega * Qv)
    # an axial rotor's loads do not depend on its azimuth (interior indices)
    assert abs(om.perform(cl, cd, aoa, p + 7)[0] - T) <= 1e-6 * abs(T)


def test_kj_uses_each_stations_airfoil():
    """KuttaJoukowski.m:262-269: station m interpolates table blade.NFoil(m).
    Three tables whose lift differs by a constant factor; the stations on
    each table get exactly that table's cl at their angle of attack."""
    tb = small_rotor()
    a = tb.alpha.copy()
    base = tb.cl.copy()
    n = a.size
    tb.alpha = np.concatenate([a, a, a])
    tb.cl = np.concatenate([base, 1.5 * base, 0.5 * base])
    tb.cd = np.concatenate([tb.cd, tb.cd, tb.cd])
    tb.foil_len = np.array([n, n, n], dtype=np.int32)
    tb.foil = np.array([1, 2, 3, 1, 2, 3], dtype=np.int32)
    om = OracleMarch(tb, 6, setup=False)
    aoa = np.linspace(2.0, 9.0, 12).reshape(6, 2)
    got = om.polar(tb.cl, aoa)
    ref = np.interp(aoa, a, base)
    fac = np.array([1.0, 1.5, 0.5, 1.0, 1.5, 0.5])[:, None]
    assert np.allclose(got, fac * ref, rtol=1e-15)
