"""The host-side preprocessing the device march reads as tables: DCMRot.m,
kinematics.m, velocity.m, ctdiff.m (winds/kinematics.py), BEM.m
(winds/bem.py), the airfoil reader of NRELrotor.m (winds/airfoils.py) and the
case builders (winds/cases.py).  Pinned by closed forms: single-axis
rotations and their order of composition, exactness of the finite
differences on polynomials, rigid rotation on circles, the analytic blade
velocity and frames, the hub offset of tilt and overhang, the momentum
balance of the converged BEM solution, and the .dat files themselves."""
import math
from pathlib import Path

import numpy as np
import pytest

from winds import airfoils as af
from winds import bem, cases, kinematics as K

REF = Path("/root/reference/modeldata/NREL_files")


def test_dcm_single_axes_and_order():
    x = np.array([1.0, 0.0, 0.0])
    # z by 90: e_x -> e_y;  then x by 90: e_y -> e_z
    assert np.allclose(K.dcm_rot(x, [[90.0]], "z"), [[0, 1, 0]], atol=0)
    assert np.allclose(K.dcm_rot(x, [[90.0, 90.0]], "zx"), [[0, 0, 1]], atol=0)
    # the other order leaves e_x on the x axis, then turns it to e_y
    assert np.allclose(K.dcm_rot(x, [[90.0, 90.0]], "xz"), [[0, 1, 0]], atol=0)
    # y by 90: e_z -> e_x (DCMRot.m:47)
    assert np.allclose(K.dcm_rot([0.0, 0.0, 1.0], [[90.0]], "y"), [[1, 0, 0]], atol=0)
    rng = np.random.default_rng(0)
    t = rng.uniform(-180, 180, (50, 6))
    v = rng.normal(size=(50, 3))
    y = K.dcm_rot(v, t, "zyxzyx")
    assert np.allclose(np.linalg.norm(y, axis=1), np.linalg.norm(v, axis=1), rtol=1e-14)
    back = K.dcm_rot(y, t, "zyxzyx", rev=True)          # the transpose undoes it
    assert np.allclose(back, v, atol=1e-13)
    # sind/cosd exact at multiples of 90 degrees, as MATLAB's
    assert K.sind(180.0) == 0.0 and K.cosd(90.0) == 0.0 and K.sind(-90.0) == -1.0


def test_ctdiff_is_exact_on_polynomials():
    x = np.linspace(0.0, 1.4, 15)
    y2 = np.stack([3 * x ** 2 - x + 2, -x ** 2], axis=1)
    d2 = np.stack([6 * x - 1, -2 * x], axis=1)
    assert np.allclose(K.ctdiff(x, y2), d2, atol=1e-12)      # every stencil: order >= 2
    y4 = (x ** 4 - 2 * x ** 3)[:, None] * np.ones((1, 2))
    d4 = (4 * x ** 3 - 6 * x ** 2)[:, None] * np.ones((1, 2))
    assert np.allclose(K.ctdiff(x, y4)[2:-2], d4[2:-2], atol=1e-11)   # interior: order 4
    # 4-D arrays differentiate along the third dimension
    y = np.zeros((2, 3, 15, 2))
    y[:, :, :, :] = (x ** 2)[None, None, :, None]
    assert np.allclose(K.ctdiff(x, y), (2 * x)[None, None, :, None], atol=1e-12)


def test_rigid_rotor_positions_frames_and_inflow():
    c = cases.synthetic_rotor(nb=3, nst=9, R=30.0, hub=1.0, tsr=6.0, U=9.0, dpsi_deg=4.0,
                              n_times=40)
    pos = K.kinematics(c.blade, c.turbine, c.platform, c.fastout)
    vel, pos = K.velocity(pos, c.blade, c.turbine, c.wind, c.fastout)
    omega = 6.0 * 9.0 / 30.0
    T = c.fastout.Time
    # quarter-chord nodes on circles about the x axis, blade b at Az + 120 b
    q = pos.quarter
    assert np.allclose(q[:, 0], 0.0, atol=1e-12)
    assert np.allclose(np.hypot(q[:, 1], q[:, 2]), c.blade.RTrail[:, None, None], rtol=1e-14)
    ang = np.arctan2(q[-1, 2], q[-1, 1])                      # [nt, nb]
    expect = np.radians(c.fastout.Azimuth)[:, None] + np.radians(120.0 * np.arange(3))[None] + np.pi / 2
    d = np.angle(np.exp(1j * (ang - expect)))
    assert np.max(np.abs(d)) <= 1e-12
    # trailing edge 0.75 chord behind the quarter chord
    assert np.allclose(np.linalg.norm(pos.trail - pos.quarter, axis=1),
                       0.75 * c.blade.ChordTrail[:, None, None], rtol=1e-13)
    # inflow = U e_x - Omega x r, to ctdiff's fourth order inside the range
    rb = pos.bound
    vrot = np.stack([np.zeros_like(rb[:, 0]), -omega * rb[:, 2], omega * rb[:, 1]], axis=1)
    ana = np.array([9.0, 0, 0])[None, :, None, None] - vrot
    err = np.abs(vel.bound - ana)[:, :, 2:-2] / 9.0
    dt = T[1] - T[0]
    assert err.max() <= 0.05 * (omega * dt) ** 4 * 60
    # blade frames orthonormal, bzn along the span, bxn along the chord
    for a, b in (("bxn", "byn"), ("byn", "bzn"), ("bxn", "bzn")):
        dot = np.sum(getattr(pos.nodes, a) * getattr(pos.nodes, b), axis=1)
        assert np.max(np.abs(dot)) <= 1e-12
    ez = pos.bound / np.linalg.norm(pos.bound, axis=1, keepdims=True)
    assert np.allclose(pos.nodes.bzn[1:], ez[1:], atol=1e-12)
    # geometric angle of attack of the rigid rotor: inflow angle - twist
    # (AeroTwst = -twist, pitch 0): atan(U / (Omega r)) + AeroTwst
    rn = c.blade.RNodes
    aoag = pos.aoag[:, 0, 5, 0]
    ref = np.degrees(np.arctan2(9.0, omega * rn)) + c.blade.AeroTwst
    assert np.allclose(aoag, ref, atol=2e-6)


def test_nrel_hub_tilt_and_precone():
    c = cases.nrel_rotor(t=(0.0, 1.0, 10.0))
    pos = K.kinematics(c.blade, c.turbine, c.platform, c.fastout)
    tilt = math.radians(5.0)
    assert np.allclose(pos.hub, [5.0 * math.cos(tilt), 0.0, 90.0 + 5.0 * math.sin(tilt)],
                       atol=1e-12)
    # the quarter-chord line of blade 0 at azimuth 0 points up, coned by
    # -2.5 deg and tilted 5 deg: its unit vector relative to the hub
    u = pos.quarter[-1, :, 0, 0] - pos.hub[0]
    u /= np.linalg.norm(u)
    cone = math.radians(-2.5)
    v = np.array([math.sin(cone), 0.0, math.cos(cone)])
    Ry = np.array([[math.cos(tilt), 0, math.sin(tilt)], [0, 1, 0],
                   [-math.sin(tilt), 0, math.cos(tilt)]])
    # rotation order of kinematics.m: precone about y, then tilt about y
    assert np.allclose(u, Ry @ v, atol=1e-12)


def test_platform_modes_move_the_hub():
    modes = (("surge", 2.0, 0.2, 0.0, 0.0, 0.0, 0.0, 0.5, 0.0),)
    c = cases.nrel_rotor(modes=modes, t=(0.0, 5.0, 20.0))
    pos = K.kinematics(c.blade, c.turbine, c.platform, c.fastout)
    vel, pos = K.velocity(pos, c.blade, c.turbine, c.wind, c.fastout)
    T = c.fastout.Time
    surge = 2.0 * np.sin(2 * np.pi * 0.2 * T) + 0.5
    assert np.allclose(pos.platform[:, 0], surge, atol=1e-12)
    assert np.allclose(pos.hub[:, 0] - pos.hub[0, 0], surge - surge[0], atol=1e-12)
    dsurge = 2.0 * 2 * np.pi * 0.2 * np.cos(2 * np.pi * 0.2 * T)
    assert np.allclose(vel.platform[2:-2, 0], dsurge[2:-2], atol=1e-5)


def test_bem_momentum_balance():
    """At convergence the blade-element thrust coefficient sigma (1-a)^2
    C_n / sin^2(phi) equals the momentum thrust coefficient 4 F a (1-a) at
    every station below the Glauert threshold (C_T <= 0.96 F), and the
    tangential balance 4 F a' (1 - a) Omega r / U = sigma C_t (1-a) / sin^2
    (phi) ... written as a' = 1 / (4 F sin cos / (sigma C_t) - 1)."""
    c = cases.nrel_rotor(t=(0.0, 1.0, 10.0))
    pos = K.kinematics(c.blade, c.turbine, c.platform, c.fastout)
    vel, pos = K.velocity(pos, c.blade, c.turbine, c.wind, c.fastout)
    r = bem.bem(c.airfoils, c.blade, c.turbine, c.fastout, vel)
    assert r.converged
    a, ap, phi = r.a[:, 3], r.ap[:, 3], r.phi[:, 3]
    cl, cd = r.cl[:, 3], r.cd[:, 3]
    rn = c.blade.RNodes
    sigma = 3 * c.blade.Chord / (2 * np.pi * rn)
    s, co = np.sin(phi), np.cos(phi)
    cn, ct_ = cl * co + cd * s, cl * s - cd * co
    F = (2 / np.pi * np.arccos(np.exp(-3 * (63.0 - rn) / (2 * rn * s))) *
         2 / np.pi * np.arccos(np.exp(-3 * (rn - 1.5) / (2 * 1.5 * s))))
    CT = sigma * (1 - a) ** 2 * cn / s ** 2
    lift = cl != 0                               # the cylinders carry no lift
    ok = lift & (CT <= 0.96 * F)
    assert ok.sum() >= 8
    assert np.allclose(CT[ok], 4 * F[ok] * a[ok] * (1 - a[ok]), rtol=2e-4)
    assert np.allclose(ap[lift], 1 / (4 * F[lift] * s[lift] * co[lift] / (sigma[lift] * ct_[lift]) - 1),
                       rtol=2e-4)
    # the inflow angle is consistent with the inductions and the kinematics
    # (phi of the last iteration, before its update of a by at most 1e-6)
    Om = c.fastout.RotSpeed[3] * np.pi / 30
    U = np.linalg.norm(vel.relhub[3])
    assert np.allclose(np.tan(phi), U * (1 - a) / (Om * rn * (1 + ap)), rtol=1e-5)


def test_airfoil_reader(tmp_path):
    # a file with a repeated angle of attack: the first of the pair goes
    lines = ["title", "second line"] + [f"   {v}   header" for v in
                                        (1, 0, 8.5, 0, 0, 0, -2, 6, 1.4, -0.5, -1.5, 0.006)]
    lines += ["-180.0  0.0  0.02  0.0", "0.0  0.1  0.01  0.0", "0.0  0.2  0.01  0.0",
              "180.0  0.0  0.02  0.0", ""]
    f = tmp_path / "x.dat"
    f.write_text("\n".join(lines))
    p = af.read_dat(f)
    assert np.array_equal(p.AoA, [-180.0, 0.0, 180.0]) and np.array_equal(p.Cl, [0.0, 0.2, 0.0])
    assert p.header["StallAoA"] == 8.5 and p.header["Cdmin"] == 0.006
    profs = af.nrel5mw_polars()
    assert [q.name for q in profs] == list(af.NREL5MW_NAMES)
    assert all(np.all(np.diff(q.AoA) > 0) and q.AoA[0] == -180 and q.AoA[-1] == 180 for q in profs)


@pytest.mark.skipif(not REF.exists(), reason="the reference's .dat files are not on this host")
def test_packaged_polars_equal_the_dat_files():
    for q in af.nrel5mw_polars():
        p = af.read_dat(REF / f"{q.name}.dat")
        assert np.array_equal(p.AoA, q.AoA) and np.array_equal(p.Cl, q.Cl)
        assert np.array_equal(p.Cd, q.Cd) and np.array_equal(p.Cm, q.Cm)


def test_nrel_blade_nodes():
    """NRELrotor.m:48-53: RNodes at the midpoints, NFoil of the nearest
    trailing node (halfway: the outer one, as interp1 'nearest'), and the
    station lengths."""
    c = cases.nrel_rotor(t=(0.0, 0.4, 10.0))
    b = c.blade
    assert b.RNodes.size == 18
    assert np.array_equal(b.NFoil, [1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8, 8])
    assert np.allclose(b.DRNodes, np.diff(cases.NREL_RTRAIL))
    assert np.allclose(b.Chord[0], 3.542) and np.allclose(b.AeroTwst[0], -13.308)


def test_tower_shadow_is_the_flow_around_a_cylinder():
    """tower_shadow.m: the doublet flow U e_x + V around the tower has no
    normal component on the cylinder surface, decays as 1/r^2, is zero above
    the tower and inside it; a downwind rotor's wake region gets -0.3 U."""
    from types import SimpleNamespace
    from winds import tower
    tur = SimpleNamespace(TowerHt=90.0, TowerBaseDiam=6.0, TowerTopDiam=4.0)
    U = 11.0
    th = np.linspace(-np.pi, np.pi, 37)[:-1]
    z = 45.0
    R = (6.0 + (4.0 - 6.0) * z / 90.0) / 2
    for rr in (R * (1 + 1e-12), 2 * R, 4 * R):
        P = np.stack([rr * np.cos(th), rr * np.sin(th), np.full(th.size, z)], 1)[:, :, None]
        V = tower.tower_shadow(P, U, 1, tur)[:, :, 0, 0]
        if rr == R * (1 + 1e-12):
            vn = (U + V[:, 0]) * np.cos(th) + V[:, 1] * np.sin(th)
            assert np.max(np.abs(vn)) <= 1e-9 * U
        mag = np.hypot(V[:, 0], V[:, 1])
        assert np.allclose(mag, U * (R / rr) ** 2, rtol=1e-12)
        assert np.all(V[:, 2] == 0)
    P = np.array([[3.0, 0.0, 95.0], [0.5, 0.5, 10.0], [-20.0, 1.0, 10.0]])[:, :, None]
    V = tower.tower_shadow(P, U, 1, tur)
    assert np.all(V[0] == 0) and np.all(V[1] == 0) and np.any(V[2] != 0)
    Pd = np.array([[10.0, 2.0, 30.0], [-10.0, 2.0, 30.0]])[:, :, None]
    Vd = tower.tower_shadow(Pd, U, 2, tur)[:, :, 0, 0]
    assert Vd[0, 0] == -0.3 * U and Vd[0, 1] == 0 and Vd[1, 0] != -0.3 * U
