"""Test cases: the structures the reference's case files return
(blade, turbine, platform, fastout, airfoils, wind; modeldata/NRELrotor.m,
modeldata/elliptical.m), and the kinematic tables the device march takes.

nrel_rotor     the NREL 5-MW rotor of modeldata/NRELrotor.m:29-163: blade
               geometry (NRELrotor.m:34-55), eight airfoil tables
               (NRELrotor.m:128-157, packaged data), rigid rotation at the
               tip speed ratio, optional platform modes (NRELrotor.m:92-123)
synthetic_rotor  a rotor of any size with a generated blade, a generated polar
               and a round root section, hub at the origin (what the
               benchmarks use)
elliptical     the elliptical wing of modeldata/elliptical.m:28-120
tables         kinematics.m + velocity.m + the initial lift coefficient
               (BEM.m, or the elliptic closed form of initials.m:111-113):
               everything winds.march.March needs
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from . import airfoils as af
from . import bem as bem_mod
from . import gen, kinematics as kin


def _time(t):
    """fastout.Time = (t(1):1/t(3):t(2))' (NRELrotor.m:76)."""
    t0, t1, freq = t
    n = int(math.floor((t1 - t0) * freq + 1e-9)) + 1
    return t0 + np.arange(n) / freq


def _blade_nodes(blade, nfoil_trail):
    """RNodes at the midpoints, AeroTwst and Chord interpolated there, NFoil
    of the nearest trailing node (a node halfway takes the outer one, as
    MATLAB's interp1 'nearest'), DRNodes (NRELrotor.m:48-53)."""
    RT = blade.RTrail
    blade.RNodes = RT[:-1] + np.diff(RT) / 2
    blade.AeroTwst = np.interp(blade.RNodes, RT, blade.AeroTwstTrail)
    blade.Chord = np.interp(blade.RNodes, RT, blade.ChordTrail)
    k = np.searchsorted(RT, blade.RNodes, side="right") - 1
    k = np.clip(k, 0, RT.size - 2)
    up = blade.RNodes >= (RT[k] + RT[k + 1]) / 2
    blade.NFoil = np.asarray(nfoil_trail)[k + up.astype(int)].astype(np.int32)
    blade.DRNodes = np.abs(np.diff(RT))
    return blade


def _rigid_fastout(time, omega, nb, pitch, yaw, wind):
    fo = SimpleNamespace()
    n = time.size
    fo.Time = time
    fo.WindVxi = np.full(n, float(wind[0]))
    fo.WindVyi = np.full(n, float(wind[1]))
    fo.WindVzi = np.full(n, float(wind[2]))
    fo.Azimuth = omega * (180 / math.pi) * time
    fo.BldPitch = np.full((n, nb), float(pitch))
    fo.NacYaw = np.full(n, float(yaw))
    for k in ("PtfmSurge", "PtfmSway", "PtfmHeave", "PtfmRoll", "PtfmPitch", "PtfmYaw"):
        setattr(fo, k, np.zeros(n))
    return fo


def _wind(fo):
    w = SimpleNamespace()
    w.infty = np.stack([fo.WindVxi, fo.WindVyi, fo.WindVzi], axis=1)
    w.time = fo.Time
    w.inftyM = np.sqrt(np.sum(w.infty ** 2, axis=1))
    return w


MODES = ("surge", "sway", "heave", "roll", "pitch", "yaw")


def _apply_modes(fo, modes):
    """Platform motions A1 sin(2 pi f1 t + p1) + A2 sin(2 pi f2 t + p2) + mean,
    starting at time t_start (NRELrotor.m:92-123).  modes: tuples
    (dof, A1, f1, p1, A2, f2, p2, mean, t_start)."""
    for dof, A1, f1, p1, A2, f2, p2, mn, ts in modes:
        tmm = fo.Time.copy()
        lt = tmm < ts
        gt = ~lt
        tmm[lt] = 0.0
        if gt.any():
            tmm[gt] = tmm[gt] - tmm[gt][0]
        v = A1 * np.sin(2 * np.pi * f1 * tmm + p1) + A2 * np.sin(2 * np.pi * f2 * tmm + p2) + mn
        d = dof.lower()
        if d not in MODES:
            raise ValueError(f"unknown platform mode {dof!r}")
        setattr(fo, "Ptfm" + d.capitalize(), v)


# NREL 5-MW blade (modeldata/NRELrotor.m:35-47)
NREL_RTRAIL = (1.5, 2.86670, 5.6, 8.33330, 11.75, 15.85, 19.95, 24.05, 28.15, 32.25, 36.35,
               40.45, 44.55, 48.65, 52.75, 56.1667, 58.90, 61.6333, 63.0)
NREL_CHORD = (3.542, 3.542, 3.854, 4.167, 4.557, 4.652, 4.458, 4.249, 4.007, 3.748, 3.502, 3.256,
              3.010, 2.764, 2.518, 2.313, 2.086, 1.419, 1.419 / 2)
NREL_TWIST = (13.308, 13.308, 13.308, 13.308, 13.308, 11.48, 10.162, 9.011, 7.795, 6.544, 5.361,
              4.188, 3.125, 2.319, 1.526, 0.863, 0.370, 0.106, 0.106 / 2)
NREL_NFOIL = (1, 1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8, 8)


def nrel_rotor(wind=(11.4, 0.0, 0.0), tsr: float = 7.0, pitch: float = 0.0, yaw: float = 0.0,
               t=(0.0, 5.0, 10.0), modes=(), overhang: float = 5.0):
    """modeldata/NRELrotor.m with the user settings of WInDS.m:22-60
    (user.rotor.wind, tsr, pitch, yaw, modes; user.t; user.toweroverhang)."""
    U = np.asarray(wind, dtype=np.float64)
    blade = SimpleNamespace(TipRad=63.0, HubRad=1.5)
    blade.RTrail = np.array(NREL_RTRAIL)
    blade.ChordTrail = np.array(NREL_CHORD)
    blade.AeroTwstTrail = -np.array(NREL_TWIST)
    blade.AR = blade.TipRad ** 2 / np.trapezoid(blade.ChordTrail, blade.RTrail)
    _blade_nodes(blade, NREL_NFOIL)
    omega = tsr * float(np.linalg.norm(U)) / blade.TipRad
    turbine = SimpleNamespace(NumBl=3, OverHang=overhang, TowerHt=90.0, Twr2Shft=0.0,
                              TowerBaseDiam=6.0, TowerTopDiam=3.87, ShftTilt=5.0,
                              PreCone=np.full(3, -2.5))
    platform = SimpleNamespace(Type="NRELRotor", TwrDraft=0.0, PtfmCm=0.0, PtfmRef=0.0,
                               PtfmDraft=0.0, PtfmDiam=0.0)
    fo = _rigid_fastout(_time(t), omega, 3, pitch, yaw, U)
    _apply_modes(fo, modes)
    fo.TipSpdRat = np.full(fo.Time.size, float(tsr))
    fo.RotSpeed = np.full(fo.Time.size, omega * 30 / math.pi)
    airfoils = SimpleNamespace(Names=list(af.NREL5MW_NAMES), profiles=af.nrel5mw_polars())
    return SimpleNamespace(blade=blade, turbine=turbine, platform=platform, fastout=fo,
                           airfoils=airfoils, wind=_wind(fo), name="NRELrotor")


def synthetic_polar(step_deg: float = 1.0):
    """Lift and drag over -180..180 deg: thin-airfoil slope 2*pi below ~12 deg,
    flat-plate 2 sin(a) cos(a) and 2 sin^2(a) beyond, smoothly blended."""
    a = np.arange(-180.0, 180.0 + 0.5 * step_deg, step_deg)
    ar = np.radians(a)
    s = 1.0 / (1.0 + (a / 12.0) ** 8)
    cl = 2.0 * np.sin(ar) * np.cos(ar) * (1.0 + 2.14159 * s)
    t = (a / 20.0) ** 4
    cd = 0.008 + 0.02 * (a / 10.0) ** 2 / (1.0 + (a / 10.0) ** 2) + 1.3 * np.sin(ar) ** 2 * t / (1.0 + t)
    return a, cl, cd


def synthetic_rotor(nb: int = 3, nst: int = 31, R: float = 63.0, hub: float = 1.5,
                    tsr: float = 7.0, U: float = 11.4, yaw_deg: float = 0.0,
                    dpsi_deg: float = 5.0, n_times: int = 64, pitch: float = 0.0,
                    spacing: str = "uniform"):
    """A rigid rotor about +x with its centre at the origin (no tower, shaft
    tilt or precone): cosine-spaced trailing nodes, chord 4.6 m at the hub
    tapering to 1.4 m at the tip, twist set so that the design inflow (axial
    induction 1/3 at tsr) meets every section at 7 degrees (AeroTwst = minus
    the twist, the reference's sign), a generated polar outboard and, as on
    the NREL 5-MW blade, a round root section (no lift, Cd 0.5) inboard of
    10% of the tip radius; time step of dpsi_deg of rotation, wind yawed by
    yaw_deg about z.  Trailing nodes are evenly spaced by default: with 31
    cosine-spaced nodes the innermost and outermost stations are 0.17 m apart
    under a 1.4-4.6 m chord, and the relaxed Kutta-Joukowski fixed point of
    KuttaJoukowski.m (relaxation 0.25, 30 iterations) then amplifies a
    change in the last bit of its inputs to 10% of cl within a few steps."""
    if spacing == "cosine":
        r = gen.cosine_stations(nst, R, hub)
    elif spacing == "uniform":
        r = np.linspace(hub, R, nst)
    else:
        raise ValueError(f"unknown spacing {spacing!r}")
    xi = (r - hub) / (R - hub)
    omega_d = tsr * U / R
    twist = np.degrees(np.arctan2(U * (2.0 / 3.0), omega_d * r)) - 7.0
    blade = SimpleNamespace(TipRad=R, HubRad=hub, RTrail=r, ChordTrail=4.6 - 3.2 * xi,
                            AeroTwstTrail=-twist)
    _blade_nodes(blade, np.where(r < 0.1 * R, 1, 2))
    y = math.radians(yaw_deg)
    wind = np.array([U * math.cos(y), U * math.sin(y), 0.0])
    omega = tsr * U / R
    dt = math.radians(dpsi_deg) / omega
    turbine = SimpleNamespace(NumBl=nb, OverHang=0.0, TowerHt=0.0, Twr2Shft=0.0,
                              TowerBaseDiam=0.0, TowerTopDiam=0.0, ShftTilt=0.0,
                              PreCone=np.zeros(nb))
    platform = SimpleNamespace(Type="Rotor", TwrDraft=0.0, PtfmCm=0.0, PtfmRef=0.0,
                               PtfmDraft=0.0, PtfmDiam=0.0)
    fo = _rigid_fastout(np.arange(n_times) * dt, omega, nb, pitch, 0.0, wind)
    fo.TipSpdRat = np.full(n_times, float(tsr))
    fo.RotSpeed = np.full(n_times, omega * 30 / math.pi)
    a, cl, cd = synthetic_polar()
    ac = np.array([-180.0, 0.0, 180.0])
    root = af.Profile("round root", ac, np.zeros(3), np.full(3, 0.5), np.zeros(3))
    prof = af.Profile("synthetic", a, cl, cd, np.zeros_like(a))
    airfoils = SimpleNamespace(Names=["round root", "synthetic"], profiles=[root, prof])
    return SimpleNamespace(blade=blade, turbine=turbine, platform=platform, fastout=fo,
                           airfoils=airfoils, wind=_wind(fo), name="synthetic", dt=dt)


def elliptical(b: float = 10.0, AR: float = 6.0, wind=(1.0, 0.0, 0.0), pitch: float = 5.0,
               yaw: float = 0.0, n: int = 20, t=(0.0, 5.0, 10.0)):
    """modeldata/elliptical.m:28-120 (constant pitch): a single wing of span
    b and aspect ratio AR, n cosine-spaced stations, a linear polar
    cl = 2 pi alpha over -90..90 deg."""
    U = np.asarray(wind, dtype=np.float64)
    S = b * b / AR
    c = 4 * S / (math.pi * b)
    theta = np.arange(n + 1) * (math.pi / n) - math.pi
    x = (b / 2) * np.cos(theta)
    tt = np.arccos(2 * x / b)
    y = c * np.sin(tt)
    blade = SimpleNamespace(TipRad=b / 2, HubRad=-b / 2, RTrail=x, ChordTrail=y,
                            AeroTwstTrail=np.zeros(x.size))
    _blade_nodes(blade, np.ones(x.size, dtype=int))
    turbine = SimpleNamespace(NumBl=1, OverHang=0.0, TowerHt=0.0, Twr2Shft=0.0, ShftTilt=0.0,
                              PreCone=np.zeros(1), TowerBaseDiam=0.0, TowerTopDiam=0.0,
                              ellip=SimpleNamespace(b=b, S=S, AR=AR))
    platform = SimpleNamespace(Type="EllipticalWing", TwrDraft=0.0, PtfmCm=0.0, PtfmRef=0.0,
                               PtfmDraft=0.0, PtfmDiam=0.0)
    time = _time(t)
    fo = _rigid_fastout(time, 0.0, 1, pitch, 0.0, U)
    fo.Azimuth = np.full(time.size, -90.0)
    fo.PtfmYaw = np.full(time.size, float(yaw))
    fo.TipSpdRat = np.ones(time.size)
    fo.RotSpeed = np.zeros(time.size)
    aoa = np.arange(101) * (math.pi / 100) - math.pi / 2
    aoa = aoa * 180 / math.pi
    prof = af.Profile("Analytical", aoa, 2 * math.pi * aoa * math.pi / 180, np.zeros(aoa.size),
                      np.zeros(aoa.size))
    airfoils = SimpleNamespace(Names=["Analytical"], profiles=[prof])
    return SimpleNamespace(blade=blade, turbine=turbine, platform=platform, fastout=fo,
                           airfoils=airfoils, wind=_wind(fo), name="elliptical")


class Tables(SimpleNamespace):
    """What winds.march.March takes: the case's blade, airfoils and per-time
    kinematic tables in the reference's layouts (include/winds_march.h)."""


def tables(case, cl_init=None, ll_extra=None, tower: bool = False, up_down: int = 1) -> Tables:
    """kinematics.m + velocity.m on the case, and the initial lift
    coefficient: BEM.m's at the first time index (initials.m:116-119), or
    for the elliptical wing cl = 2 pi alpha / (1 + 2/AR) at the geometric
    angle of attack (initials.m:111-113).  tower=True adds the tower's
    induced velocity at the lifting line (tower_shadow.m, user.towerflag;
    up_down = user.towerupdown: 1 upwind, 2 downwind rotor) to ll_extra."""
    from . import tower as tw
    blade, turbine, platform, fo = case.blade, case.turbine, case.platform, case.fastout
    pos = kin.kinematics(blade, turbine, platform, fo)
    vel, pos = kin.velocity(pos, blade, turbine, case.wind, fo)
    if tower:
        tt = tw.tower_table(pos, case.wind, turbine, up_down)
        ll_extra = tt if ll_extra is None else np.asfortranarray(ll_extra + tt)
    nb = int(turbine.NumBl)
    nt = fo.Time.size
    if cl_init is None:
        if platform.Type == "EllipticalWing":
            aoa = pos.aoag[:, 0, 0, 0]
            cl_init = 2 * math.pi / (1 + 2 / turbine.ellip.AR) * aoa * math.pi / 180
            bres = None
        else:
            bres = bem_mod.bem(case.airfoils, blade, turbine, fo, vel)
            cl_init = bres.cl[:, 0]
    else:
        bres = None
    profs = case.airfoils.profiles
    pitch = np.asarray(fo.BldPitch, dtype=np.float64).reshape(nt, -1)
    if pitch.shape[1] == 1:
        pitch = np.repeat(pitch, nb, axis=1)
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
    return Tables(
        name=getattr(case, "name", ""), nb=nb, nst=blade.RTrail.size, ns=blade.RNodes.size,
        n_times=nt, tip_radius=float(blade.TipRad), hub_radius=float(blade.HubRad),
        r_node=F(blade.RNodes), chord_node=F(blade.Chord), twist_node=F(blade.AeroTwst),
        dr_node=F(blade.DRNodes), foil=np.ascontiguousarray(blade.NFoil, dtype=np.int32),
        cl_init=F(cl_init),
        foil_len=np.array([p.AoA.size for p in profs], dtype=np.int32),
        alpha=F(np.concatenate([p.AoA for p in profs])),
        cl=F(np.concatenate([p.Cl for p in profs])),
        cd=F(np.concatenate([p.Cd for p in profs])),
        time=F(fo.Time), wind=F(case.wind.infty), tsr=F(fo.TipSpdRat),
        rot_speed=F(fo.RotSpeed), azimuth=F(fo.Azimuth), pitch=F(pitch), hub=F(pos.hub),
        quarter=F(pos.quarter), trail=F(pos.trail), bound=F(pos.bound),
        bxn=F(pos.nodes.bxn), byn=F(pos.nodes.byn), bzn=F(pos.nodes.bzn), vblade=F(vel.blade),
        ll_extra=None if ll_extra is None else F(ll_extra),
        dt=float(fo.Time[1] - fo.Time[0]), pos=pos, vel=vel, bem=bres)
