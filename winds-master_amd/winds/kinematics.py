"""Rotor kinematics: inertial positions, velocities and blade frames of the
lifting line and the wake's first two node rows at every time index.

Host-side restatement of the reference's preprocessing (the MATLAB functions
run once before the time loop, WInDS.m:112-116):

  dcm_rot     core/DCMRot.m:31-89     direction cosines of a rotation sequence
  kinematics  core/kinematics.m:27-109 positions of the blade stations
  velocity    core/velocity.m:30-73   blade-station velocities and frames
  ctdiff      numerical/ctdiff.m:24-44 time derivative by finite differences

Arrays keep the reference's shapes and column-major meaning: positions
[n, 3, nt, nb], time series [nt] or [nt, k].  The device time march reads the
results as tables (winds.march.Tables), so any kinematics a caller can state
-- platform surge/sway/heave/roll/pitch/yaw, nacelle yaw, shaft tilt, precone,
time-varying pitch and wind -- drives the same march.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np


def sind(x):
    """sin of degrees, exact at multiples of 90 degrees (MATLAB's sind)."""
    x = np.asarray(x, dtype=np.float64)
    r = np.mod(x, 360.0)
    out = np.sin(np.radians(r))
    out = np.where(r == 0.0, 0.0, out)
    out = np.where(r == 90.0, 1.0, out)
    out = np.where(r == 180.0, 0.0, out)
    out = np.where(r == 270.0, -1.0, out)
    return out


def cosd(x):
    """cos of degrees, exact at multiples of 90 degrees (MATLAB's cosd)."""
    return sind(np.asarray(x, dtype=np.float64) + 90.0)


def _axis_matrix(axis: str, ang):
    """[n, 3, 3] rotation about x, y or z by ang degrees (DCMRot.m:45-51)."""
    s, c = sind(ang), cosd(ang)
    n = s.shape[0]
    R = np.zeros((n, 3, 3))
    if axis == "x":
        R[:, 0, 0] = 1.0
        R[:, 1, 1], R[:, 1, 2] = c, -s
        R[:, 2, 1], R[:, 2, 2] = s, c
    elif axis == "y":
        R[:, 0, 0], R[:, 0, 2] = c, s
        R[:, 1, 1] = 1.0
        R[:, 2, 0], R[:, 2, 2] = -s, c
    elif axis == "z":
        R[:, 0, 0], R[:, 0, 1] = c, -s
        R[:, 1, 0], R[:, 1, 1] = s, c
        R[:, 2, 2] = 1.0
    else:
        raise ValueError(f"unknown rotation axis {axis!r}")
    return R


def dcm_rot(x, t, rotseq: str, A=None, rev: bool = False):
    """core/DCMRot.m: y = A x with A = R_n ... R_1 for the rotation sequence
    rotseq (one axis letter per column of t, angles in degrees, one row per
    time), optionally starting from A and transposed (rev).  x is [3], [1, 3]
    or [nt, 3]; returns [nt, 3]."""
    t = np.atleast_2d(np.asarray(t, dtype=np.float64))
    if t.shape[1] != len(rotseq):
        raise ValueError("one rotation angle column per axis letter")
    n = t.shape[0]
    M = np.broadcast_to(np.eye(3), (n, 3, 3)).copy() if A is None else np.array(A, dtype=np.float64)
    for k, ax in enumerate(rotseq.lower()):
        M = np.einsum("nij,njk->nik", _axis_matrix(ax, t[:, k]), M)
    if rev:
        M = np.transpose(M, (0, 2, 1))
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    if x.shape[0] < n:
        x = np.broadcast_to(x, (n, 3))
    elif x.shape[0] > n:
        M = np.broadcast_to(M, (x.shape[0], 3, 3))
    return np.einsum("nij,nj->ni", M, x)


def ctdiff(x, y):
    """numerical/ctdiff.m: d y / d x along time (rows of a 2-D y, the third
    dimension of a 4-D y) on the uniform grid x: second-order one-sided at the
    ends, second-order centred next to them, fourth-order centred inside."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    dx = abs(np.mean(np.diff(x)))
    ax = 0 if y.ndim == 2 else 2
    n = y.shape[ax]
    if n < 5:
        raise ValueError("ctdiff needs at least 5 time levels")
    Y = np.moveaxis(y, ax, 0)
    dy = np.zeros_like(Y)
    dy[0] = (-Y[2] + 4 * Y[1] - 3 * Y[0]) / (2 * dx)
    dy[1] = (Y[2] - Y[0]) / (2 * dx)
    dy[2:-2] = (-Y[4:] + 8 * Y[3:-1] - 8 * Y[1:-3] + Y[:-4]) / (12 * dx)
    dy[-2] = (Y[-1] - Y[-3]) / (2 * dx)
    dy[-1] = (3 * Y[-1] - 4 * Y[-2] + Y[-3]) / (2 * dx)
    return np.moveaxis(dy, 0, ax)


def _col(a, n):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    return np.broadcast_to(a, (n,)) if a.size == 1 else a


def kinematics(blade, turbine, platform, fastout):
    """core/kinematics.m:27-109: inertial positions of the blade stations.

    Returns pos with platform [nt,3], hub [nt,3], and lead/quarter/trail
    [nst,3,nt,nb], bound/colloc/end [ns,3,nt,nb]."""
    T = np.asarray(fastout.Time, dtype=np.float64).reshape(-1)
    nt = T.size
    nb = int(turbine.NumBl)
    pos = SimpleNamespace()
    pos.platform = np.stack([_col(fastout.PtfmSurge, nt), _col(fastout.PtfmSway, nt),
                             _col(fastout.PtfmHeave, nt)], axis=1)
    tilt = float(turbine.ShftTilt)
    hub_nominal = np.array([turbine.OverHang * float(cosd(tilt)), 0.0,
                            platform.PtfmRef + turbine.TowerHt + turbine.Twr2Shft +
                            turbine.OverHang * float(sind(tilt))])
    hub_rotseq = np.stack([_col(fastout.PtfmYaw, nt), _col(fastout.PtfmPitch, nt),
                           _col(fastout.PtfmRoll, nt), _col(fastout.NacYaw, nt)], axis=1)
    pos.hub = pos.platform + dcm_rot(hub_nominal, hub_rotseq, "zyxz")

    RT = np.asarray(blade.RTrail, dtype=np.float64).reshape(-1)
    CT = np.asarray(blade.ChordTrail, dtype=np.float64).reshape(-1)
    TT = np.asarray(blade.AeroTwstTrail, dtype=np.float64).reshape(-1)
    RN = np.asarray(blade.RNodes, dtype=np.float64).reshape(-1)
    CN = np.asarray(blade.Chord, dtype=np.float64).reshape(-1)
    TN = np.asarray(blade.AeroTwst, dtype=np.float64).reshape(-1)
    nst, ns = RT.size, RN.size
    z_st, z_ns = np.zeros(nst), np.zeros(ns)
    local = {
        "lead": np.stack([-0.25 * CT, z_st, RT], 1),
        "quarter": np.stack([z_st, z_st, RT], 1),
        "trail": np.stack([0.75 * CT, z_st, RT], 1),
        "bound": np.stack([z_ns, z_ns, RN], 1),
        "colloc": np.stack([0.25 * CN, z_ns, RN], 1),
        "end": np.stack([0.75 * CN, z_ns, RN], 1),
    }
    rotor_rotseq = np.concatenate([_col(fastout.Azimuth, nt)[:, None], np.full((nt, 1), tilt),
                                   hub_rotseq[:, ::-1]], axis=1)
    elliptical = getattr(platform, "Type", "") == "EllipticalWing"
    rseq = "zzyxxyzxyz" if elliptical else "zzzyxxyzxyz"
    pitch = np.asarray(fastout.BldPitch, dtype=np.float64).reshape(nt, -1)
    if pitch.shape[1] == 1:
        pitch = np.repeat(pitch, nb, axis=1)
    precone = np.asarray(turbine.PreCone, dtype=np.float64).reshape(-1)
    if precone.size == 1:
        precone = np.repeat(precone, nb)
    az_step = 360.0 / nb
    Az = np.concatenate([[0.0], np.cumsum(np.full(nb - 1, az_step))])
    for name in local:
        n = nst if name in ("lead", "quarter", "trail") else ns
        setattr(pos, name, np.zeros((n, 3, nt, nb), order="F"))
    pos.blade_rotseq = []
    for b in range(nb):
        cols = [pitch[:, b:b + 1], np.full((nt, 1), precone[b]), np.full((nt, 1), Az[b]),
                rotor_rotseq]
        if not elliptical:
            cols = [np.full((nt, 1), 90.0)] + cols
        bseq = np.concatenate(cols, axis=1)
        pos.blade_rotseq.append(bseq)
        for name, vecs in local.items():
            twist = TT if name in ("lead", "quarter", "trail") else TN
            out = getattr(pos, name)
            for i in range(vecs.shape[0]):
                seq = np.concatenate([np.full((nt, 1), twist[i]), bseq], axis=1)
                out[i, :, :, b] = (dcm_rot(vecs[i], seq, rseq) + pos.hub).T
    return pos


def _unit(v, axis=1):
    # a zero vector (the elliptical wing's zero-chord tips) gives NaN, as in
    # the reference; those frames are never used by the march
    with np.errstate(invalid="ignore", divide="ignore"):
        return v / np.sqrt(np.sum(v * v, axis=axis, keepdims=True))


def velocity(pos, blade, turbine, wind, fastout):
    """core/velocity.m:30-73: kinematic inflow at the lifting line and the
    blade frames.  Adds pos.nodes (bxt, byt, bzt at the trailing nodes; bxn,
    byn, bzn at the bound nodes, [n,3,nt,nb]) and pos.aoag (geometric angle
    of attack, [ns,1,nt,nb]); returns vel with platform/hub/relhub [nt,3],
    bound (inflow = free stream - blade motion) and blade (the same in the
    blade frame) [ns,3,nt,nb]."""
    T = np.asarray(fastout.Time, dtype=np.float64).reshape(-1)
    winf = np.asarray(wind.infty, dtype=np.float64).reshape(T.size, 3)
    vel = SimpleNamespace()
    vel.platform = ctdiff(T, pos.platform)
    vel.hub = ctdiff(T, pos.hub)
    vel.relhub = vel.platform + vel.hub + winf
    vb = ctdiff(T, pos.bound)
    vel.bound = -vb + np.transpose(winf)[None, :, :, None]
    nodes = SimpleNamespace()
    nodes.bxt = _unit(pos.trail - pos.quarter)
    bzt = _unit(np.diff(pos.quarter, axis=0))
    nodes.bzt = np.concatenate([bzt[:1], bzt], axis=0)
    nodes.byt = np.cross(nodes.bzt, nodes.bxt, axis=1)
    nodes.bxn = _unit(pos.end - pos.bound)
    bzn = _unit(np.diff(pos.bound, axis=0))
    nodes.bzn = np.concatenate([bzn[:1], bzn], axis=0)
    nodes.byn = np.cross(nodes.bzn, nodes.bxn, axis=1)
    pos.nodes = nodes
    vel.blade = np.stack([np.sum(nodes.bxn * vel.bound, axis=1),
                          np.sum(nodes.byn * vel.bound, axis=1),
                          np.sum(nodes.bzn * vel.bound, axis=1)], axis=1)
    aoag = np.degrees(np.arctan2(-vel.blade[:, 1:2], vel.blade[:, 0:1]))
    aoag[np.isnan(aoag) | (np.abs(aoag) == 180.0)] = 0.0
    pos.aoag = aoag
    return vel, pos
