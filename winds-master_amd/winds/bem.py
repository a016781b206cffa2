"""Blade-element momentum solution for the initial circulation.

Host-side restatement of core/BEM.m:41-181, which initials.m:116-119 calls
once to set the lift coefficient of every station at the first time index
(the bound circulation is then 0.5 |U| chord cl, initials.m:124).

Per station and time: inflow angle phi = atan2(U (1-a), Omega r (1+a')),
angle of attack phi - (twist + pitch) with the reference's twist sign
(-AeroTwst), lift and drag from the station's airfoil table, Prandtl tip and
hub losses, the axial induction from momentum balance with the modified
Glauert correction where C_T > 0.96 F, the tangential induction, the skewed
wake correction for yawed or tilted platforms, and under-relaxation 0.1, to
|da|, |da'| <= 1e-6 or 200 iterations.  Where BEM.m takes real() of a
complex intermediate (the square roots of the initial guess and of the
Glauert branch), the real part of the complex result is taken here as well.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .airfoils import interp1


def _sign(x):
    return np.sign(x)


def bem(airfoils, blade, turbine, fastout, vel, tol: float = 1e-6, iters: int = 200,
        wt: float = 0.1):
    """core/BEM.m: returns cl, cd, phi (rad), aoa (deg), a, each [ns, nt], and
    the iteration count."""
    RN = np.asarray(blade.RNodes, dtype=np.float64).reshape(-1)
    chord = np.asarray(blade.Chord, dtype=np.float64).reshape(-1)
    nfoil = np.asarray(blade.NFoil).reshape(-1).astype(int)
    ns = RN.size
    T = np.asarray(fastout.Time, dtype=np.float64).reshape(-1)
    nt = T.size
    Uinf = np.sqrt(np.sum(np.asarray(vel.relhub) ** 2, axis=1))
    Om = np.asarray(fastout.RotSpeed, dtype=np.float64).reshape(-1) * (2 * np.pi / 60)
    twst = -np.asarray(blade.AeroTwst, dtype=np.float64).reshape(-1) * np.pi / 180
    pitch = np.asarray(fastout.BldPitch, dtype=np.float64).reshape(nt, -1)[:, 0]
    ptch = pitch * np.pi / 180
    rP = -np.asarray(fastout.PtfmPitch, dtype=np.float64).reshape(-1) * np.pi / 180
    rY = (np.asarray(fastout.PtfmYaw, dtype=np.float64).reshape(-1) +
          np.asarray(fastout.NacYaw, dtype=np.float64).reshape(-1)) * np.pi / 180
    rP = np.broadcast_to(rP, (nt,))
    rY = np.broadcast_to(rY, (nt,))
    # MATLAB's `if v` on a vector holds when every element is nonzero
    if np.all(_sign(rP) == 0):
        sp = _sign(rY)
    elif np.all(_sign(rY) == 0):
        sp = _sign(rP)
    else:
        sp = _sign(rP) * _sign(rY)
    gamma = sp * np.arccos(np.cos(rP) * np.cos(rY))
    psi = np.pi - np.arctan2(np.cos(rP) * np.sin(rY), np.sin(rP))

    Om = np.tile(Om, (ns, 1))
    Uinf = np.tile(Uinf, (ns, 1))
    ptch = np.tile(ptch, (ns, 1))
    gamma = np.tile(gamma, (ns, 1))
    psi = np.tile(psi, (ns, 1))
    twst = np.tile(twst[:, None], (1, nt))
    sigmap = np.tile((turbine.NumBl * chord / (2 * np.pi * RN))[:, None], (1, nt))
    R = np.tile(RN[:, None], (1, nt))
    lam = Om * R / Uinf

    root = np.lib.scimath.sqrt(4 - 4 * np.pi * lam * sigmap + np.pi * lam ** 2 * sigmap *
                               (8 * (twst + ptch) + np.pi * sigmap))
    a = np.real(0.25 * (2 + np.pi * lam * sigmap - root))
    ap = np.zeros_like(a)
    a0 = np.zeros_like(a)
    ap0 = np.zeros_like(a)
    phi = np.zeros_like(a)
    aoa = np.zeros_like(a)
    cl = np.zeros_like(a)
    cd = np.zeros_like(a)
    ct = np.zeros_like(a)
    f = np.zeros_like(a)
    da = np.ones_like(a)
    dap = np.ones_like(a)
    ncv = (da > tol) | (dap > tol)
    profiles = airfoils.profiles
    j = 0
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        for j in range(1, iters + 1):
            a0[ncv] = a[ncv]
            ap0[ncv] = ap[ncv]
            phi[ncv] = np.arctan2(Uinf[ncv] * (1 - a[ncv]), Om[ncv] * R[ncv] * (1 + ap[ncv]))
            aoa[ncv] = (phi[ncv] - (twst[ncv] + ptch[ncv])) * 180 / np.pi
            for k, prof in enumerate(profiles, start=1):
                rows = nfoil == k
                cl[rows, :] = interp1(prof.AoA, prof.Cl, aoa[rows, :])
                cd[rows, :] = interp1(prof.AoA, prof.Cd, aoa[rows, :])
            s, c = np.sin(phi), np.cos(phi)
            cn = cl * c + cd * s
            ct[ncv] = (sigmap * (1 - a) ** 2 * cn / s ** 2)[ncv]
            ftip = 2 / np.pi * np.arccos(np.exp(-(turbine.NumBl * (blade.TipRad - R) /
                                                  (2 * R * s))))
            fhub = 2 / np.pi * np.arccos(np.exp(-(turbine.NumBl * (R - blade.HubRad) /
                                                  (2 * blade.HubRad * s))))
            f[ncv] = (fhub * ftip)[ncv]
            a[ncv] = np.real((1 + 4 * f * s ** 2 / (sigmap * cn)) ** -1)[ncv]
            ncvf = (ct > 0.96 * f) & ((da > tol) | (dap > tol))
            gl = np.lib.scimath.sqrt(ct * (50 - 36 * f) + 12 * f * (3 * f - 4))
            a[ncvf] = np.real((18 * f - 20 - 3 * gl) / (36 * f - 50))[ncvf]
            ap[ncv] = ((4 * f * c * s / (sigmap * (cl * s - cd * c)) - 1) ** -1)[ncv]
            if np.all(np.abs(gamma) > 1e-8):
                corr = 1 + 15 * np.pi / 32 * R / blade.TipRad * np.tan(0.5 * (0.6 * a + 1) * gamma) * np.cos(psi)
                a[ncv] = (a * corr)[ncv]
            da[ncv] = np.abs(a0[ncv] - a[ncv])
            dap[ncv] = np.abs(ap0[ncv] - ap[ncv])
            if wt > 0:
                a[ncv] = a0[ncv] + wt * (a[ncv] - a0[ncv])
                ap[ncv] = ap0[ncv] + wt * (ap[ncv] - ap0[ncv])
            ncv = (da > tol) | (dap > tol)
            if not ncv.any():
                break
    return SimpleNamespace(cl=cl, cd=cd, phi=phi, aoa=aoa, a=a, ap=ap, iterations=j,
                           converged=not ncv.any())
