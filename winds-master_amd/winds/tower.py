"""Tower shadow: the tower's induced velocity at the lifting line.

Host-side restatement of core/tower_shadow.m:1-121, which
KuttaJoukowski.m:232-238 adds to the wake's induced velocity at the bound
points when user.towerflag is 'true'.  The tower is a vertical cylinder on
the z axis from z = 0 to the tower height, its diameter linear from the base
to the top; the flow around it is the potential flow of a doublet of
strength k = 2 pi U R^2 in the free stream U along x:

    u_r = -k cos(theta) / (2 pi r^2),  u_theta = -k sin(theta) / (2 pi r^2)

(r, theta: distance from the tower axis and azimuth about it, measured from
+x), turned into x and y components.  Points above the tower top, below its
base or inside the cylinder get no velocity.  For a downwind rotor
(up_down = 2) the points with y < 5 R and x > R (the reference's wake
region, tower_shadow.m:89) get a deficit of -0.3 U along x instead.  U is the
x component of the free stream at that time (the reference indexes
wind.infty(time) linearly, which reads column 1).

Quirk not reproduced: in tower_shadow.m the downwind deficit is written into
a new third dimension of u_rel for blades 2 and 3 (u_rel(ind,1,nb)), so only
blade 1 receives it; here every blade does.
"""
from __future__ import annotations

import numpy as np


def tower_shadow(P, Uinf: float, up_down: int, turbine):
    """P: bound points [ns, 3, nb]; returns V_ind [ns, 3, 1, nb]."""
    P = np.asarray(P, dtype=np.float64)
    ns, _, nb = P.shape
    V = np.zeros((ns, 3, 1, nb))
    zhub = float(turbine.TowerHt)
    D = np.array([turbine.TowerBaseDiam, turbine.TowerTopDiam], dtype=np.float64)
    zt = np.array([0.0, zhub])
    for b in range(nb):
        x, y, z = P[:, 0, b], P[:, 1, b], P[:, 2, b]
        on = (z <= zhub) & (z >= 0.0)
        if not on.any():
            continue
        x, y, z = x[on], y[on], z[on]
        U = np.full(x.shape, float(Uinf))
        r = np.sqrt(x * x + y * y)
        th = np.arctan2(y, x)
        R = np.interp(z, zt, D) / 2
        k = 2 * np.pi * U * R * R
        with np.errstate(divide="ignore", invalid="ignore"):
            ur = -(k * np.cos(th)) / (2 * np.pi * r * r)
            ut = -(k * np.sin(th)) / (2 * np.pi * r * r)
        u = ur * np.cos(th) - ut * np.sin(th)
        v = ur * np.sin(th) + ut * np.cos(th)
        if up_down == 2:
            dw = (y < 2.5 * R * 2) & (x > R)
            u = np.where(dw, -0.3 * U, u)
            v = np.where(dw, 0.0, v)
        inside = r < R
        u = np.where(inside, 0.0, u)
        v = np.where(inside, 0.0, v)
        idx = np.nonzero(on)[0]
        V[idx, 0, 0, b] = u
        V[idx, 1, 0, b] = v
    return V


def tower_table(pos, wind, turbine, up_down: int = 1):
    """vel.uind_tower at every time index: [ns, 3, nt, nb] from pos.bound
    [ns, 3, nt, nb] and wind.infty [nt, 3] (KuttaJoukowski.m:232-238)."""
    B = np.asarray(pos.bound)
    ns, _, nt, nb = B.shape
    out = np.zeros((ns, 3, nt, nb), order="F")
    inf = np.asarray(wind.infty, dtype=np.float64).reshape(nt, 3)
    for p in range(nt):
        out[:, :, p, :] = tower_shadow(B[:, :, p, :], inf[p, 0], up_down, turbine)[:, :, 0, :]
    return out
