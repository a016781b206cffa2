"""Seeded synthetic rotor wakes in the reference's array layout.

The reference stores one time level of the wake lattice as
``wake.domain{t}`` of shape ``[nst, 3, ages, nb]`` (core/initials.m:44-45,87),
circulations as ``gamma.shed{t}`` ``[ns, 1, ages, nb]`` and
``gamma.trail{t}`` ``[nst, 1, ages-1, nb]`` (initials.m:73-76,105-106), and
squared core radii with the same shapes as the circulations.  The callers
build segments from the lattice as

* shed:  F1 = domain(1:end-1,:,:,:), F2 = domain(2:end,:,:,:)   (WInDS.m:150)
* trail: F1 = domain(:,:,2:end,:),   F2 = domain(:,:,1:end-1,:) (WInDS.m:153)

and evaluate at every node of the lattice.  This module produces such states
for a helical rotor wake, deterministic from integer seeds (SplitMix64 words,
Box-Muller normals), all arrays float64 in Fortran (column-major) order.

Circulation follows the reference's Kelvin bookkeeping (WInDS.m:131-140):
trail(:,k) = diff([0; G(:,k); 0]) and shed(:,2:end) = diff of the span
cumulative sum along age, where G(:,k) is the bound circulation when trailing
age k was shed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs for the given seed (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller over pairs of uniform01 draws."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2.0 * math.pi * u2)
    z[1::2] = rad * np.sin(2.0 * math.pi * u2)
    return z[:n]


@dataclass
class RotorWake:
    """One time level of a rotor wake in the reference's layout."""

    domain: np.ndarray       # [nst, 3, ages, nb]
    gamma_shed: np.ndarray   # [ns, 1, ages, nb]
    gamma_trail: np.ndarray  # [nst, 1, ages-1, nb]
    rc_shed: np.ndarray      # [ns, 1, ages, nb]   (r_c^2)
    rc_trail: np.ndarray     # [nst, 1, ages-1, nb] (r_c^2)
    dt: float
    omega: float
    params: dict = field(default_factory=dict)

    @property
    def nst(self) -> int:
        return self.domain.shape[0]

    @property
    def ages(self) -> int:
        return self.domain.shape[2]

    @property
    def nb(self) -> int:
        return self.domain.shape[3]

    @property
    def n_targets(self) -> int:
        return self.nst * self.ages * self.nb

    @property
    def n_segments(self) -> int:
        return (self.nst - 1) * self.ages * self.nb + self.nst * (self.ages - 1) * self.nb

    def shed_segments(self):
        """(F1, F2, gamma, rc) of the shed filaments, WInDS.m:150-151."""
        d = self.domain
        return (np.asfortranarray(d[:-1]), np.asfortranarray(d[1:]),
                self.gamma_shed, self.rc_shed)

    def trail_segments(self):
        """(F1, F2, gamma, rc) of the trailing filaments, WInDS.m:153-154."""
        d = self.domain
        return (np.asfortranarray(d[:, :, 1:]), np.asfortranarray(d[:, :, :-1]),
                self.gamma_trail, self.rc_trail)


def cosine_stations(nst: int, R: float, hub: float) -> np.ndarray:
    j = np.arange(nst, dtype=np.float64)
    return hub + (R - hub) * (1.0 - np.cos(math.pi * j / (nst - 1))) / 2.0


def bound_gamma(r_mid: np.ndarray, R: float, hub: float, peak: float,
                noise: float, ages: int, nb: int, seed: int) -> np.ndarray:
    """Elliptic bound circulation with seeded +-noise, shape [ns, ages, nb]."""
    xi = (2.0 * r_mid - R - hub) / (R - hub)
    base = peak * np.sqrt(np.clip(1.0 - xi * xi, 0.0, None))
    ns = r_mid.size
    u = 2.0 * uniform01(seed, ns * ages * nb) - 1.0
    u = u.reshape((ns, ages, nb), order="F")
    return base[:, None, None] * (1.0 + noise * u)


def kelvin_from_bound(G: np.ndarray):
    """Trailing and shed circulation from the bound history G[ns, A-1, nb].

    trail(:,k) = diff([0; G(:,k); 0]) (WInDS.m:133-136);
    shed(:,1) = G(:,1); shed(:,2:end) = diff(cat(3, cumsum(trail(1:end-1)), 0), 1, 3)
    (WInDS.m:139-140).
    """
    ns, na, nb = G.shape
    z = np.zeros((1, na, nb))
    trail = np.diff(np.concatenate([z, G, z], axis=0), axis=0)  # [nst, A-1, nb]
    C = np.cumsum(trail[:-1], axis=0)                            # [ns, A-1, nb]
    shed = np.empty((ns, na + 1, nb))
    shed[:, 0] = G[:, 0]
    shed[:, 1:] = np.diff(np.concatenate([C, np.zeros((ns, 1, nb))], axis=1), axis=1)
    return trail, shed


def rotor_wake(nb: int = 3, nst: int = 18, ages: int = 72, R: float = 63.0,
               hub: float = 1.5, tsr: float = 7.0, U: float = 11.4,
               a: float = 0.3, dpsi_deg: float = 5.0, jitter: float = 0.01,
               gamma_peak: float = 60.0, gamma_noise: float = 0.05,
               r0: float = 0.05, alpha: float = 1.25613, delta: float = 100.0,
               nu: float = 1e-5, yaw_deg: float = 0.0, seed_pos: int = 1,
               seed_gamma: int = 2, psi0: float = 0.0) -> RotorWake:
    """Helical wake of a rigid rotor with seeded jitter and circulation noise.

    Node (j, k, b) sits at station radius r_j, azimuth
    psi_b - k*Omega*dt with psi_b = psi0 + 2 pi b / nb, and
    axial distance U(1-a)*k*dt along the convection direction, which is the
    +x axis yawed by yaw_deg about z.  Every coordinate gets N(0, jitter*R).
    Core: r_c^2 = r0^2 + 4*alpha*delta*nu*t_age, with t_age = k*dt for shed
    filaments and (k+1/2)*dt for trailing filaments.
    """
    omega = tsr * U / R
    dt = math.radians(dpsi_deg) / omega
    r = cosine_stations(nst, R, hub)
    k = np.arange(ages, dtype=np.float64)
    psi_b = psi0 + 2.0 * math.pi * np.arange(nb) / nb
    psi = psi_b[None, :] - (k * omega * dt)[:, None]            # [ages, nb]
    s = U * (1.0 - a) * k * dt                                   # [ages]
    yaw = math.radians(yaw_deg)
    dom = np.empty((nst, 3, ages, nb))
    dom[:, 0] = (s * math.cos(yaw))[None, :, None] + 0.0 * psi[None]
    dom[:, 1] = r[:, None, None] * np.cos(psi)[None] + (s * math.sin(yaw))[None, :, None]
    dom[:, 2] = r[:, None, None] * np.sin(psi)[None]
    if jitter > 0.0:
        dom += (jitter * R) * normal(seed_pos, dom.size).reshape(dom.shape, order="F")
    dom = np.asfortranarray(dom)

    r_mid = 0.5 * (r[:-1] + r[1:])
    G = bound_gamma(r_mid, R, hub, gamma_peak, gamma_noise, ages - 1, nb, seed_gamma)
    trail, shed = kelvin_from_bound(G)
    ns = nst - 1
    gamma_shed = np.asfortranarray(shed.reshape(ns, 1, ages, nb))
    gamma_trail = np.asfortranarray(trail.reshape(nst, 1, ages - 1, nb))

    c = 4.0 * alpha * delta * nu
    t_shed = k * dt
    t_trail = (k[:-1] + 0.5) * dt
    rc_shed = np.broadcast_to((r0 * r0 + c * t_shed)[None, None, :, None],
                              (ns, 1, ages, nb))
    rc_trail = np.broadcast_to((r0 * r0 + c * t_trail)[None, None, :, None],
                               (nst, 1, ages - 1, nb))
    return RotorWake(domain=dom, gamma_shed=gamma_shed, gamma_trail=gamma_trail,
                     rc_shed=np.asfortranarray(rc_shed), rc_trail=np.asfortranarray(rc_trail),
                     dt=dt, omega=omega,
                     params=dict(nb=nb, nst=nst, ages=ages, R=R, hub=hub, tsr=tsr, U=U,
                                 a=a, yaw_deg=yaw_deg, seed_pos=seed_pos,
                                 seed_gamma=seed_gamma, psi0=psi0))


def config1() -> RotorWake:
    """3 blades x 18 cosine stations x 72 ages (1 rev at 5 deg): 3,888 nodes."""
    return rotor_wake(nb=3, nst=18, ages=72)


def config2() -> RotorWake:
    """3 blades x 31 stations x 1,440 ages (20 revs): 133,920 nodes."""
    return rotor_wake(nb=3, nst=31, ages=1440)


def config3_params(n_wakes: int = 64, seed0: int = 100):
    """Per-wake parameters of the uneven batch: blades in {2,3}, 12-40
    stations, 200-2000 ages, TSR U[4,10], yaw U[0,30] deg; wake i uses seed
    seed0+i for its draws, positions and circulation."""
    out = []
    for i in range(n_wakes):
        s = seed0 + i
        u = uniform01(1000003 * s + 17, 5)
        out.append(dict(nb=2 + int(u[0] * 2), nst=12 + int(u[1] * 29),
                        ages=200 + int(u[2] * 1801), tsr=4.0 + 6.0 * u[3],
                        yaw_deg=30.0 * u[4], seed_pos=2 * s + 1, seed_gamma=2 * s + 2))
    return out


def config3(n_wakes: int = 64, seed0: int = 100):
    return [rotor_wake(**p) for p in config3_params(n_wakes, seed0)]
