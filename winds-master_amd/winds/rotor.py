"""Rigid rotor description for the free-wake time march.

Plays the role of the reference's case files (modeldata/NRELrotor.m:29-55:
blade.RTrail, ChordTrail, AeroTwstTrail, RNodes = midpoints, Chord/AeroTwst
interpolated at RNodes, DRNodes = diff(RTrail)) with a synthetic blade and a
synthetic airfoil polar, so every input is generated here and no case data
files are needed.

Frames (rotor at the origin, rotor plane y-z, wind along +x before yaw):
blade b at azimuth psi = psi_b + Omega t has radial unit e_r = (0, cos psi,
sin psi) and direction of motion e_t = (0, -sin psi, cos psi); a section of
twist+pitch theta has its chord (leading to trailing edge) along
e_c = -cos(theta) e_t + sin(theta) e_x.  The quarter-chord line is r e_r, the
trailing edge r e_r + 0.75 c e_c (core/kinematics.m:46-53: blade_quarter,
blade_trail, blade_bound, blade_end), and the blade frame is bxn = e_c,
bzn = e_r, byn = bzn x bxn (core/velocity.m:58-63).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import gen


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


@dataclass
class Rotor:
    nb: int
    r_trail: np.ndarray      # [nst] trailing-node radii
    chord_trail: np.ndarray  # [nst]
    twist_trail: np.ndarray  # [nst] degrees
    alpha: np.ndarray        # polar angle of attack, degrees
    cl: np.ndarray
    cd: np.ndarray
    R: float
    hub: float
    tsr: float
    U: float
    yaw_deg: float = 0.0
    pitch_deg: float = 0.0
    dpsi_deg: float = 5.0

    @property
    def nst(self) -> int:
        return self.r_trail.size

    @property
    def ns(self) -> int:
        return self.r_trail.size - 1

    @property
    def omega(self) -> float:
        return self.tsr * self.U / self.R

    @property
    def dt(self) -> float:
        return math.radians(self.dpsi_deg) / self.omega

    @property
    def uinf(self) -> np.ndarray:
        y = math.radians(self.yaw_deg)
        return np.array([self.U * math.cos(y), self.U * math.sin(y), 0.0])

    @property
    def r_node(self) -> np.ndarray:
        return self.r_trail[:-1] + np.diff(self.r_trail) / 2

    @property
    def chord_node(self) -> np.ndarray:
        return np.interp(self.r_node, self.r_trail, self.chord_trail)

    @property
    def twist_node(self) -> np.ndarray:
        return np.interp(self.r_node, self.r_trail, self.twist_trail)

    @property
    def dr_node(self) -> np.ndarray:
        return np.abs(np.diff(self.r_trail))


def synthetic_rotor(nb: int = 3, nst: int = 31, R: float = 63.0, hub: float = 1.5,
                    tsr: float = 7.0, U: float = 11.4, yaw_deg: float = 0.0,
                    dpsi_deg: float = 5.0) -> Rotor:
    """Cosine-spaced stations; chord 4.6 m at the hub tapering to 1.4 m at the
    tip; twist 13.3 deg (1 - xi)^2."""
    r = gen.cosine_stations(nst, R, hub)
    xi = (r - hub) / (R - hub)
    a, cl, cd = synthetic_polar()
    return Rotor(nb=nb, r_trail=r, chord_trail=4.6 - 3.2 * xi, twist_trail=13.3 * (1 - xi) ** 2,
                 alpha=a, cl=cl, cd=cd, R=R, hub=hub, tsr=tsr, U=U, yaw_deg=yaw_deg,
                 dpsi_deg=dpsi_deg)
