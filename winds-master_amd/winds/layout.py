"""Host-side conversions between MATLAB-shaped arrays and the device layout.

Segment k of a [ns,3,nt,nb] endpoint pair is k = m + n*ns + p*ns*nt, the
order of crunch_f / crunch_rg (cuda/source/BiotSavartMex.cu:316-357); target j
of P follows crunch_p (:361-379).  Device records are 8 doubles per segment
(ax ay az bx by bz gamma rc) and targets are SoA [3, n].
"""
from __future__ import annotations

import numpy as np


def segments(F1, F2, gamma, rc) -> np.ndarray:
    """[n, 8] segment records from [ns,3,nt,nb] endpoints and [ns,1,nt,nb] data."""
    F1 = np.asarray(F1, dtype=np.float64)
    F2 = np.asarray(F2, dtype=np.float64)
    n = F1.shape[0] * int(np.prod(F1.shape[2:]))
    out = np.empty((n, 8))
    for c in range(3):
        out[:, c] = F1[:, c].reshape(-1, order="F")
        out[:, 3 + c] = F2[:, c].reshape(-1, order="F")
    out[:, 6] = np.asarray(gamma, dtype=np.float64).reshape(-1, order="F")
    out[:, 7] = np.asarray(rc, dtype=np.float64).reshape(-1, order="F")
    return out


def targets(P) -> np.ndarray:
    """[3, n] SoA targets from a [ns,3,nt,nb] array."""
    P = np.asarray(P, dtype=np.float64)
    return np.stack([P[:, c].reshape(-1, order="F") for c in range(3)])


def uind_to_cm(u3: np.ndarray, pshape) -> np.ndarray:
    """[3, n] SoA velocities back to MATLAB's [ns,3,nt,nb] (reassem_uind)."""
    ns = pshape[0]
    rest = tuple(pshape[2:])
    out = np.empty((ns, 3) + rest, order="F")
    for c in range(3):
        out[:, c] = u3[c].reshape((ns,) + rest, order="F")
    return out


def wake_segments(wake, chained: bool = False) -> np.ndarray:
    """Shed then trailing segments of a RotorWake, WInDS.m:150-154.

    chained=False keeps MATLAB's column-major order of each call.  chained=True
    lists the trailing filaments line by line (blade, station, age from the
    oldest down), the order the device march packs them in: consecutive
    records are neighbouring filaments, as the shed filaments of one row
    already are.  The set of segments is the same."""
    shed = segments(*wake.shed_segments())
    trail = segments(*wake.trail_segments())
    if chained:
        nst, _, na, nb = np.shape(wake.gamma_trail)
        t3 = trail.reshape(nb, na, nst, 8)                 # [b][a][m]
        trail = np.ascontiguousarray(t3[:, ::-1, :, :].transpose(0, 2, 1, 3).reshape(-1, 8))
    return np.concatenate([shed, trail])


def wake_targets(wake) -> np.ndarray:
    return targets(wake.domain)
