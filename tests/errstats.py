"""Error metrics for induced-velocity fields.

Per-target relative error e_i = |u_i - ref_i| / |ref_i| over targets with
|ref_i| >= floor * max_j |ref_j| (the relative error is meaningless where the
reference velocity vanishes); reported as max and RMS, plus the normwise
max_i |u_i - ref_i| / max_j |ref_j|.
"""
import numpy as np


def as_vectors(u):
    """[ns,3,nt,nb] or [3,n] -> [n,3]."""
    u = np.asarray(u)
    if u.ndim == 2 and u.shape[0] == 3:
        return u.T
    return np.moveaxis(u, 1, -1).reshape(-1, 3)


def stats(u, ref, floor=1e-6):
    a, b = as_vectors(u), as_vectors(ref)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    m = nb >= floor * nb.max()
    e = d[m] / nb[m]
    return {
        "max_rel": float(e.max()) if e.size else 0.0,
        "rms_rel": float(np.sqrt(np.mean(e * e))) if e.size else 0.0,
        "normwise": float(d.max() / nb.max()) if nb.max() > 0 else float(d.max()),
        "n": int(m.sum()),
    }
