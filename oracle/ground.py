"""CPU restatement of the WInDS ground-effect vortex-panel model.

TEST INFRASTRUCTURE ONLY (the checker of winds.ground); the engine never
imports it.  Follows the reference files step by step, vectorised over the
points as they are:

  nodes            ground/Main_Driver.m:7-21 (meshgrid of Np+1 nodes per side
                   on z = 0; here lo + k (hi - lo) / Np)
  matrix + solve   ground/VP_Influence_Coefficient_Matrix_v7.m:9-100
                   (control points at the panel centres, row i = k + (j-1) Np;
                   ring j's four segment terms t_k c_k added 1-2-3-4, /4/pi;
                   normal n_j = c2(1,:)/|c2(1,:)| at control point 1;
                   A(:, j) = ring . n_j; V_n = -V_ext . n; Gamma = inv(A) V_n)
  induced          ground/VP_Induced_Velocity_LM.m:36-81 (V_ind accumulated
                   over panels in order, Gamma(j) * ring)

Arrays: rows are points ([N, 3]) as in the reference.
"""
from __future__ import annotations

import numpy as np


def nodes(np_side, lo, hi):
    return lo + np.arange(np_side + 1) * ((hi - lo) / np_side)


def centres(xp, yp):
    """VP_Influence_Coefficient_Matrix_v7.m:11-18 (z = 0)."""
    Np = xp.size - 1
    P = np.zeros((Np * Np, 3))
    i = 0
    for j in range(Np):
        for k in range(Np):
            P[i, 0] = xp[k] + 0.5 * (xp[k + 1] - xp[k])
            P[i, 1] = yp[j] + 0.5 * (yp[j + 1] - yp[j])
            P[i, 2] = 0.0 + 0.5 * (0.0 - 0.0)
            i += 1
    return P


def _terms(P, xp, yp, n, m, co):
    """The four segment terms A1..A4 of ring (n, m) at points P [N, 3] and
    c2 (VP_*.m:62-86 / :47-71)."""
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    corners = ((xp[m], yp[n]), (xp[m + 1], yp[n]), (xp[m + 1], yp[n + 1]), (xp[m], yp[n + 1]))
    r = [np.stack([x - cx, y - cy, z - 0.0], axis=1) for cx, cy in corners]
    nn = [np.sqrt(ri[:, 0] * ri[:, 0] + ri[:, 1] * ri[:, 1] + ri[:, 2] * ri[:, 2]) for ri in r]
    A, C = [], []
    for k in range(4):
        a, b = k, (k + 1) % 4
        ra, rb = r[a], r[b]
        dot = ra[:, 0] * rb[:, 0] + ra[:, 1] * rb[:, 1] + ra[:, 2] * rb[:, 2]
        t = (nn[a] + nn[b]) / (nn[a] * nn[b] * (nn[a] * nn[b] + dot) + co)
        c = np.stack([ra[:, 1] * rb[:, 2] - ra[:, 2] * rb[:, 1],
                      ra[:, 2] * rb[:, 0] - ra[:, 0] * rb[:, 2],
                      ra[:, 0] * rb[:, 1] - ra[:, 1] * rb[:, 0]], axis=1)
        A.append(c * t[:, None])
        C.append(c)
    return A, C[1]


def ring_velocity(P, xp, yp, n, m, co):
    """(A1+A2+A3+A4)/4/pi: ring (n, m)'s velocity per unit circulation."""
    A, _ = _terms(P, xp, yp, n, m, co)
    return (((A[0] + A[1]) + A[2]) + A[3]) / 4 / np.pi


def influence(xp, yp, co_frac):
    """(A, n_vec, P): VP_Influence_Coefficient_Matrix_v7.m:55-95."""
    Np = xp.size - 1
    dx, dy = xp[1] - xp[0], yp[1] - yp[0]
    co = co_frac * (dx + dy) / 2
    P = centres(xp, yp)
    N = Np * Np
    A = np.zeros((N, N))
    nv = np.zeros((N, 3))
    j = 0
    for n in range(Np):
        for m in range(Np):
            T, c2 = _terms(P, xp, yp, n, m, co)
            nv[j] = c2[0] / np.sqrt(c2[0, 0] ** 2 + c2[0, 1] ** 2 + c2[0, 2] ** 2)
            ring = (((T[0] + T[1]) + T[2]) + T[3]) / 4 / np.pi
            A[:, j] = ring[:, 0] * nv[j, 0] + ring[:, 1] * nv[j, 1] + ring[:, 2] * nv[j, 2]
            j += 1
    return A, nv, P


def solve(A, nv, v_ext):
    """Gamma = inv(A) V_n with V_n = -V_ext . n (:89, :99-100)."""
    Vn = -(v_ext[:, 0] * nv[:, 0] + v_ext[:, 1] * nv[:, 1] + v_ext[:, 2] * nv[:, 2])
    return np.linalg.inv(A) @ Vn


def induced(P, gamma, xp, yp, co_frac):
    """VP_Induced_Velocity_LM.m:36-81."""
    Np = xp.size - 1
    dx, dy = xp[1] - xp[0], yp[1] - yp[0]
    co = co_frac * (dx + dy) / 2
    V = np.zeros((P.shape[0], 3))
    j = 0
    for n in range(Np):
        for m in range(Np):
            V = V + gamma[j] * ring_velocity(P, xp, yp, n, m, co)
            j += 1
    return V
