"""Pins of the CPU restatement of the ground-effect panel model (oracle/ground.py).

The reference holds no tests or golden vectors for ground/; the restatement is
pinned by what the mathematics fixes:

- a square vortex ring (no cut-off): its axial velocity at height z is
  Gamma a^2 / (2 pi (z^2 + a^2/4) sqrt(z^2 + a^2/2)), 2 sqrt(2) Gamma / (pi a)
  at the centre, in +z for the 1-2-3-4 (counter-clockwise) ring;
- the ring's velocity at arbitrary points against the textbook straight
  segment, Gamma / (4 pi h) (cos t1 - cos t2) along d x (P - A) (angles to the
  segment, h the distance from its line), summed over the four sides: a sign,
  an index or a transposed cross product fails it;
- uniform circulation over the grid equals one loop round the perimeter
  (inner sides cancel);
- the influence matrix is symmetric (uniform square grid) and positive
  definite, its normals are +z, and its columns are the induced-velocity
  routine at the control points;
- the solved circulation cancels the external normal velocity at every
  control point when the panels' own velocity is evaluated by the separate
  induced-velocity routine (VP_Induced_Velocity_LM.m).
"""
import math

import numpy as np

from oracle import ground as G


def textbook_segment(P, A, B):
    """Velocity per unit circulation of the straight segment A->B at P."""
    d = B - A
    L = np.linalg.norm(d)
    e = d / L
    ra, rb = P - A, P - B
    c1 = ra @ e / np.linalg.norm(ra, axis=1)
    c2 = rb @ e / np.linalg.norm(rb, axis=1)
    perp = np.cross(e, ra)            # e x (P - A): the swirl direction
    h = np.linalg.norm(perp, axis=1)
    return (c1 - c2)[:, None] / (4 * math.pi * h[:, None]) * perp / h[:, None]


def textbook_loop(P, corners):
    v = np.zeros_like(P)
    for k in range(len(corners)):
        v += textbook_segment(P, corners[k], corners[(k + 1) % len(corners)])
    return v


def test_square_ring_on_axis_closed_form():
    a = 2.0
    xp = np.array([-1.0, 1.0])
    yp = np.array([-1.0, 1.0])
    z = np.array([0.0, 0.1, 0.7, 3.0, 25.0])
    P = np.stack([np.zeros_like(z), np.zeros_like(z), z], axis=1)
    v = G.ring_velocity(P, xp, yp, 0, 0, 0.0)
    w = a * a / (2 * math.pi * (z * z + a * a / 4) * np.sqrt(z * z + a * a / 2))
    assert abs(w[0] - 2 * math.sqrt(2) / (math.pi * a)) < 1e-15
    np.testing.assert_allclose(v[:, 2], w, rtol=1e-13)
    assert np.max(np.abs(v[:, :2])) < 1e-16


def test_ring_matches_textbook_segments():
    rng = np.random.default_rng(3)
    xp = np.array([1.0, 2.5])
    yp = np.array([-0.5, 0.3])
    P = rng.uniform(-3, 4, (200, 3))
    v = G.ring_velocity(P, xp, yp, 0, 0, 0.0)
    corners = [np.array([1.0, -0.5, 0.0]), np.array([2.5, -0.5, 0.0]),
               np.array([2.5, 0.3, 0.0]), np.array([1.0, 0.3, 0.0])]
    ref = textbook_loop(P, corners)
    np.testing.assert_allclose(v, ref, rtol=1e-11, atol=1e-13 * np.max(np.abs(ref)))


def test_uniform_circulation_is_the_perimeter_loop():
    rng = np.random.default_rng(5)
    xp = G.nodes(6, -3.0, 3.0)
    yp = G.nodes(6, -2.0, 1.0)
    P = np.column_stack([rng.uniform(-4, 4, 50), rng.uniform(-3, 2, 50), rng.uniform(0.2, 3, 50)])
    v = G.induced(P, np.ones(36), xp, yp, 0.0)
    corners = [np.array([-3.0, -2.0, 0.0]), np.array([3.0, -2.0, 0.0]),
               np.array([3.0, 1.0, 0.0]), np.array([-3.0, 1.0, 0.0])]
    ref = textbook_loop(P, corners)
    assert np.max(np.abs(v - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_influence_matrix_structure():
    xp = G.nodes(8, -40.0, 120.0)
    yp = G.nodes(8, -30.0, 30.0)
    A, nv, P = G.influence(xp, yp, 0.05)
    N = 64
    assert A.shape == (N, N)
    assert np.array_equal(nv, np.tile([0.0, 0.0, 1.0], (N, 1)))
    assert np.max(np.abs(A - A.T)) <= 1e-14 * np.max(np.abs(A))
    assert np.min(np.linalg.eigvalsh(0.5 * (A + A.T))) > 0
    # control points: panel centres in the reference's order (x fastest)
    assert np.allclose(P[1], [xp[1] + 10.0, yp[0] + 3.75, 0.0])
    assert np.allclose(P[8], [xp[0] + 10.0, yp[1] + 3.75, 0.0])
    # columns: ring j's normal velocity at every control point
    for j in (0, 9, 63):
        ring = G.induced(P, np.eye(N)[j], xp, yp, 0.05)
        assert np.allclose(A[:, j], ring[:, 2], rtol=0, atol=1e-15 * np.max(np.abs(A)))


def test_single_panel_matrix_is_the_centre_velocity():
    xp = np.array([0.0, 3.0])
    yp = np.array([0.0, 3.0])
    A, _, _ = G.influence(xp, yp, 0.0)
    assert abs(A[0, 0] - 2 * math.sqrt(2) / (math.pi * 3.0)) < 1e-15


def test_solution_cancels_the_normal_velocity():
    rng = np.random.default_rng(11)
    xp = G.nodes(7, -500.0, 1500.0)
    yp = G.nodes(7, -375.0, 375.0)
    A, nv, P = G.influence(xp, yp, 0.05)
    v_ext = np.column_stack([np.full(49, 11.4), rng.normal(0, 1, 49), rng.normal(0, 2, 49)])
    gam = G.solve(A, nv, v_ext)
    vind = G.induced(P, gam, xp, yp, 0.05)
    net = vind[:, 2] + v_ext[:, 2]
    assert np.max(np.abs(net)) <= 1e-12 * np.max(np.abs(v_ext[:, 2]))
