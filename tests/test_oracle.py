"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors for the Biot-Savart path (its only
parity check, tests/testbiotmexreal.m:38-46, is commented out and its .mat
fixture is absent), so each oracle function is pinned here by closed forms,
invariants and a __float128 brute force, chosen so that a dropped term, a
wrong sign, a swapped index or a transposed operand fails at least one test.
"""
import math

import numpy as np
import pytest

from oracle import bs as O
from winds import gen, layout


def seg(a, b):
    """One segment as MATLAB-shaped F1, F2 [1,3,1,1]."""
    return (np.asarray(a, float).reshape(1, 3, 1, 1, order="F"),
            np.asarray(b, float).reshape(1, 3, 1, 1, order="F"))


def pts(*p):
    """Targets as [n,3,1,1]."""
    return np.asfortranarray(np.asarray(p, float).reshape(len(p), 3, 1, 1))


def one(x=1.0):
    return np.full((1, 1, 1, 1), x)


def vec(u, i=0):
    return np.array([u[i, 0, 0, 0], u[i, 1, 0, 0], u[i, 2, 0, 0]])


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------

@pytest.mark.parametrize("mode", ["mex", "matlab"])
def test_single_segment_closed_form(mode):
    # segment on the x axis from x=-1.3 to x=2.1, target at (0.4, h, 0)
    G, h = 7.5, 0.8
    a, b, x0 = -1.3, 2.1, 0.4
    F1, F2 = seg([a, 0, 0], [b, 0, 0])
    P = pts([x0, h, 0])
    u = O.bs_mex(F1, F2, P, one(G), 0.0) if mode == "mex" else \
        O.bs_matlab(F1, F2, P, one(G), one(0.01), 0.0)
    # |u| = G/(4 pi h) (cos t1 - cos t2), t measured from +x
    cos1 = (x0 - a) / math.hypot(x0 - a, h)
    cos2 = (x0 - b) / math.hypot(x0 - b, h)
    expect = G / (4 * math.pi * h) * (cos1 - cos2)
    v = vec(u)
    assert v[0] == 0.0 and v[1] == 0.0
    # right-hand rule: dl along +x, target at +y -> +z
    assert v[2] > 0
    assert abs(v[2] - expect) <= 1e-14 * expect


@pytest.mark.parametrize("mode", ["mex", "matlab"])
def test_long_segment_tends_to_infinite_line(mode):
    # (the reference's r1 r2 (r1 r2 + r1.r2) cancels like (L/h)^2 eps, so the
    # half-length stays moderate)
    G, h, a = 3.0, 0.5, 200.0
    F1, F2 = seg([-a, 0, 0], [a, 0, 0])
    P = pts([0, 0, h])  # dl x r: x-hat x z-hat = -y-hat
    u = O.bs_mex(F1, F2, P, one(G), 0.0) if mode == "mex" else \
        O.bs_matlab(F1, F2, P, one(G), one(0.0), 0.0)
    v = vec(u)
    finite = G / (2 * math.pi * h) * a / math.hypot(a, h)
    assert abs(v[1] + finite) < 1e-9 * finite
    assert abs(v[1] + G / (2 * math.pi * h)) < 1e-5 * finite
    assert v[0] == 0.0 and v[2] == 0.0


@pytest.mark.parametrize("N", [3, 4, 7, 64])
def test_ngon_ring_centre(N):
    # regular N-gon ring of circumradius R in the y-z plane, circulating
    # counter-clockwise seen from +x: |u| = N G tan(pi/N) / (2 pi R) along +x
    G, R = 2.0, 1.7
    th = 2 * math.pi * np.arange(N + 1) / N
    ring = np.stack([np.zeros(N + 1), R * np.cos(th), R * np.sin(th)], axis=1)
    F1 = np.asfortranarray(ring[:-1].reshape(N, 3, 1, 1))
    F2 = np.asfortranarray(ring[1:].reshape(N, 3, 1, 1))
    g = np.full((N, 1, 1, 1), G)
    P = pts([0, 0, 0])
    expect = N * G * math.tan(math.pi / N) / (2 * math.pi * R)
    for u in (O.bs_mex(F1, F2, P, g, 0.0), O.bs_matlab(F1, F2, P, g, g * 0, 0.0)):
        v = vec(u)
        assert abs(v[0] - expect) <= 1e-14 * expect
        assert abs(v[1]) < 1e-15 * expect and abs(v[2]) < 1e-15 * expect


def test_cutoff_closed_form_perpendicular_bisector():
    # target on the perpendicular bisector of a segment of length l at
    # distance h: |u| = G 2 rho l h / (4 pi (2 h^2 rho^2 + d l^2)), rho^2 = h^2 + l^2/4
    G, l, h, d = 4.0, 2.0, 0.05, 0.01
    F1, F2 = seg([-l / 2, 0, 0], [l / 2, 0, 0])
    P = pts([0, h, 0])
    rho = math.sqrt(h * h + l * l / 4)
    expect = G * 2 * rho * l * h / (4 * math.pi * (2 * h * h * rho * rho + d * l * l))
    for u in (O.bs_mex(F1, F2, P, one(G), d), O.bs_matlab(F1, F2, P, one(G), one(0), d)):
        assert abs(vec(u)[2] - expect) <= 1e-13 * expect


@pytest.mark.parametrize("n", [1, 2, 3])
def test_vatistas_factor_on_bisector(n):
    # Vatistas-n multiplies the bare kernel by h^2 (rc^n + h^2n)^(-1/n), rc = r_c^2
    G, l, h, rc = 4.0, 2.0, 0.07, 0.05 ** 2
    F1, F2 = seg([-l / 2, 0, 0], [l / 2, 0, 0])
    P = pts([0, h, 0])
    bare = O.bs_matlab(F1, F2, P, one(G), one(rc), 0.0)
    visc = O.bs_matlab(F1, F2, P, one(G), one(rc), f"visc{n}")
    fac = h * h * (rc ** n + h ** (2 * n)) ** (-1.0 / n)
    assert abs(vec(visc)[2] - fac * vec(bare)[2]) <= 1e-13 * abs(vec(bare)[2])
    # and the quad brute force agrees
    q = O.bs_quad(F1, F2, P, one(G), one(rc), f"visc{n}")
    # (h^2 = r1^2 - (L.r1)^2/|L|^2 cancels by r1^2/h^2 ~ 200 here)
    assert abs(vec(q)[2] - vec(visc)[2]) <= 1e-13 * abs(vec(q)[2])


# --------------------------------------------------------------------------
# special cases
# --------------------------------------------------------------------------

def test_endpoint_and_collinear_targets_give_zero():
    F1, F2 = seg([0, 0, 0], [1, 0, 0])
    P = pts([0, 0, 0], [1, 0, 0], [2.5, 0, 0], [-3, 0, 0], [0.5, 0, 0])
    u = O.bs_mex(F1, F2, P, one(5.0), 1e-4)  # d > 0: den stays > 0 at the endpoints
    assert np.all(u == 0.0)
    for d in (0.0, 1e-4, "visc1", "visc2"):
        um = O.bs_matlab(F1, F2, P, one(5.0), one(1e-3), d)
        assert np.all(um == 0.0), d


def test_mex_without_cutoff_keeps_nan_at_endpoint():
    # the MEX path never zeroes (BiotSavartMex.cu:87): d = 0 at an endpoint is 0*inf
    F1, F2 = seg([0, 0, 0], [1, 0, 0])
    u = O.bs_mex(F1, F2, pts([0, 0, 0]), one(5.0), 0.0)
    assert np.all(np.isnan(u))
    assert np.all(O.bs_matlab(F1, F2, pts([0, 0, 0]), one(5.0), one(0), 0.0) == 0.0)


def test_far_cutoff_co():
    F1, F2 = seg([0, 0, 0], [1, 0, 0])
    P = pts([0.5, 3.0, 0], [0.5, 30.0, 0])
    u_inf = O.bs_matlab(F1, F2, P, one(1.0), one(0), 0.0, co=math.inf)
    u_10 = O.bs_matlab(F1, F2, P, one(1.0), one(0), 0.0, co=10.0)
    assert u_10[0, 2, 0, 0] == u_inf[0, 2, 0, 0] != 0
    assert u_10[1, 2, 0, 0] == 0.0 and u_inf[1, 2, 0, 0] != 0


# --------------------------------------------------------------------------
# invariants on a random wake
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def small_wake():
    return gen.rotor_wake(nb=2, nst=6, ages=9, jitter=0.01)


def test_linearity_in_gamma(small_wake):
    w = small_wake
    F1, F2, g, rc = w.trail_segments()
    g2 = np.asfortranarray(np.cos(np.arange(g.size)).reshape(g.shape, order="F"))
    for f in (lambda gg: O.bs_mex(F1, F2, w.domain, gg, 1e-4),
              lambda gg: O.bs_matlab(F1, F2, w.domain, gg, rc, "visc2")):
        lhs = f(2.5 * g - 0.5 * g2)
        rhs = 2.5 * f(g) - 0.5 * f(g2)
        assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(lhs))


def test_swap_endpoints_flips_sign_exactly(small_wake):
    F1, F2, g, rc = small_wake.shed_segments()
    a = O.bs_mex(F1, F2, small_wake.domain, g, 1e-4)
    b = O.bs_mex(F2, F1, small_wake.domain, g, 1e-4)
    assert np.array_equal(a, -b)


def test_rigid_motion_equivariance(small_wake):
    w = small_wake
    F1, F2, g, rc = w.trail_segments()
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    t = np.array([12.0, -4.0, 7.5])

    def move(X):
        Y = np.einsum("ij,sjtb->sitb", Q, X) + t[None, :, None, None]
        return np.asfortranarray(Y)

    for d in (1e-4, "visc1"):
        u = O.bs_matlab(F1, F2, w.domain, g, rc, d)
        v = O.bs_matlab(move(F1), move(F2), move(w.domain), g, rc, d)
        Ru = np.einsum("ij,sjtb->sitb", Q, u)
        assert np.max(np.abs(v - Ru)) <= 1e-11 * np.max(np.abs(u))


def test_superposition_shed_plus_trail(small_wake):
    w = small_wake
    s = w.shed_segments()
    t = w.trail_segments()
    us = O.bs_mex(s[0], s[1], w.domain, s[2], 1e-4)
    ut = O.bs_mex(t[0], t[1], w.domain, t[2], 1e-4)
    # one call over every segment, in packed order k = m + n*ns + p*ns*nt
    recs = np.concatenate([layout.segments(*s), layout.segments(*t)])
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    g = recs[:, 6]
    uall = O.bs_mex(F1, F2, w.domain, g.reshape(-1, 1, 1, 1), 1e-4)
    assert np.max(np.abs(uall - (us + ut))) <= 1e-13 * np.max(np.abs(uall))


def test_divergence_free_away_from_filaments():
    # div u = 0 for a closed ring, by central differences at a point off the ring
    N, R, G = 24, 1.0, 3.0
    th = 2 * math.pi * np.arange(N + 1) / N
    ring = np.stack([np.zeros(N + 1), R * np.cos(th), R * np.sin(th)], axis=1)
    F1 = np.asfortranarray(ring[:-1].reshape(N, 3, 1, 1))
    F2 = np.asfortranarray(ring[1:].reshape(N, 3, 1, 1))
    g = np.full((N, 1, 1, 1), G)
    x0 = np.array([0.3, 0.4, -0.2])
    e = 1e-4
    P = []
    for i in range(3):
        for s in (+1, -1):
            p = x0.copy()
            p[i] += s * e
            P.append(p)
    u = O.bs_mex(F1, F2, pts(*P), g, 0.0)
    div = sum((u[2 * i, i, 0, 0] - u[2 * i + 1, i, 0, 0]) / (2 * e) for i in range(3))
    scale = np.max(np.abs(u)) / e
    assert abs(div) < 1e-6 * scale


def test_fp64_oracle_against_quad_bruteforce(small_wake):
    w = small_wake
    for segs in (w.shed_segments(), w.trail_segments()):
        F1, F2, g, rc = segs
        for d in (1e-4, "visc1", "visc2", "visc3"):
            ref = O.bs_quad(F1, F2, w.domain, g, rc, d)
            got = O.bs_matlab(F1, F2, w.domain, g, rc, d)
            scale = np.max(np.abs(ref))
            assert np.max(np.abs(got - ref)) <= 1e-13 * scale, d
        got = O.bs_mex(F1, F2, w.domain, g, 1e-4)
        ref = O.bs_quad(F1, F2, w.domain, g, rc, 1e-4)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_mex_and_matlab_cutoff_agree(small_wake):
    F1, F2, g, rc = small_wake.trail_segments()
    a = O.bs_mex(F1, F2, small_wake.domain, g, 1e-4)
    b = O.bs_matlab(F1, F2, small_wake.domain, g, rc, 1e-4)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))


def test_threads_do_not_change_bits(small_wake):
    F1, F2, g, rc = small_wake.trail_segments()
    a = O.bs_mex(F1, F2, small_wake.domain, g, 1e-4, threads=1)
    b = O.bs_mex(F1, F2, small_wake.domain, g, 1e-4, threads=4)
    assert np.array_equal(a, b)
    r = O.bs_mex_range(F1, F2, small_wake.domain, g, 1e-4, 3, 17, threads=2)
    pk = np.stack([a[:, c].reshape(-1, order="F") for c in range(3)], axis=1)
    assert np.array_equal(r, pk[3:17])


def test_filament_length():
    F1, F2 = seg([1, 2, 3], [4, 6, 3])
    assert O.filament_length(F1, F2)[0, 0, 0, 0] == 5.0


# --------------------------------------------------------------------------
# generator: Kelvin bookkeeping of WInDS.m:131-140
# --------------------------------------------------------------------------

def test_generator_kelvin_relations(small_wake):
    w = small_wake
    tr = w.gamma_trail[:, 0]          # [nst, A-1, nb]
    sh = w.gamma_shed[:, 0]           # [ns, A, nb]
    # trailing circulation sums to zero over the span at every age
    assert np.max(np.abs(tr.sum(axis=0))) < 1e-12
    G = np.cumsum(tr[:-1], axis=0)    # bound history recovered by cumsum
    assert np.allclose(sh[:, 0], G[:, 0], atol=1e-13)
    assert np.allclose(sh[:, 1:-1], G[:, 1:] - G[:, :-1], atol=1e-13)
    assert np.allclose(sh[:, -1], -G[:, -1], atol=1e-13)
    # total shed circulation of each station telescopes to zero
    assert np.max(np.abs(sh.sum(axis=1))) < 1e-12


def test_generator_config_sizes():
    c1 = gen.config1()
    assert c1.n_targets == 3888 and c1.n_segments == 7506
    p = gen.config3_params()
    assert len(p) == 64
    assert all(2 <= q["nb"] <= 3 and 12 <= q["nst"] <= 40 and 200 <= q["ages"] <= 2000 for q in p)
    assert gen.rotor_wake(nb=3, nst=31, ages=4).n_segments == 30 * 4 * 3 + 31 * 3 * 3


def test_generator_is_deterministic():
    a = gen.rotor_wake(nb=2, nst=5, ages=6)
    b = gen.rotor_wake(nb=2, nst=5, ages=6)
    assert np.array_equal(a.domain, b.domain)
    assert np.array_equal(a.gamma_shed, b.gamma_shed)
    # SplitMix64 known answers for seed 0 (first three outputs)
    z = gen.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
