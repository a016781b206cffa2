"""Device ground-effect panels against the CPU restatement (oracle/ground.py).

The influence matrix and the induced velocity follow the reference's
arithmetic in the same order (FP64, no contraction), so they must agree with
the oracle to the last bits (<= 1e-15 relative); the circulation comes from
a conjugate-gradient solve instead of inv(A) * V_n and must agree to 1e-11
relative.  Sizes: the reference's example grid (Main_Driver.m: 50 x 50
panels over [-500, 1500] x [-375, 375], cut-off 0.05) and small ragged
cases (one panel, a non-square extent)."""
import numpy as np
import pytest

from conftest import flush_print
import winds
from oracle import ground as G

pytestmark = pytest.mark.gpu


def make(np_side, x0, x1, y0, y1, co):
    from winds.ground import Ground
    return Ground(np_side, x0, x1, y0, y1, co)


@pytest.mark.parametrize("np_side,ext,co", [(1, (0.0, 3.0, 0.0, 3.0), 0.0),
                                            (7, (-500.0, 1500.0, -375.0, 375.0), 0.05),
                                            (12, (-20.0, 100.0, -7.0, 13.0), 0.1)])
def test_matrix_and_solve_match_oracle(need_gpu, np_side, ext, co):
    flush_print(f"[ground] Np={np_side} ext={ext} co={co}")
    g = make(np_side, *ext, co)
    xp, yp = G.nodes(np_side, ext[0], ext[1]), G.nodes(np_side, ext[2], ext[3])
    A0, nv, P0 = G.influence(xp, yp, co)
    P, A = g.geometry()
    P, A = P.cpu().numpy(), A.cpu().numpy()
    assert np.array_equal(P, P0)
    assert np.max(np.abs(A - A0)) <= 1e-15 * np.max(np.abs(A0))
    rng = np.random.default_rng(np_side)
    N = np_side * np_side
    v_ext = np.column_stack([np.full(N, 11.4), rng.normal(0, 1, N), rng.normal(-1, 2, N)])
    gam = g.solve(v_ext).cpu().numpy()
    ref = G.solve(A0, nv, v_ext)
    err = np.max(np.abs(gam - ref)) / np.max(np.abs(ref))
    flush_print(f"[ground] gamma err {err:.2e} in {g.last_iterations} iterations "
                f"(residual {g.last_residual:.1e})")
    assert err <= 1e-11
    assert g.last_residual <= 1e-13


def test_reference_grid_no_penetration(need_gpu):
    """Main_Driver.m's 50 x 50 grid: the device circulation cancels the
    external normal velocity at all 2,500 control points, checked with the
    oracle's induced-velocity routine (independent of the matrix)."""
    g = make(50, -500.0, 1500.0, -375.0, 375.0, 0.05)
    xp, yp = G.nodes(50, -500.0, 1500.0), G.nodes(50, -375.0, 375.0)
    P0 = G.centres(xp, yp)
    rng = np.random.default_rng(2)
    v_ext = np.column_stack([np.full(2500, 11.4), rng.normal(0, 0.5, 2500),
                             -1.0 + 0.5 * np.sin(P0[:, 0] / 200.0)])
    gam = g.solve(v_ext)
    flush_print(f"[ground 50x50] {g.last_iterations} iterations, residual {g.last_residual:.1e}")
    assert g.last_iterations < 2500
    vind = G.induced(P0, gam.cpu().numpy(), xp, yp, 0.05)
    assert np.max(np.abs(vind[:, 2] + v_ext[:, 2])) <= 1e-11


def test_induced_velocity_matches_oracle(need_gpu):
    g = make(20, -500.0, 1500.0, -375.0, 375.0, 0.05)
    xp, yp = G.nodes(20, -500.0, 1500.0), G.nodes(20, -375.0, 375.0)
    rng = np.random.default_rng(9)
    gam = rng.normal(0, 3, 400)
    P = np.column_stack([rng.uniform(-600, 1600, 3000), rng.uniform(-400, 400, 3000),
                         rng.uniform(-10, 200, 3000)])
    P[:10, 2] = 0.0                                   # on the plane
    P[10, :] = [xp[3], yp[4], 0.0]                    # on a panel corner
    u = g.induced(gam, P).cpu().numpy()
    ref = G.induced(P, gam, xp, yp, 0.05)
    assert np.all(np.isfinite(u))
    assert np.max(np.abs(u - ref)) <= 1e-15 * np.max(np.abs(ref))


def test_bad_arguments():
    from winds.ground import Ground
    with pytest.raises(winds.WindsError, match="panels per side"):
        Ground(0, 0.0, 1.0, 0.0, 1.0)
    with pytest.raises(winds.WindsError, match="panels per side"):
        Ground(129, 0.0, 1.0, 0.0, 1.0)
    with pytest.raises(winds.WindsError, match="extent"):
        Ground(4, 1.0, 1.0, 0.0, 1.0)
    g = Ground(2, 0.0, 1.0, 0.0, 1.0)
    with pytest.raises(ValueError, match="one per panel"):
        g.solve(np.zeros((3, 3)))
    with pytest.raises(ValueError, match="one per panel"):
        g.induced(np.zeros(5), np.zeros((2, 3)))
    assert g.induced(np.zeros(4), np.zeros((0, 3))).shape == (0, 3)
