"""Parity of the HIP engine with the CPU oracle, through the C ABI.

FP64 with MEX semantics must be bitwise equal to the restated CPU path of
BiotSavartMex (BiotSavart_nogpu, BiotSavartMex.cu:268-282): same operation
order, same sequential sum, correctly rounded sqrt and division on both
sides.  FP64 with BiotSavart.m semantics must agree within 1e-12 relative.
FP32 compensated must agree within the bounds written in each test.
"""
import math

import numpy as np
import pytest

from conftest import flush_print
from errstats import stats
from oracle import bs as O
import winds
from winds import gen, layout

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(need_gpu):
    return gen.config1()


def calls(w):
    return {"shed": w.shed_segments(), "trail": w.trail_segments()}


# --------------------------------------------------------------------------
# FP64
# --------------------------------------------------------------------------

@pytest.mark.parametrize("type", ["fuln", "fuls", "full"])
def test_fp64_mex_bitexact_config1(c1, type):
    for name, (F1, F2, g, rc) in calls(c1).items():
        flush_print(f"[fp64 mex {type}] {name}")
        u, _ = winds.BiotSavartMex(F1, F2, c1.domain, g, rc, 1e-4, "none", 1000.0, type, "true",
                                   (256, 1, 0))
        ref = O.bs_mex(F1, F2, c1.domain, g, 1e-4)
        assert np.array_equal(u, ref), (name, np.max(np.abs(u - ref)))


@pytest.mark.parametrize("d", [1e-4, 0.0, "visc1", "visc2", "visc3"])
def test_fp64_matlab_config1(c1, d):
    for name, (F1, F2, g, rc) in calls(c1).items():
        flush_print(f"[fp64 matlab {d}] {name}")
        u, L = winds.BiotSavart(F1, F2, c1.domain, g, rc, d, 1000.0, "full")
        ref = O.bs_matlab(F1, F2, c1.domain, g, rc, d, co=1000.0)
        s = stats(u, ref)
        flush_print(f"   {s}  exact={np.array_equal(u, ref)}")
        assert s["max_rel"] <= 1e-12 and s["normwise"] <= 1e-12
        # 'full' returns the squared filament length (BiotSavart.m:71)
        assert np.allclose(L, O.filament_length(F1, F2) ** 2, rtol=1e-15, atol=0)


def test_fp64_matlab_cutoff_bitexact(c1):
    F1, F2, g, rc = c1.trail_segments()
    u, _ = winds.BiotSavart(F1, F2, c1.domain, g, rc, 1e-4, 1000.0, "full")
    ref = O.bs_matlab(F1, F2, c1.domain, g, rc, 1e-4, co=1000.0)
    assert np.array_equal(u, ref)


@pytest.mark.parametrize("m,p,q,nt", [(1, 1, 1, 1), (5, 1, 1, 3), (5, 3, 1, 64), (7, 9, 2, 65),
                                      (3, 40, 3, 130), (18, 17, 1, 200)])
def test_fp64_matlab_nested_sum_shapes_bitexact(need_gpu, m, p, q, nt):
    """The nested sum sum(sum(sum(., 1), 3), 4) at shapes with fewer
    columns than a workgroup's waves, a single column, partial rounds and
    partial target groups: bitwise equal to the oracle's loop."""
    rng = np.random.default_rng(1000 * m + 10 * p + q)
    F1 = np.asfortranarray(rng.normal(size=(m, 3, p, q)))
    F2 = np.asfortranarray(F1 + 0.3 * rng.normal(size=(m, 3, p, q)))
    g = np.asfortranarray(rng.normal(size=(m, 1, p, q)))
    rc = np.asfortranarray(0.01 + 0.01 * rng.random(size=(m, 1, p, q)))
    P = np.asfortranarray(rng.normal(size=(nt, 3, 1, 1)) * 2.0)
    for d in (1e-4, "visc2"):
        u, _ = winds.BiotSavart(F1, F2, P, g, rc, d, 1000.0, "full")
        ref = O.bs_matlab(F1, F2, P, g, rc, d, co=1000.0)
        if d == 1e-4:
            assert np.array_equal(u, ref), np.max(np.abs(u - ref))
        else:
            assert stats(u, ref)["max_rel"] <= 1e-12


# --------------------------------------------------------------------------
# FP32 compensated
# --------------------------------------------------------------------------

@pytest.mark.parametrize("cmod,d", [("none", 1e-4), ("cutm", 1e-4), ("visc", 1), ("visc", 2),
                                    ("visc", 3)])
def test_fp32c_config1(c1, cmod, d):
    for name, (F1, F2, g, rc) in calls(c1).items():
        flush_print(f"[fp32c {cmod}{d}] {name}")
        u, _ = winds.BiotSavartMex(F1, F2, c1.domain, g, rc, d, cmod, 1000.0, "fulc", "true",
                                   (256, 1, 0))
        if cmod == "none":
            ref = O.bs_mex(F1, F2, c1.domain, g, d)
        else:
            ref = O.bs_matlab(F1, F2, c1.domain, g, rc, d if cmod == "cutm" else f"visc{d}",
                              co=1000.0)
        s = stats(u, ref)
        flush_print(f"   {s}")
        # FP32 pair arithmetic, compensated sums, FP64 near pairs.  A shed-only
        # or trail-only call leaves many targets with |u| near 1e-3 of the
        # largest, where the relative error is the absolute error magnified;
        # the 1e-6 per-target bound is asserted on whole evaluations below.
        assert s["max_rel"] <= 5e-6 and s["rms_rel"] <= 3e-7 and s["normwise"] <= 1e-7


@pytest.mark.parametrize("chained", [False, True])
@pytest.mark.parametrize("core", ["vatistas2", "vatistas1", "vatistas3", "cutoff"])
def test_fp32c_config1_whole_wake(c1, core, chained):
    import torch
    from winds.device import Evaluator
    recs = layout.wake_segments(c1, chained=chained)
    tg = torch.from_numpy(layout.wake_targets(c1)).cuda().contiguous()
    sg = torch.from_numpy(recs).cuda().contiguous()
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    if core == "cutoff":
        o = winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4,
                       precision=winds.PREC_FP32C)
        ref = O.bs_mex(F1, F2, c1.domain, recs[:, 6], 1e-4)
    else:
        n = int(core[-1])
        o = winds.opts(core=winds.CORE_VATISTAS, vat_n=n, semantics=winds.SEM_MATLAB,
                       precision=winds.PREC_FP32C)
        ref = O.bs_matlab(F1, F2, c1.domain, recs[:, 6], recs[:, 7], f"visc{n}")
    ev = Evaluator(0)
    u = layout.uind_to_cm(ev.eval(tg, sg, o).cpu().numpy(), c1.domain.shape)
    s = stats(u, ref)
    flush_print(f"[fp32c whole wake {core} chained={chained}] {s}")
    assert s["max_rel"] <= 1e-6 and s["rms_rel"] <= 1e-7 and s["normwise"] <= 1e-7
    # the same call again gives the same bits (fixed summation orders,
    # near-pair sums added in list order)
    u2 = layout.uind_to_cm(ev.eval(tg, sg, o).cpu().numpy(), c1.domain.shape)
    assert np.array_equal(u, u2)
    # segment ranges that start inside a filament (the kernel then forms r1
    # from the range's first start point instead of carrying it)
    for split in (2, 5, 7):
        o.split = split
        us = layout.uind_to_cm(ev.eval(tg, sg, o).cpu().numpy(), c1.domain.shape)
        ss = stats(us, ref)
        flush_print(f"   split {split}: {ss}")
        assert ss["max_rel"] <= 1e-6 and ss["rms_rel"] <= 1e-7 and ss["normwise"] <= 1e-7


@pytest.mark.parametrize("co", [25.0, 60.0])
def test_fp32c_far_cutoff_whole_wake(c1, co):
    """The far cut-off co (BiotSavart.m:111) drops a pair's whole term when
    r1 > co and r2 > co.  With co inside the wake many pairs sit at the step;
    pairs within FP32 rounding of it take the FP64 branch, so the FP32 sum
    keeps or drops the same terms as the FP64 reference."""
    import torch
    from winds.device import Evaluator
    recs = layout.wake_segments(c1, chained=True)
    tg = torch.from_numpy(layout.wake_targets(c1)).cuda().contiguous()
    sg = torch.from_numpy(recs).cuda().contiguous()
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    ref = O.bs_matlab(F1, F2, c1.domain, recs[:, 6], recs[:, 7], "visc1", co=co)
    o = winds.opts(core=winds.CORE_VATISTAS, vat_n=1, semantics=winds.SEM_MATLAB, co=co,
                   precision=winds.PREC_FP32C)
    u = layout.uind_to_cm(Evaluator(0).eval(tg, sg, o).cpu().numpy(), c1.domain.shape)
    s = stats(u, ref)
    flush_print(f"[fp32c far cut-off co={co}] {s}")
    assert s["max_rel"] <= 1e-6 and s["normwise"] <= 1e-7


# --------------------------------------------------------------------------
# edge cases of the boundary
# --------------------------------------------------------------------------

def seg(a, b):
    return (np.asarray(a, float).reshape(1, 3, 1, 1, order="F"),
            np.asarray(b, float).reshape(1, 3, 1, 1, order="F"))


@pytest.mark.parametrize("type", ["fuln", "fulc"])
def test_empty_inputs(need_gpu, type):
    P = np.ones((4, 3, 2, 1), order="F")
    F = np.zeros((0, 3, 5, 1), order="F")
    u, L = winds.BiotSavartMex(F, F, P, np.zeros(0), np.zeros(0), 1e-4, "none", 1.0, type)
    assert u.shape == (4, 3, 2, 1) and np.all(u == 0)
    F1, F2 = seg([0, 0, 0], [1, 0, 0])
    u, L = winds.BiotSavartMex(F1, F2, np.zeros((0, 3, 1, 1)), [1.0], [0.0], 1e-4, "none", 1.0, type)
    assert u.shape == (0, 3, 1, 1)


def test_ragged_target_shapes(c1):
    F1, F2, g, rc = c1.shed_segments()
    P4 = c1.domain[:, :, 5:6, 1:2]
    P2 = np.asfortranarray(P4[:, :, 0, 0])          # 2-D P is padded to [n,3,1,1]
    u2, _ = winds.BiotSavartMex(F1, F2, P2, g, rc, 1e-4, "none", 1.0, "fuln")
    u4, _ = winds.BiotSavartMex(F1, F2, P4, g, rc, 1e-4, "none", 1.0, "fuln")
    assert u2.shape == (P2.shape[0], 3, 1, 1)
    assert np.array_equal(u2[:, :, 0, 0], u4[:, :, 0, 0])
    one = np.asfortranarray(c1.domain[3:4, :, 7:8, 2:3])
    u1, _ = winds.BiotSavartMex(F1, F2, one, g, rc, 1e-4, "none", 1.0, "fuln")
    assert np.array_equal(u1, O.bs_mex(F1, F2, one, g, 1e-4))


@pytest.mark.parametrize("type", ["fuln", "fulc"])
def test_singular_targets(need_gpu, type):
    F1, F2 = seg([0, 0, 0], [1, 0, 0])
    P = np.asfortranarray(np.array([[0, 0, 0], [1, 0, 0], [2.5, 0, 0], [0.5, 0, 0],
                                    [0.5, 0.3, 0]], float).reshape(5, 3, 1, 1))
    g, rc = [5.0], [1e-3]
    # MEX semantics, d > 0: exact zeros at endpoints and on the line
    u, _ = winds.BiotSavartMex(F1, F2, P, g, rc, 1e-4, "none", 1.0, type)
    ref = O.bs_mex(F1, F2, P, g, 1e-4)
    assert np.all(u[:4] == 0) and np.all(ref[:4] == 0)
    assert abs(u[4, 2, 0, 0] - ref[4, 2, 0, 0]) <= 1e-12 * abs(ref[4, 2, 0, 0])
    # MEX semantics, d = 0: the reference's NaN at an endpoint is kept
    u, _ = winds.BiotSavartMex(F1, F2, P[:1], g, rc, 0.0, "none", 1.0, type)
    assert np.all(np.isnan(u))
    # BiotSavart.m semantics: zeroed (BiotSavart.m:111)
    for cmod, d in (("cutm", 0.0), ("visc", 1), ("visc", 2)):
        u, _ = winds.BiotSavartMex(F1, F2, P, g, rc, d, cmod, 1000.0, type)
        assert np.all(u[:4] == 0), (cmod, d)
        assert np.isfinite(u).all()


@pytest.mark.parametrize("type", ["fuln", "fulc"])
def test_zero_length_segment_and_co(need_gpu, type):
    F1 = np.asfortranarray(np.array([[0, 0, 0], [0, 0, 0]], float).reshape(2, 3, 1, 1))
    F2 = np.asfortranarray(np.array([[0, 0, 0], [1, 0, 0]], float).reshape(2, 3, 1, 1))
    P = np.asfortranarray(np.array([[0.5, 2.0, 0], [0.5, 50.0, 0], [0, 0, 0]], float)
                          .reshape(3, 3, 1, 1))
    g, rc = [3.0, 2.0], [1e-3, 1e-3]
    for cmod, d in (("cutm", 1e-4), ("visc", 2)):
        u, _ = winds.BiotSavartMex(F1, F2, P, g, rc, d, cmod, 10.0, type)
        dd = d if cmod == "cutm" else f"visc{d}"
        ref = O.bs_matlab(F1, F2, P, g, rc, dd, co=10.0)
        assert np.isfinite(u).all()
        assert u[1, 2, 0, 0] == 0.0 and ref[1, 2, 0, 0] == 0.0   # beyond co
        assert np.all(u[2] == 0)
        tol = 0 if type == "fuln" else 1e-6
        assert abs(u[0, 2, 0, 0] - ref[0, 2, 0, 0]) <= tol * abs(ref[0, 2, 0, 0])


@pytest.mark.parametrize("type", ["fuln", "fulc"])
def test_nan_circulation(need_gpu, type):
    """One segment of a chained filament carries a NaN circulation.  MEX
    semantics propagate it to every target, as biot_p2p's sum does
    (BiotSavartMex.cu:68-92); BiotSavart.m semantics zero that segment's
    NaN terms (BiotSavart.m:111) and the rest of the sum stands.  The FP32
    path sends a NaN record to the FP64 arithmetic, so it behaves the same."""
    rng = np.random.default_rng(5)
    ns = 40
    X = np.cumsum(rng.normal(size=(ns + 1, 3)), axis=0)
    F1 = np.asfortranarray(X[:-1].reshape(ns, 3, 1, 1))
    F2 = np.asfortranarray(X[1:].reshape(ns, 3, 1, 1))
    g = rng.uniform(1.0, 5.0, ns)
    g[17] = np.nan
    rc = np.full(ns, 1e-3)
    P = np.asfortranarray((X[:-1] + 0.5 * (X[1:] - X[:-1])
                           + rng.normal(size=(ns, 3)) * 0.5).reshape(ns, 3, 1, 1))
    u, _ = winds.BiotSavartMex(F1, F2, P, g, rc, 1e-4, "none", 1.0, type)
    assert np.all(np.isnan(u))
    for cmod, d, dd in (("cutm", 1e-4, 1e-4), ("visc", 2, "visc2")):
        u, _ = winds.BiotSavartMex(F1, F2, P, g, rc, d, cmod, 1000.0, type)
        ref = O.bs_matlab(F1, F2, P, g.reshape(ns, 1, 1, 1), rc.reshape(ns, 1, 1, 1), dd, co=1000.0)
        assert np.isfinite(u).all(), cmod
        s = stats(u, ref)
        if type == "fuln":
            assert s["normwise"] <= 1e-12, (cmod, s)
        else:
            assert s["max_rel"] <= 1e-6, (cmod, s)


@pytest.mark.parametrize("threads", [64, 128, 512, 1024, 0])
def test_block_sizes_do_not_change_fp64_bits(c1, threads):
    F1, F2, g, rc = c1.trail_segments()
    u, _ = winds.BiotSavartMex(F1, F2, c1.domain, g, rc, 1e-4, "none", 1.0, "fuln", "true",
                               (threads, 1, 0))
    assert np.array_equal(u, O.bs_mex(F1, F2, c1.domain, g, 1e-4))


def test_length_mode(c1):
    F1, F2, g, rc = c1.trail_segments()
    for t in ("leng", "length"):
        u, L = winds.BiotSavartMex(F1, F2, c1.domain, g, rc, 1e-4, "none", 1.0, t)
        assert np.all(u == 0)
        assert np.array_equal(L, O.filament_length(F1, F2))
    u, L = winds.BiotSavart(F1, F2, c1.domain, g, rc, 1e-4, 1000.0, "length")
    assert np.array_equal(L, O.filament_length(F1, F2))


# --------------------------------------------------------------------------
# device-resident API
# --------------------------------------------------------------------------

def test_device_api_matches_host_api(c1):
    import torch
    from winds.device import Evaluator
    ev = Evaluator(0)
    F1, F2, g, rc = c1.trail_segments()
    segs = torch.from_numpy(layout.segments(F1, F2, g, rc)).cuda()
    tg = torch.from_numpy(layout.targets(c1.domain)).cuda().contiguous()
    u = ev.eval(tg, segs, winds.opts(d=1e-4)).cpu().numpy()
    ref = O.bs_mex(F1, F2, c1.domain, g, 1e-4)
    assert np.array_equal(layout.uind_to_cm(u, c1.domain.shape), ref)
    # FP32 compensated: the split of the segment range changes only rounding
    outs = []
    for split in (1, 3, 8):
        o = winds.opts(precision=winds.PREC_FP32C, d=1e-4, split=split)
        outs.append(ev.eval(tg, segs, o).cpu().numpy())
        s = stats(layout.uind_to_cm(outs[-1], c1.domain.shape), ref)
        flush_print(f"[split {split}] {s}")
        assert s["normwise"] <= 1e-6
    for x in outs[1:]:
        assert np.max(np.abs(x - outs[0])) <= 1e-7 * np.max(np.abs(outs[0]))


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_fp32c_targets_inside_segment_sphere_take_fp64(need_gpu, core):
    """A target inside the sphere on a segment as diameter (r1.r2 < 0) is a
    near pair of the FP32 kernel: the near radius is at least 0.51|L|, also
    when a smaller fraction is asked for.  Its term is then the FP64 term,
    bit for bit, and the FP32 arithmetic never meets r1 r2 + r1.r2 ~ 0."""
    import torch
    from winds.device import Evaluator
    ev = Evaluator(0)
    if core == "cutoff":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    rng = np.random.default_rng(11)
    n = 4096
    for trial in range(6):
        A = rng.uniform(-800.0, 800.0, 3)
        L = rng.normal(size=3) * rng.uniform(0.2, 20.0)
        F1 = A.reshape(1, 3, 1, 1)
        F2 = (A + L).reshape(1, 3, 1, 1)
        g = np.array([rng.uniform(-60.0, 60.0)])
        rc = np.array([rng.uniform(1e-4, 1e-2)])  # squared core radius
        sg = torch.from_numpy(layout.segments(F1, F2, g, rc)).cuda()
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rho = 0.999 * rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0)
        P = (A + 0.5 * L) + 0.5 * np.linalg.norm(L) * rho[:, None] * d
        assert np.all(np.einsum("ij,ij->i", P - A, P - A - L) < 0.0)
        tg = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
        flush_print(f"[inside sphere {core}] trial {trial}, |L| = {np.linalg.norm(L):.3f}")
        u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
        for frac in (0.51, 0.3):
            o = winds.opts(precision=winds.PREC_FP32C, near_factor=1e-3, near_len_frac=frac, **kw)
            u32 = ev.eval(tg, sg, o).cpu().numpy()
            assert np.array_equal(u32, u64), (trial, frac, np.max(np.abs(u32 - u64)))


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_fp32c_chained_filament_near_pairs(need_gpu, core):
    """A filament of chained segments whose lengths (and so near radii)
    alternate by a factor of up to 60, with targets inside each segment's
    sphere a hair off its line (where r1 r2 + r1.r2 nearly cancels and an
    FP32 term would be wrong in its leading digits) and just beyond each end
    point.  The near test of a chained record takes its start-point half
    from the record before, with that record's end-point threshold raised to
    cover both; every such pair must still take the FP64 term, so each
    target's velocity (dominated by that term) matches FP64 to 1e-6."""
    import torch
    from winds.device import Evaluator
    ev = Evaluator(0)
    if core == "cutoff":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    rng = np.random.default_rng(23)
    ns = 41
    lens = np.where(np.arange(ns) % 2 == 0, rng.uniform(0.2, 0.5, ns), rng.uniform(6.0, 12.0, ns))
    dirs = rng.normal(size=(ns, 3)) * 0.3 + np.array([1.0, 0.2, 0.0])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    X = np.vstack([rng.uniform(-500.0, 500.0, 3), np.zeros((ns, 3))])
    for j in range(ns):
        X[j + 1] = X[j] + lens[j] * dirs[j]
    F1 = X[:-1].reshape(ns, 3, 1, 1)
    F2 = X[1:].reshape(ns, 3, 1, 1)
    g = rng.uniform(20.0, 60.0, ns)
    rc = np.full(ns, 1e-8)                      # squared core radius: R = 0.51|L|
    recs = layout.segments(F1, F2, g, rc)
    assert np.all(recs[1:, 0:3] == recs[:-1, 3:6])   # one chained filament
    P = []
    for j in range(ns):
        A, B = X[j], X[j + 1]
        L = B - A
        e1 = np.cross(L, [0.3, -0.7, 0.5])
        e1 /= np.linalg.norm(e1)
        for t in np.linspace(0.1, 0.9, 9):      # inside the sphere, off the line
            for h in (1e-4, 1e-3, 1e-2):
                P.append(A + t * L + h * np.linalg.norm(L) * e1)
        for sgn, E in ((-1.0, A), (1.0, B)):    # just beyond each end point
            P.append(E + sgn * 0.2 * L + 0.05 * np.linalg.norm(L) * e1)
    P = np.array(P)
    sg = torch.from_numpy(recs).cuda()
    tg = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    flush_print(f"[chained near pairs {core}] {P.shape[0]} targets x {ns} segments")
    u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
    u32 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, near_factor=1e-3,
                                     **kw)).cpu().numpy()
    n64 = np.linalg.norm(u64, axis=0)
    rel = np.linalg.norm(u32 - u64, axis=0) / n64
    assert np.all(n64 > 0)
    assert rel.max() <= 1e-6, (rel.max(), int(rel.argmax()))


# --------------------------------------------------------------------------
# full single-wake size: 133,920 targets x 263,427 segments
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c2_device(need_gpu):
    import torch
    w = gen.config2()
    recs = layout.wake_segments(w, chained=True)      # the order bench.py times
    tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
    sg = torch.from_numpy(recs).cuda().contiguous()
    return w, recs, tg, sg


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_config2_fp32c_within_1e6_of_fp64(c2_device, core):
    from winds.device import Evaluator
    w, recs, tg, sg = c2_device
    ev = Evaluator(0)
    if core == "cutoff":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    flush_print(f"[config2 {core}] fp64")
    u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
    flush_print(f"[config2 {core}] fp32c")
    u32 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, **kw)).cpu().numpy()
    s = stats(u32, u64)
    flush_print(f"   {s}")
    assert s["max_rel"] <= 1e-6 and s["rms_rel"] <= 1e-7 and s["normwise"] <= 1e-6
    # the FP64 engine itself against the CPU oracle on a sample of targets
    idx = np.random.default_rng(5).choice(w.n_targets, size=24, replace=False)
    P = np.asfortranarray(layout.wake_targets(w)[:, idx].T.reshape(-1, 3, 1, 1))
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    if core == "cutoff":
        ref = O.bs_mex(F1, F2, P, recs[:, 6], 1e-4)
    else:
        ref = O.bs_matlab(F1, F2, P, recs[:, 6], recs[:, 7], "visc2")
    assert np.array_equal(u64[:, idx].T, ref[:, :, 0, 0])


# --------------------------------------------------------------------------
# FP32 cancellation guard, general Vatistas order at large n, co < 0
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def wake60(need_gpu):
    """Wake 60 of the 64-wake uneven batch at full size (3 blades x 20
    stations x 1,449 ages: 86,940 targets x 169,473 segments).  One target
    of it (|u| = 8e-5 of the wake's largest) has FP32 terms whose root sum
    of squares is ~400 |u|: no FP32 arithmetic reaches 1e-6 there."""
    import torch
    w = gen.rotor_wake(**gen.config3_params()[60])
    tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
    sg = torch.from_numpy(layout.wake_segments(w, chained=True)).cuda().contiguous()
    return w, tg, sg


@pytest.mark.parametrize("core", ["vatistas2", "cutoff"])
def test_guard_holds_1e6_per_target_on_cancelling_targets(wake60, core):
    from winds.device import Evaluator
    w, tg, sg = wake60
    ev = Evaluator(0)
    if core == "cutoff":
        kw = dict(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, d=1e-4)
    else:
        kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    flush_print(f"[guard {core}] fp64")
    u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
    flush_print(f"[guard {core}] fp32c, guard on")
    u32 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, **kw)).cpu().numpy()
    n_fix = ev.guard_count()
    s = stats(u32, u64)
    flush_print(f"   guard on: {s}, {n_fix} targets re-evaluated in FP64")
    assert s["max_rel"] <= 1e-6 and s["rms_rel"] <= 1e-7 and s["normwise"] <= 1e-7
    assert 0 < n_fix <= 2e-3 * w.n_targets
    # the same call again: the same bits (the list order may differ, each
    # listed target's FP64 sum does not)
    again = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, **kw)).cpu().numpy()
    assert np.array_equal(u32, again)
    flush_print(f"[guard {core}] fp32c, guard off")
    off = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, guard_kappa=-1.0,
                                     **kw)).cpu().numpy()
    assert ev.guard_count() == 0
    so = stats(off, u64)
    flush_print(f"   guard off: {so}")
    # the guard changes only the listed targets, and only towards FP64
    changed = np.any(off != u32, axis=0)
    assert changed.sum() <= n_fix
    if core == "vatistas2":
        assert so["max_rel"] > 1e-6     # what the guard is for


def test_guard_uses_every_fix_path(need_gpu):
    """kappa = 1e-6 lists every target of a small wake: more than the split
    re-evaluation's capacity, so the later list entries take the one-
    workgroup path.  Both paths add the same chunks in the same order: the
    result is the FP64 chunked sum for every listed target whichever path
    it took, and repeated calls agree bit for bit."""
    import torch
    from winds.device import Evaluator
    w = gen.rotor_wake(nb=3, nst=18, ages=120)
    tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
    sg = torch.from_numpy(layout.wake_segments(w, chained=True)).cuda().contiguous()
    ev = Evaluator(0)
    kw = dict(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB)
    u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
    o = winds.opts(precision=winds.PREC_FP32C, guard_kappa=1e-6, **kw)
    a = ev.eval(tg, sg, o).cpu().numpy()
    n_fix = ev.guard_count()
    flush_print(f"[guard all paths] {n_fix} of {w.n_targets} listed")
    assert n_fix == w.n_targets > 2048
    b = ev.eval(tg, sg, o).cpu().numpy()
    assert np.array_equal(a, b)
    s = stats(a, u64)
    flush_print(f"   {s}")
    assert s["max_rel"] <= 1e-11


@pytest.mark.parametrize("n", [8, 10, 64])
def test_general_vatistas_order_large_n(c1, n):
    """Vatistas order n in the general (run-time n) arithmetic.  h^2n leaves
    the FP32 range at n = 10 beyond h ~ 84 m and the FP64 range at n = 64
    beyond h ~ 259 m (there the reference's term is h^2 * inf^(-1/n) = 0).
    Targets are the wake's nodes plus nodes moved 150-900 m off the rotor
    plane, against the FP64 engine (pow, as the reference) and the oracle."""
    import torch
    from winds.device import Evaluator
    recs = layout.wake_segments(c1, chained=True)
    T = layout.wake_targets(c1)
    rng = np.random.default_rng(n)
    far = T[:, ::7] + rng.uniform(150.0, 900.0, size=(1, T[:, ::7].shape[1])) * \
        np.array([[0.0], [0.6], [0.8]])
    T = np.ascontiguousarray(np.concatenate([T, far], axis=1))
    tg = torch.from_numpy(T).cuda()
    sg = torch.from_numpy(recs).cuda().contiguous()
    ev = Evaluator(0)
    kw = dict(core=winds.CORE_VATISTAS, vat_n=n, semantics=winds.SEM_MATLAB)
    u64 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP64, **kw)).cpu().numpy()
    u32 = ev.eval(tg, sg, winds.opts(precision=winds.PREC_FP32C, **kw)).cpu().numpy()
    assert np.isfinite(u32).all()
    s = stats(u32, u64)
    flush_print(f"[vatistas n={n}] {s}, guard {ev.guard_count()}")
    assert s["max_rel"] <= 1e-6 and s["normwise"] <= 1e-7
    idx = np.random.default_rng(3).choice(T.shape[1], size=32, replace=False)
    P = np.asfortranarray(T[:, idx].T.reshape(-1, 3, 1, 1))
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    ref = O.bs_matlab(F1, F2, P, recs[:, 6], recs[:, 7], f"visc{n}")
    assert stats(u64[:, idx], ref[:, :, 0, 0].T)["max_rel"] <= 1e-12


@pytest.mark.parametrize("cmod,d", [("cutm", 1e-4), ("visc", 2)])
def test_negative_co_zeroes_every_term(c1, cmod, d):
    """r1 > co and r2 > co holds for every pair when co < 0, so BiotSavart.m
    zeroes every term (BiotSavart.m:111); the FP32 path does the same."""
    F1, F2, g, rc = c1.trail_segments()
    for typ in ("fuln", "fulc"):
        u, _ = winds.BiotSavartMex(F1, F2, c1.domain, g, rc, d, cmod, -2.0, typ)
        assert np.all(u == 0), (cmod, typ)
