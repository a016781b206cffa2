"""Far-wake surrogate: training data from the engine, training, and the march
configuration (include/winds_surrogate.h, include/winds_march.h).

The network maps a query point of the near wake or the lifting line plus a
descriptor of the far wake (every segment with an endpoint of age >= split)
to the far wake's induced velocity there.  Training data come from the
engine's own free-wake marches at seeded operating points over TSR 4-10 and
yaw 0-30 deg (marched_training_set), or from prescribed rigid helices
(build_training_set, which do not teach the rolled-up, expanding wake of a
march); the labels are direct sums of the engine's FP32-compensated kernel
over the far segments, and the inputs are built by the same device code the
march uses.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, gen, layout

NBINS, PER_BIN, NGLOBAL = 8, 5, 5
NDESC = NBINS * PER_BIN + NGLOBAL
NIN = 4 + NDESC
NOUT = 3


class SurrGeom(C.Structure):
    _fields_ = [("nb", C.c_int32), ("nst", C.c_int32), ("ages", C.c_int32),
                ("split", C.c_int32), ("max_ages", C.c_int32), ("reserved", C.c_int32),
                ("R", C.c_double), ("U", C.c_double), ("psi", C.c_double),
                ("hub", C.c_double * 3)]


def _bind(L):
    if getattr(L, "_winds_surr_bound", False):
        return L
    P = C.c_void_p
    G = C.POINTER(SurrGeom)
    L.winds_surr_features_dev.argtypes = [G, P, P, P, C.c_int64, P, P, P, P]
    L.winds_surr_labels_dev.argtypes = [G, P, C.c_int64, P, P, P]
    L.winds_surr_apply_dev.argtypes = [G, P, C.c_int64, P, P, P]
    for f in ("features", "labels", "apply"):
        getattr(L, f"winds_surr_{f}_dev").restype = C.c_int
    L._winds_surr_bound = True
    return L


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------- layouts

def march_layout(w: gen.RotorWake):
    """(X [3, N], gt [nst*(A-1)*nb]) in the march layout j = m + a*nst + b*nst*A."""
    X = np.stack([w.domain[:, c].reshape(-1, order="F") for c in range(3)])
    gt = np.asarray(w.gamma_trail).reshape(-1, order="F")
    return np.ascontiguousarray(X), np.ascontiguousarray(gt)


def far_segments(w: gen.RotorWake, split: int) -> np.ndarray:
    """[n, 8] records of every segment with an endpoint of age >= split."""
    F1, F2, g, rc = w.shed_segments()
    sh = layout.segments(F1[:, :, split:], F2[:, :, split:], g[:, :, split:], rc[:, :, split:])
    F1, F2, g, rc = w.trail_segments()
    k = split - 1
    tr = layout.segments(F1[:, :, k:], F2[:, :, k:], g[:, :, k:], rc[:, :, k:])
    return np.concatenate([sh, tr])


def query_points(w: gen.RotorWake, target_ages: int) -> np.ndarray:
    """Near-wake nodes of age < target_ages and the span midpoints of the
    age-0 nodes (lifting-line points), SoA [3, n]."""
    near = layout.targets(w.domain[:, :, :target_ages])
    d0 = w.domain[:, :, 0, :]
    mid = 0.5 * (d0[:-1] + d0[1:])                               # [ns, 3, nb]
    ll = np.stack([mid[:, c, :].reshape(-1, order="F") for c in range(3)])
    return np.ascontiguousarray(np.concatenate([near, ll], axis=1))


def geom(nb, nst, ages, split, max_ages, R, U, psi, hub=(0.0, 0.0, 0.0)) -> SurrGeom:
    g = SurrGeom()
    g.nb, g.nst, g.ages, g.split, g.max_ages = nb, nst, ages, split, max_ages
    g.R, g.U, g.psi = R, U, psi
    for i in range(3):
        g.hub[i] = float(hub[i])
    return g


def features(g: SurrGeom, X, gt, P, in_norm=None):
    """Raw (in_norm None) or normalised input rows [n, NIN] on the device."""
    import torch
    L = _bind(_lib.lib())
    n = P.shape[1]
    feat = torch.empty((n, NIN), dtype=torch.float32, device=P.device)
    desc = torch.empty(64, dtype=torch.float32, device=P.device)
    _lib.check(L.winds_surr_features_dev(C.byref(g), _ptr(X), _ptr(gt), _ptr(P), n,
                                         _ptr(in_norm), _ptr(desc), _ptr(feat), _stream()))
    return feat


def labels(g: SurrGeom, u3, out_norm=None):
    import torch
    L = _bind(_lib.lib())
    n = u3.shape[1]
    T = torch.empty((n, NOUT), dtype=torch.float32, device=u3.device)
    _lib.check(L.winds_surr_labels_dev(C.byref(g), _ptr(u3), n, _ptr(out_norm), _ptr(T),
                                       _stream()))
    return T


def apply(g: SurrGeom, Y, u3, out_norm=None):
    L = _bind(_lib.lib())
    _lib.check(L.winds_surr_apply_dev(C.byref(g), _ptr(Y), Y.shape[0], _ptr(out_norm), _ptr(u3),
                                      _stream()))
    return u3


# ---------------------------------------------------------------- data

@dataclass
class TrainingSet:
    F: "object"          # raw inputs  [n, NIN] float32 (device)
    Y: "object"          # raw targets [n, 3]   float32 (device)
    wakes: int
    interactions: int
    seconds: float


def wake_sample(rng, split, max_ages, nb=3, R=63.0, hub=1.5, U=11.4, dpsi_deg=5.0,
                nst_range=(18, 31), tsr_range=(4.0, 10.0), yaw_range=(0.0, 30.0)):
    """One seeded synthetic wake with a far part."""
    nst = int(rng.integers(nst_range[0], nst_range[1] + 1))
    ages = int(rng.integers(split + 60, max_ages + 1))
    tsr = float(rng.uniform(*tsr_range))
    yaw = float(rng.uniform(*yaw_range))
    a = float(rng.uniform(0.1, 0.4))
    peak = float(rng.uniform(30.0, 90.0))
    psi0 = float(rng.uniform(0.0, 2.0 * math.pi))
    seeds = rng.integers(1, 2**62, size=2)
    w = gen.rotor_wake(nb=nb, nst=nst, ages=ages, R=R, hub=hub, tsr=tsr, U=U, a=a,
                       dpsi_deg=dpsi_deg, jitter=0.002, gamma_peak=peak, yaw_deg=yaw,
                       seed_pos=int(seeds[0]), seed_gamma=int(seeds[1]), psi0=psi0)
    return w, psi0


def build_training_set(n_samples: int, opts, seed: int = 7, split: int = 144,
                       target_ages: int = 72, max_ages: int = 1440, nb: int = 3,
                       R: float = 63.0, U: float = 11.4, hub: float = 1.5,
                       dpsi_deg: float = 5.0, nst_range=(18, 31), log=None) -> TrainingSet:
    """Direct-sum samples of the far wake's velocity at near-wake nodes and
    lifting-line points of seeded wakes (TSR U[4,10], yaw U[0,30] deg, axial
    induction U[0.1,0.4], peak circulation U[30,90] m^2/s, 18-31 stations,
    split+60 .. max_ages ages)."""
    import torch
    from .device import Evaluator
    ev = Evaluator(0)
    rng = np.random.default_rng(seed)
    Fs, Ys = [], []
    n = 0
    inter = 0
    nw = 0
    t0 = time.perf_counter()
    while n < n_samples:
        w, psi0 = wake_sample(rng, split, max_ages, nb=nb, R=R, hub=hub, U=U,
                              dpsi_deg=dpsi_deg, nst_range=nst_range)
        F, Y, k = wake_samples(w, psi0, opts, split, target_ages, max_ages, R, U, ev)
        Fs.append(F)
        Ys.append(Y)
        n += F.shape[0]
        inter += k
        nw += 1
        if log and nw % 50 == 0:
            log(f"  training wakes {nw}: {n} samples")
    F = torch.cat(Fs)[:n_samples].contiguous()
    Y = torch.cat(Ys)[:n_samples].contiguous()
    torch.cuda.synchronize()
    return TrainingSet(F, Y, nw, inter, time.perf_counter() - t0)


def march_psi(tables, p: int) -> float:
    """Angle of the frame turning with the rotor at time index p: blade 0's
    azimuth (fastout.Azimuth) plus 90 degrees, so that psi = 0 puts blade 0
    along +y (the march's own convention, psi_of in march.hip)."""
    return math.radians(float(tables.azimuth[p - 1])) + 0.5 * math.pi


def march_wake(m) -> tuple:
    """The wake of a winds.march.March at its next time index p as a
    RotorWake (circulation of time p: time p-1's values moved one age older,
    age 0 repeated; cores zero, which the cut-off kernels ignore), the
    turning-frame angle and the rotor centre at p."""
    r = m.tables
    inf = m.info()
    p, Ag = inf["p"], inf["ages_solved"]
    nodes, gs, gt = m.state()
    A = nodes.shape[2]
    gs = gs[: r.ns * Ag * r.nb].reshape((r.ns, 1, Ag, r.nb), order="F")
    gt = gt[: r.nst * (Ag - 1) * r.nb].reshape((r.nst, 1, Ag - 1, r.nb), order="F")
    keep = A - 1 if Ag == A else Ag       # a truncated wake drops its oldest age
    gs_p = np.concatenate([gs[:, :, :1], gs[:, :, :keep]], axis=2)
    gt_p = np.concatenate([gt[:, :, :1], gt[:, :, :keep - 1]], axis=2)
    w = gen.RotorWake(domain=np.asfortranarray(nodes), gamma_shed=np.asfortranarray(gs_p),
                      gamma_trail=np.asfortranarray(gt_p),
                      rc_shed=np.zeros_like(gs_p, order="F"), rc_trail=np.zeros_like(gt_p, order="F"),
                      dt=r.dt, omega=float(r.rot_speed[p - 1]) * math.pi / 30)
    return w, march_psi(r, p), np.asarray(r.hub)[p - 1]


def wake_samples(w, psi, opts, split, target_ages, max_ages, R, U, ev=None,
                 hub=(0.0, 0.0, 0.0)):
    """(raw inputs, raw targets, interactions) of one wake: the far segments'
    direct sum at the near-wake nodes and lifting-line points."""
    import torch
    from .device import Evaluator
    ev = ev or Evaluator(0)
    X, gt = march_layout(w)
    P = query_points(w, target_ages)
    segs = far_segments(w, split)
    Xd = torch.from_numpy(X).cuda()
    gtd = torch.from_numpy(gt).cuda()
    Pd = torch.from_numpy(P).cuda()
    Sd = torch.from_numpy(np.ascontiguousarray(segs)).cuda()
    u = ev.eval(Pd, Sd, opts)
    g = geom(w.nb, w.nst, w.ages, split, max_ages, R, U, psi, hub)
    return features(g, Xd, gtd, Pd), labels(g, u), P.shape[1] * segs.shape[0]


def march_training_set(m, opts, steps: int, every: int, split: int, target_ages: int,
                       log=None) -> TrainingSet:
    """Direct-sum samples from the engine's own march: advance m by `steps`
    steps and take one snapshot of the far wake's velocity at the near
    targets every `every` steps once the wake is longer than split + 1."""
    import torch
    from .device import Evaluator
    ev = Evaluator(0)
    r = m.tables
    max_ages = m.cfg.max_ages
    Fs, Ys, inter, snaps = [], [], 0, 0
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(every, steps - done)
        m.step(k)
        m.sync()
        done += k
        if m.info()["ages"] <= split + 1:
            continue
        w, psi, hub = march_wake(m)
        U = float(np.linalg.norm(np.asarray(r.wind)[m.info()["p"] - 1]))
        F, Y, n = wake_samples(w, psi, opts, split, target_ages, max_ages, r.tip_radius, U, ev,
                               hub=hub)
        Fs.append(F)
        Ys.append(Y)
        inter += n
        snaps += 1
        if log and snaps % 10 == 0:
            log(f"  march snapshots {snaps} (step {done})")
    torch.cuda.synchronize()
    return TrainingSet(torch.cat(Fs).contiguous(), torch.cat(Ys).contiguous(), snaps, inter,
                       time.perf_counter() - t0)


def operating_points(n: int, seed: int = 17, tsr_range=(4.0, 10.0),
                     yaw_range=(0.0, 30.0)) -> list:
    """n seeded (TSR, yaw in degrees) pairs, stratified: the TSR x yaw box
    cut into 4 x (n // 4) cells (n a multiple of 4), one uniform point in
    each, so that every region of the box has training wakes."""
    if n % 4:
        raise ValueError("operating_points: n must be a multiple of 4")
    rng = np.random.default_rng(seed)
    nt, ny = 4, n // 4
    (t0, t1), (y0, y1) = tsr_range, yaw_range
    return [(t0 + (t1 - t0) * (i + rng.uniform()) / nt, y0 + (y1 - y0) * (j + rng.uniform()) / ny)
            for i in range(nt) for j in range(ny)]


def marched_training_set(points, opts, max_ages: int = 1440, split: int = 144,
                         target_ages: int = 72, sample_revs: int = 2, every: int = 6,
                         nb: int = 3, nst: int = 31, U: float = 11.4, direct_far: bool = True,
                         log=None) -> TrainingSet:
    """Direct-sum samples of the engine's own marches, one per operating
    point (tsr, yaw_deg) of `points`: a synthetic rotor (winds.cases.
    synthetic_rotor, designed for that TSR) marched with PCC until its wake
    has max_ages ages, then sample_revs more revolutions with one snapshot of
    the far wake's velocity at the near targets every `every` steps.  With
    direct_far the wake is developed by the surrogate's direct-far mode (near
    sum every step, older nodes' velocities refreshed once per revolution:
    ~10x cheaper per step at 1,440 ages); the snapshots are full direct sums
    either way."""
    import torch
    from . import cases
    from .march import March
    per_rev = 72  # the synthetic rotor turns 5 degrees per step
    Fs, Ys, inter, snaps = [], [], 0, 0
    t0 = time.perf_counter()
    for i, (tsr, yaw) in enumerate(points):
        tb = cases.tables(cases.synthetic_rotor(nb=nb, nst=nst, tsr=tsr, U=U, yaw_deg=yaw,
                                                n_times=max_ages + (sample_revs + 1) * per_rev))
        m = March(tb, max_ages, opts, integrator="pcc")
        if direct_far:
            m.set_surrogate(MarchSurrogate(mode=1, split=split, target_ages=target_ages,
                                           check_every=per_rev))
        m.step(max_ages)
        ts = march_training_set(m, opts, steps=sample_revs * per_rev, every=every, split=split,
                                target_ages=target_ages)
        del m
        Fs.append(ts.F)
        Ys.append(ts.Y)
        inter += ts.interactions
        snaps += ts.wakes
        if log:
            log(f"  marched wake {i + 1}/{len(points)} (TSR {tsr:.2f}, yaw {yaw:.1f} deg): "
                f"{ts.F.shape[0]} samples, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    return TrainingSet(torch.cat(Fs).contiguous(), torch.cat(Ys).contiguous(), snaps, inter,
                       time.perf_counter() - t0)


def concat(*sets: TrainingSet) -> TrainingSet:
    import torch
    return TrainingSet(torch.cat([s.F for s in sets]).contiguous(),
                       torch.cat([s.Y for s in sets]).contiguous(),
                       sum(s.wakes for s in sets), sum(s.interactions for s in sets),
                       sum(s.seconds for s in sets))


# ---------------------------------------------------------------- normalisation

def fit_norm(F, Y):
    """in_norm [2, NIN] = (mean, 1/std) and out_norm [2, 3] = (mean, std),
    float32 numpy; std floored at 1e-2 so constant inputs stay bounded."""
    import torch
    fm = F.double().mean(0)
    fs = F.double().std(0).clamp_min(1e-2)
    ym = Y.double().mean(0)
    ys = Y.double().std(0).clamp_min(1e-6)
    in_norm = torch.stack([fm, 1.0 / fs]).float().cpu().numpy()
    out_norm = torch.stack([ym, ys]).float().cpu().numpy()
    return np.ascontiguousarray(in_norm), np.ascontiguousarray(out_norm)


def normalise(F, Y, in_norm, out_norm):
    import torch
    inn = torch.from_numpy(in_norm).to(F.device)
    on = torch.from_numpy(out_norm).to(F.device)
    Fn = ((F - inn[0]) * inn[1]).contiguous()
    Yn = ((Y - on[0]) / on[1]).contiguous() if Y is not None else None
    return Fn, Yn


def train(net, Fn, Yn, steps: int, batch: int, lr: float = 1e-3, lr_end: float = 3e-5,
          seed: int = 7, log=None):
    """Adam on shuffled minibatches with a cosine learning-rate decay.
    Returns (seconds, final mean loss over the last 100 steps)."""
    import torch
    n = Fn.shape[0]
    g = torch.Generator(device=Fn.device)
    g.manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=Fn.device)
    acc = torch.zeros(1, dtype=torch.float64, device=Fn.device)
    perm = torch.randperm(n, device=Fn.device, generator=g)
    pos = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        if pos + batch > n:
            perm = torch.randperm(n, device=Fn.device, generator=g)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        xb = Fn.index_select(0, idx)
        yb = Yn.index_select(0, idx)
        r = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * it / max(1, steps)))
        # the loss (an atomic sum in the step) only where it is read
        logged = log and (it + 1) % 1000 == 0
        want = it >= steps - 100 or logged
        net.train_step(xb, yb, r, loss if want else None)
        if it >= steps - 100:
            acc += loss
        if logged:
            log(f"  train step {it + 1}: loss {float(loss.item()):.3e}")
    torch.cuda.synchronize()
    return time.perf_counter() - t0, float(acc.item()) / min(100, steps)


def train_two_phase(net, synthetic: TrainingSet, march: TrainingSet, steps: int, batch: int,
                    log=None, between=None):
    """Normalisation from both sets; half the steps on the synthetic wakes
    (rate 1e-3 -> 3e-5), half on the march's own samples (3e-4 -> 1e-5).
    between(net, in_norm, out_norm), if given, runs after the first phase.
    Returns (in_norm, out_norm, seconds of each phase, final mean loss of
    each phase)."""
    both = concat(synthetic, march)
    in_norm, out_norm = fit_norm(both.F, both.Y)
    del both
    Fs, Ys = normalise(synthetic.F, synthetic.Y, in_norm, out_norm)
    s1, l1 = train(net, Fs, Ys, steps=steps // 2, batch=batch, lr=1e-3, lr_end=3e-5, log=log)
    del Fs, Ys
    if between is not None:
        between(net, in_norm, out_norm)
    Fm, Ym = normalise(march.F, march.Y, in_norm, out_norm)
    s2, l2 = train(net, Fm, Ym, steps=steps - steps // 2, batch=batch, lr=3e-4, lr_end=1e-5,
                   log=log)
    return in_norm, out_norm, (s1, s2), (l1, l2)


def predict(net, Fn, out_norm, chunk: int | None = None):
    """Network output in raw label units (u / U in the turning frame)."""
    import torch
    chunk = chunk or net.max_batch
    on = torch.from_numpy(out_norm).to(Fn.device)
    ys = [net.forward(Fn[i:i + chunk].contiguous()) for i in range(0, Fn.shape[0], chunk)]
    return torch.cat(ys) * on[1] + on[0]


def normwise_error(Yp, Y):
    """max_i |y_i - y_ref_i| / max_i |y_ref_i| over the 3-vectors."""
    import torch
    e = torch.linalg.vector_norm((Yp - Y).double(), dim=1).max()
    r = torch.linalg.vector_norm(Y.double(), dim=1).max()
    return float(e / r)


def rms_error(Yp, Y):
    """sqrt(mean |y_i - y_ref_i|^2) / sqrt(mean |y_ref_i|^2) over the 3-vectors."""
    import torch
    e = ((Yp - Y).double() ** 2).sum(1).mean().sqrt()
    r = (Y.double() ** 2).sum(1).mean().sqrt()
    return float(e / r)


# ---------------------------------------------------------------- march setup

@dataclass
class MarchSurrogate:
    """Configuration of winds_march_set_surrogate."""
    mode: int = 0                 # 0 network, 1 direct far sum, 2 far wake dropped
    split: int = 144              # 2 revolutions at 5 deg steps
    target_ages: int = 72         # near targets: 1 revolution
    check_every: int = 72         # one direct-sum check per revolution
    train_iters: int = 40
    tol: float = 0.01
    lr: float = 1e-4
