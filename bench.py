"""Benchmark of the induced-velocity engine on one MI355X.

A step is one induced-velocity evaluation of the single large wake
(3 blades x 31 cosine stations x 1,440 ages = 133,920 targets; shed and
trailing filaments together = 263,427 segments; Vatistas n=2 core with the
time-grown core radius; BiotSavart.m semantics) in FP32 with compensated
sums and FP64 near pairs.  Inputs are resident in HBM before timing starts.

Prints one JSON line: G interactions/s, the FP32/FP64 error of the result
against the FP64 engine (itself checked against the CPU oracle on a sample),
the roofline of the dominant kernel, and the CPU baseline (the oracle's
restatement of the reference's CPU loop, timed on this host).
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

T_START = time.monotonic()
ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "winds-master_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Biot–Savart G interactions/s & free-wake step ms vs FP64 plain, at stated err"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA peak
FP64_PEAK_TFLOPS = 78.6
FLOP_VATISTAS2 = 59        # algorithmic FLOPs per interaction, Vatistas n=2 form
# HBM bytes per launch of bs_fp32c<2,1,false> on this workload: FETCH_SIZE /
# 0.500 + WRITE_SIZE / 1.000 (rocprofv3 --pmc, factors from the plain
# streaming kernel bs_reduce_split), profiles/bs_fp32c_vatistas2_single_wake_133920x263427.md
TRAFFIC_BYTES = 1_179_900 * 1024
FLOP_CUTOFF = 47           # cut-off (MEX) form


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Budget:
    """Wall-clock budget of the whole run (seconds from process start).  A
    stage whose estimated time does not fit in what is left is skipped and
    named under "skipped" in the output line, so the line is printed in time
    whatever the host's speed."""

    def __init__(self, seconds: float):
        self.seconds = seconds
        self.skipped = []

    def left(self) -> float:
        return self.seconds - (time.monotonic() - T_START)

    def fits(self, stage: str, est_s: float, reserve_s: float = 20.0) -> bool:
        ok = est_s + reserve_s <= self.left()
        if not ok:
            self.skipped.append(f"{stage} (estimated {est_s:.0f} s, {self.left():.0f} s left)")
            log(f"skipping {stage}: estimated {est_s:.0f} s, {self.left():.0f} s left")
        return ok


def host_info() -> dict:
    """Host cores this process may use and the CPU model string."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    # OMP_NUM_THREADS, where set, is the host share given to this process
    # (16 per GPU on the GPU hosts); otherwise every core of the affinity mask
    threads = aff
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(aff, int(omp))
    return {"threads": max(1, threads), "affinity_cores": aff, "os_cpu_count": os.cpu_count(),
            "omp_num_threads": omp, "cpu_model": model}


def cpu_baseline(wake, budget_s: float = 12.0) -> dict:
    """The oracle's restatement of BiotSavart_nogpu (BiotSavartMex.cu:268-282,
    FP64, cut-off core): config 1 (3,888 x 7,506) in full on one thread and
    on all the host cores this process has, and a bounded sample of the
    single large wake (the GPU workload) on all of them."""
    import numpy as np
    from oracle import bs as O
    from winds import gen, layout

    hi = host_info()
    threads = hi["threads"]
    res = {}
    # config 1, whole evaluation (shed + trailing segments, all nodes)
    c1 = gen.config1()
    r1 = layout.wake_segments(c1)
    F1c = np.asfortranarray(r1[:, 0:3].reshape(-1, 3, 1, 1))
    F2c = np.asfortranarray(r1[:, 3:6].reshape(-1, 3, 1, 1))
    g1 = r1[:, 6].copy()
    inter1 = c1.n_targets * r1.shape[0]
    for label, th in (("config1_all", threads), ("config1_one", 1)):
        O.bs_mex(F1c, F2c, c1.domain, g1, 1e-4, threads=th)
        reps, t0 = 0, time.perf_counter()
        while True:
            O.bs_mex(F1c, F2c, c1.domain, g1, 1e-4, threads=th)
            reps += 1
            if time.perf_counter() - t0 > 2.0:
                break
        dt = (time.perf_counter() - t0) / reps
        res[label] = dict(rate=inter1 / dt / 1e9, ms=dt * 1e3, threads=th)
    # the single large wake: first n targets x all segments
    recs = layout.wake_segments(wake)
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    g = recs[:, 6].copy()
    P = wake.domain
    nseg = recs.shape[0]
    nt = wake.n_targets
    n = 16 * threads
    t0 = time.perf_counter()
    O.bs_mex_range(F1, F2, P, g, 1e-4, 0, n, threads=threads)
    dt = time.perf_counter() - t0
    n2 = int(min(nt, max(n, n * budget_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    O.bs_mex_range(F1, F2, P, g, 1e-4, 0, n2, threads=threads)
    dt = time.perf_counter() - t0
    rate = n2 * nseg / dt / 1e9
    a1, o1 = res["config1_all"], res["config1_one"]
    return {
        "value": round(rate, 4),
        "unit": "G interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {n2} of {nt} targets x all {nseg} segments of the single large wake, "
                   f"cut-off core (BiotSavart_nogpu arithmetic, FP64, the oracle's C "
                   f"restatement), {dt:.1f} s on {threads} OpenMP threads ({hi['cpu_model']}); "
                   f"config 1 in full: {a1['ms']:.1f} ms on {threads} threads, "
                   f"{o1['ms']:.1f} ms on 1 thread"),
        "host": hi,
        "config1_full_ms": {"threads": threads, "all_cores": round(a1["ms"], 2),
                            "one_thread": round(o1["ms"], 2),
                            "interactions": inter1},
        "config1_G_int_s": {"all_cores": round(a1["rate"], 4), "one_thread": round(o1["rate"], 4)},
    }


def batch_bench(o32, o64, steps: int) -> dict:
    """64 uneven wakes (blades 2-3, 12-40 stations, 200-2000 ages, TSR 4-10,
    yaw 0-30 deg) in one persistent launch, autotuned, replayed from a hipGraph."""
    import numpy as np
    import torch

    from winds import gen, layout
    from winds.batch import Batch
    from winds.device import Evaluator

    t0 = time.perf_counter()
    wakes = gen.config3()
    b = Batch(wakes, o32)
    info = b.info()
    log(f"batch: {len(wakes)} wakes, {info} (setup {time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    tune, tune_ms = b.autotune()
    log(f"batch autotune: {tune} {tune_ms:.1f} ms ({time.perf_counter() - t0:.1f} s)")
    b.eval(use_graph=True)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = max(3, steps // 4)
    e0.record(st)
    for _ in range(n):
        b.eval(use_graph=True)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    rate = info["interactions"] / (ms * 1e-3) / 1e9
    log(f"batch: {ms:.1f} ms per batch evaluation, {rate:.1f} G int/s")
    n_fix = b.guard_count()
    u32 = b.out.clone()
    # every wake against its own FP64 evaluation: the FP64 batch (one launch,
    # each wake bitwise equal to its single-wake FP64 evaluation)
    b64 = Batch(wakes, o64)
    b64.eval(use_graph=False)
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    u64 = b64.eval(use_graph=False)
    torch.cuda.synchronize()
    ms64 = (time.perf_counter() - f0) * 1e3
    worst = {"max_rel": 0.0, "rms_rel": 0.0, "normwise": 0.0, "wake_of_max": -1}
    for i, (a, r) in enumerate(zip(b64.split(u32), b64.split(u64))):
        ref = r.cpu().numpy()
        u = a.cpu().numpy()
        nrm = np.linalg.norm(ref, axis=0)
        dif = np.linalg.norm(u - ref, axis=0)
        m = nrm >= 1e-6 * nrm.max()
        e = dif[m] / nrm[m]
        if float(e.max()) > worst["max_rel"]:
            worst["max_rel"] = float(e.max())
            worst["wake_of_max"] = i
        worst["rms_rel"] = max(worst["rms_rel"], float(np.sqrt(np.mean(e * e))))
        worst["normwise"] = max(worst["normwise"], float(dif.max() / nrm.max()))
    log(f"batch error (all {len(wakes)} wakes vs FP64): {worst}; guard re-evaluated {n_fix} "
        f"targets; FP64 batch {ms64:.0f} ms")
    del b64
    return {"value": round(rate, 2), "unit": "G interactions/s", "ms_per_batch": round(ms, 3),
            "wakes": len(wakes), "targets": info["targets"], "segments": info["segments"],
            "interactions": info["interactions"], "work_items": info["items"],
            "tune": tune, "graph": True,
            "error_all_wakes_vs_fp64": worst,
            "guard_fp64_targets": n_fix,
            "fp64_batch_ms": round(ms64, 1),
            "fp64_batch_G_int_s": round(info["interactions"] / (ms64 * 1e-3) / 1e9, 2)}


def march_opts(prec):
    import winds
    # WInDS defaults: cut-off core d = 1e-4 (WInDS.m:42) through the MEX path
    return winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=prec, d=1e-4)


MLP_HIDDEN, MLP_LAYERS = 256, 4


def mlp_weights(n_in: int, hidden: int, layers: int, n_out: int) -> int:
    return n_in * hidden + (layers - 1) * hidden * hidden + hidden * n_out


def surrogate_bench(n_points: int, train_steps: int, march_set=None, batch: int = 4096):
    """Far-wake network: direct-sum training data from the engine's own
    marches at n_points stratified operating points (TSR 4-10 x yaw 0-30 deg,
    1,440-age wakes, split at 144 ages = 2 revolutions), training with the
    engine's BF16 MFMA kernels; its error on march_set, the benchmark march's
    own direct sums (TSR 7, no yaw: an operating point it was not trained
    on); inference rate.  Returns (report, network, in_norm, out_norm)."""
    import torch

    import winds
    from winds import surrogate as S
    from winds.mlp import MLP

    o = march_opts(winds.PREC_FP32C)
    sc = S.MarchSurrogate()
    pts = S.operating_points(n_points, seed=17)
    ts = S.marched_training_set(pts, o, max_ages=1440, split=sc.split,
                                target_ages=sc.target_ages, every=8, log=log)
    log(f"surrogate data: {len(pts)} marched wakes, {ts.wakes} snapshots, {ts.F.shape[0]} samples, "
        f"{ts.interactions / 1e9:.1f} G interactions, {ts.seconds:.1f} s")
    rows = 31 * sc.target_ages * 3 + 30 * 3
    net = MLP(S.NIN, MLP_HIDDEN, MLP_LAYERS, S.NOUT, max(batch, rows), seed=7)
    in_norm, out_norm = S.fit_norm(ts.F, ts.Y)
    Fn, Yn = S.normalise(ts.F, ts.Y, in_norm, out_norm)
    secs, loss = S.train(net, Fn, Yn, steps=train_steps, batch=batch, lr=1e-3, log=log)
    del Fn, Yn
    W = mlp_weights(S.NIN, MLP_HIDDEN, MLP_LAYERS, S.NOUT)
    # training: forward 2W, weight gradients 2W, input gradients 2W (less the first layer)
    train_flop = (6 * W - 2 * S.NIN * MLP_HIDDEN) * batch
    train_tflops = train_flop * train_steps / secs / 1e12
    Ft, _ = S.normalise(ts.F[:200_000].contiguous(), None, in_norm, out_norm)
    Yt = S.predict(net, Ft, out_norm)
    err_train, rms_train = S.normwise_error(Yt, ts.Y[:200_000]), S.rms_error(Yt, ts.Y[:200_000])
    err_march = rms_march = None
    if march_set is not None:
        Fm, _ = S.normalise(march_set.F, None, in_norm, out_norm)
        Ym = S.predict(net, Fm, out_norm)
        err_march, rms_march = S.normwise_error(Ym, march_set.Y), S.rms_error(Ym, march_set.Y)
    log(f"surrogate training: {train_steps} steps x {batch}: {secs:.1f} s "
        f"({train_tflops:.1f} TFLOP/s), loss {loss:.3e}, normwise error train "
        f"{err_train:.3e}, held-out operating point {err_march}; rms {rms_train:.3e} / "
        f"{rms_march}")
    # inference rate on max_batch rows
    B = net.max_batch
    Xb = Ft[:B].contiguous()
    Yb = torch.empty((B, S.NOUT), dtype=torch.float32, device=Xb.device)
    for _ in range(3):
        net.forward(Xb, Yb)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    e0.record(st)
    for _ in range(reps):
        net.forward(Xb, Yb)
    e1.record(st)
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    inf_rate = B / (fwd_ms * 1e-3)
    inf_tflops = 2 * W * B / (fwd_ms * 1e-3) / 1e12
    log(f"surrogate inference: {B} rows in {fwd_ms:.3f} ms, {inf_rate / 1e6:.1f} M samples/s, "
        f"{inf_tflops:.1f} TFLOP/s")
    rep = {
        "network": f"{S.NIN} inputs -> {MLP_LAYERS} x {MLP_HIDDEN} SiLU -> 3, BF16 weights and "
                   f"activations, FP32 accumulation (v_mfma_f32_32x32x16_bf16)",
        "parameters": net.num_params(),
        "split_age": sc.split, "target_ages": sc.target_ages,
        "training": ("direct sums of the engine's own marches at stratified operating "
                     "points (TSR 4-10 x yaw 0-30 deg; 1,440-age wakes developed by the "
                     "direct-far march, 2 revolutions sampled every 8 steps); the benchmark "
                     "march's operating point (TSR 7, yaw 0) is not among them"),
        "operating_points": [[round(t, 3), round(y, 2)] for t, y in pts],
        "samples": int(ts.F.shape[0]), "snapshots": ts.wakes,
        "data_G_interactions": round(ts.interactions / 1e9, 2),
        "data_s": round(ts.seconds, 2),
        "train_steps": train_steps, "batch": batch, "train_s": round(secs, 2),
        "train_samples_per_s": round(train_steps * batch / secs, 1),
        "train_tflops": round(train_tflops, 2),
        "final_loss": loss,
        "error_metrics": "normwise = max |error| / max |label|; rms = rms |error| / rms |label| "
                         "(3-vectors of far-wake velocity over U)",
        "normwise_err_train": err_train, "rms_err_train": rms_train,
        "heldout": "the benchmark march's own direct sums (TSR 7, yaw 0)",
        "heldout_samples": int(march_set.F.shape[0]) if march_set is not None else 0,
        "heldout_snapshots": march_set.wakes if march_set is not None else 0,
        "normwise_err_heldout": err_march, "rms_err_heldout": rms_march,
        "inference_rows": B, "inference_ms": round(fwd_ms, 4),
        "inference_samples_per_s": round(inf_rate, 1),
        "inference_tflops": round(inf_tflops, 2),
        "roofline_inference": {"bound": "mfma", "achieved": round(inf_tflops, 2),
                               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": round(inf_tflops / BF16_PEAK_TFLOPS, 4)},
    }
    return rep, net, in_norm, out_norm


def march_bench(revs: int, dev_revs: int, surrogate: bool, n_points: int, train_steps: int,
                check_every: int = 72, fp64_revs: int = 2, total_steps: int = 3600,
                budget: "Budget | None" = None):
    """Free-wake time march of a 3-blade, 31-node rotor at 5 deg steps with
    PCC (2 wake evaluations per step), an FP64 Kutta-Joukowski solve and a
    1,440-age wake (~10 rotor diameters):
    (1) a cold start of dev_revs revolutions with FP32 compensated wake sums
        (then one revolution of direct-sum snapshots of its full-length far
        wake measures the network, trained on other operating points, here);
    (2) step time at the full length from that state: FP32, FP64, surrogate
        (network), surrogate with the direct far sum instead of the network;
    (3) `revs` more revolutions from that state: FP32, the surrogate (with
        its online training at each check) and the surrogate with the network
        as trained on the other operating points (no online training), and
        the first fp64_revs of them in FP64 as the reference; the surrogate
        march goes on to total_steps steps in all (the workload's 3,600), its
        later revolutions set beside the FP32 march's last ones.
    Stages that do not fit in the budget are skipped (the budget names them).
    Returns (march report, surrogate report)."""
    import numpy as np

    import winds
    from winds import cases
    from winds import surrogate as S
    from winds.march import March

    per_rev = 72
    steps = revs * per_rev
    dev_steps = dev_revs * per_rev
    surr_steps = max(steps, total_steps - dev_steps) if surrogate else 0
    tb = cases.tables(cases.synthetic_rotor(nb=3, nst=31,
                                            n_times=dev_steps + max(steps, surr_steps) + 16))
    sc = S.MarchSurrogate(check_every=check_every)
    out = {"rotor": "3 blades x 31 nodes, R 63 m, TSR 7, U 11.4 m/s, 5 deg steps, PCC, "
                    "kinematics.m/velocity.m tables, BEM start",
           "max_ages": 1440, "develop_revs": dev_revs}
    o32 = march_opts(winds.PREC_FP32C)
    # (1) develop
    t0 = time.perf_counter()
    m = March(tb, 1440, o32, integrator="pcc")
    m.step(dev_steps)
    m.sync()
    dev_s = time.perf_counter() - t0
    log(f"march cold start {dev_revs} revs: {dev_s:.1f} s")
    inf = m.info()
    nodes, _, _ = m.state()
    A, Ag = inf["ages"], inf["ages_solved"]
    gs, gt = m.state_circulation(Ag)
    state = (inf["p"], nodes, gs[:, :, :A - 1], gt[:, :, :A - 2])
    out["develop_s_fp32c"] = round(dev_s, 2)
    # the network's data: one 1,440-step march per operating point in the
    # direct-far mode, ~0.23 of the FP32 cold start's time each
    if surrogate and budget is not None and not budget.fits(
            "far-wake network (training data, training, surrogate marches)",
            n_points * dev_s * 0.3 + 20):
        surrogate, surr_steps = False, 0
    msamp = None
    if surrogate:
        # the network's held-out data: direct sums of this march's far wake
        # at full length, one revolution from the developed state
        msamp = S.march_training_set(m, o32, steps=per_rev, every=8, split=sc.split,
                                     target_ages=sc.target_ages, log=log)
    del m

    surr_rep, net_args = None, None
    if surrogate:
        surr_rep, net, in_norm, out_norm = surrogate_bench(n_points, train_steps, msamp)
        net_args = (net, in_norm, out_norm)

    def fork(prec, surr_mode=None, train_iters=40):
        f = March(tb, 1440, march_opts(prec), integrator="pcc")
        f.set_state(*state)
        if surr_mode is not None:
            cfg = S.MarchSurrogate(mode=surr_mode, check_every=check_every,
                                   train_iters=train_iters)
            if surr_mode == 0:
                f.set_surrogate(cfg, *net_args)
            else:
                f.set_surrogate(cfg)
        return f

    # (2) step time at the full length
    runs = [("fp32c", winds.PREC_FP32C, 10, None), ("fp64", winds.PREC_FP64, 2, None)]
    if surrogate:
        runs += [("surrogate", winds.PREC_FP32C, 2 * check_every, 0),
                 ("surrogate_direct_far", winds.PREC_FP32C, 2 * check_every, 1)]
    for name, prec, nsteps, mode in runs:
        f = fork(prec, mode)
        f.step(1)
        f.sync()
        t0 = time.perf_counter()
        f.step(nsteps)
        f.sync()
        ms = (time.perf_counter() - t0) / nsteps * 1e3
        out[f"step_ms_{name}"] = round(ms, 2)
        log(f"march step at 1440 ages, {name}: {ms:.1f} ms")
        if mode is not None:
            out[f"{name}_stats_steady"] = f.surr_stats()
            out[f"{name}_speedup_vs_fp32c"] = round(out["step_ms_fp32c"] / ms, 2)
            log(f"  {name} {out[f'{name}_speedup_vs_fp32c']}x vs fp32c; "
                f"{out[f'{name}_stats_steady']}")
        del f
    # (2b) the same steps replayed from recorded HIP graphs (winds_march_set_graphs),
    # and a short-wake march where launches weigh more: NREL 5-MW, 72 ages
    def timed(f, warm, n):
        f.step(warm)
        f.sync()
        t0 = time.perf_counter()
        f.step(n)
        f.sync()
        return (time.perf_counter() - t0) / n * 1e3

    f = fork(winds.PREC_FP32C)
    f.set_graphs(True)
    out["step_ms_fp32c_graph"] = round(timed(f, 9, 12), 2)
    del f
    short = cases.tables(cases.nrel_rotor(t=(0.0, 30.0, 10.0)))
    g_ms = {}
    for mode in (False, True):
        f = March(short, 72, march_opts(winds.PREC_FP32C), integrator="pcc")
        f.set_graphs(mode)
        g_ms[mode] = timed(f, 90, 120)
        del f
    out["short_wake_step_ms"] = {"case": "NREL 5-MW, 18 stations, 72 ages, FP32C, PCC",
                                 "eager": round(g_ms[False], 3), "graph": round(g_ms[True], 3)}
    log(f"march step with graphs at 1440 ages: {out['step_ms_fp32c_graph']} ms; short wake "
        f"{out['short_wake_step_ms']}")
    # (3) revs more revolutions from the developed state
    hist = {}
    plan = [("fp32c", winds.PREC_FP32C, None, steps, 40)]
    if surrogate:
        plan += [("surrogate", winds.PREC_FP32C, 0, surr_steps, 40),
                 ("surrogate_pretrained_only", winds.PREC_FP32C, 0, steps, 0)]
    plan += [("fp64", winds.PREC_FP64, None, fp64_revs * per_rev, 40)]
    for name, prec, mode, nst, iters in plan:
        ms_step = out[f"step_ms_{'surrogate' if mode is not None else name}"]
        if budget is not None and not budget.fits(f"continued march {name} ({nst} steps)",
                                                  nst * ms_step * 1.2e-3 + 3):
            continue
        t0 = time.perf_counter()
        f = fork(prec, mode, iters)
        done = 0
        while done < nst:
            k = min(per_rev, nst - done)
            f.step(k)
            f.sync()
            done += k
            log(f"  continued march {name}: {done}/{nst} steps")
        hist[name] = f.perf()
        el = time.perf_counter() - t0
        out[f"continue_s_{name}"] = round(el, 2)
        log(f"march {nst} more steps {name}: {el:.1f} s")
        if mode is not None:
            out[f"{name}_stats_continue"] = f.surr_stats()
        del f

    def errs(a, b, n):
        a, b = a[:n], b[:n]
        rel = lambda i: float(np.max(np.abs(a[:, i] - b[:, i]) / np.abs(b[:, i])))
        meanr = lambda i: float(abs(np.mean(a[:, i]) - np.mean(b[:, i])) / abs(np.mean(b[:, i])))
        return {"thrust_max_rel": rel(1), "power_max_rel": rel(3), "thrust_mean_rel": meanr(1),
                "power_mean_rel": meanr(3), "steps": int(n)}

    n64 = fp64_revs * per_rev if "fp64" in hist else 0
    b64 = hist.get("fp64")
    if b64 is not None:
        out["thrust_N_mean_fp64"] = float(np.mean(b64[:, 1]))
        out["power_W_mean_fp64"] = float(np.mean(b64[:, 3]))
        out["kj_iterations_mean_fp64"] = float(np.mean(b64[:, 4]))
        if "fp32c" in hist:
            out["fp32c_vs_fp64"] = errs(hist["fp32c"], b64, n64)
            log(f"fp32c vs fp64 over {fp64_revs} revs at full length: {out['fp32c_vs_fp64']}")
    for name in ("surrogate", "surrogate_pretrained_only"):
        if name not in hist:
            continue
        if b64 is not None:
            out[f"{name}_vs_fp64"] = errs(hist[name], b64, n64)
            log(f"{name} vs fp64 over {fp64_revs} revs: {out[f'{name}_vs_fp64']}")
        if "fp32c" in hist:
            out[f"{name}_vs_fp32c"] = errs(hist[name], hist["fp32c"], steps)
            log(f"{name} vs fp32c over {revs} revs: {out[f'{name}_vs_fp32c']}")
    if "surrogate" in hist and "fp32c" in hist:
        # the surrogate's last revolution against the FP32 march's last one
        # (the rotor is periodic: same azimuths, same inflow)
        hs, hf = hist["surrogate"], hist["fp32c"]
        last_s, last_f = hs[-per_rev:], hf[-per_rev:]
        out["surrogate_last_rev_vs_fp32c_last_rev"] = {
            "surrogate_revs": int(surr_steps // per_rev), "fp32c_revs": int(revs),
            "thrust_mean_rel": float(abs(np.mean(last_s[:, 1]) - np.mean(last_f[:, 1]))
                                     / abs(np.mean(last_f[:, 1]))),
            "power_mean_rel": float(abs(np.mean(last_s[:, 3]) - np.mean(last_f[:, 3]))
                                    / abs(np.mean(last_f[:, 3])))}
        log(f"surrogate revolution {surr_steps // per_rev} vs fp32c revolution {revs}: "
            f"{out['surrogate_last_rev_vs_fp32c_last_rev']}")
    ran = lambda name, n: n if name in hist else 0
    out.update({"revs": revs, "steps": steps,
                "config5_steps_covered": (f"{dev_steps + ran('fp32c', steps)} of 3600 steps in "
                                          f"FP32 ({dev_steps} cold start + "
                                          f"{ran('fp32c', steps)} continued), "
                                          f"{dev_steps + ran('surrogate', surr_steps)} with the "
                                          f"surrogate ({dev_steps} FP32 cold start + "
                                          f"{ran('surrogate', surr_steps)}), "
                                          f"{n64} in FP64 from the developed 1,440-age state")})
    if "fp32c" in hist:
        out["thrust_N_mean_fp32c"] = float(np.mean(hist["fp32c"][:, 1]))
        out["power_W_mean_fp32c"] = float(np.mean(hist["fp32c"][:, 3]))
    return out, surr_rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-wake batch")
    ap.add_argument("--no-march", action="store_true", help="skip the free-wake time march")
    ap.add_argument("--march-revs", type=int, default=10)
    ap.add_argument("--fp64-revs", type=int, default=2,
                    help="revolutions of the FP64 reference march from the developed state")
    ap.add_argument("--develop-revs", type=int, default=20,
                    help="cold-start revolutions before the full-length measurements")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the far-wake network")
    ap.add_argument("--surrogate-points", type=int, default=16,
                    help="operating points whose marched wakes train the far-wake network "
                         "(a multiple of 4)")
    ap.add_argument("--surrogate-steps", type=int, default=8000)
    ap.add_argument("--check-every", type=int, default=72,
                    help="surrogate steps between direct-sum checks")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 2 GiB device-to-device copy (reference point for "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--ages", type=int, default=1440)
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="seconds from process start; later stages that would not finish "
                         "within it are skipped and listed under \"skipped\"")
    args = ap.parse_args()
    budget = Budget(args.time_budget)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine "
              "is single-GPU in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import winds
    from winds import gen, layout
    from winds.device import Evaluator
    from oracle import bs as O

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    wake = gen.rotor_wake(nb=3, nst=31, ages=args.ages)
    # trailing filaments line by line: the engine's own order (march.hip)
    recs = layout.wake_segments(wake, chained=True)
    nt, nseg = wake.n_targets, recs.shape[0]
    inter = nt * nseg
    log(f"workload: {nt} targets x {nseg} segments = {inter / 1e9:.3f} G interactions")
    tg = torch.from_numpy(layout.wake_targets(wake)).to(dev).contiguous()
    sg = torch.from_numpy(recs).to(dev).contiguous()
    ev = Evaluator(0)
    o32 = winds.opts(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB,
                     precision=winds.PREC_FP32C, co=math.inf)
    o64 = winds.opts(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB,
                     precision=winds.PREC_FP64, co=math.inf)
    out = torch.empty((3, nt), dtype=torch.float64, device=dev)

    # FP64 engine result (reference for the error), checked on a sample;
    # the second evaluation is timed (the first loads the code)
    u64 = ev.eval(tg, sg, o64).clone()
    torch.cuda.synchronize()
    f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    f0.record(torch.cuda.current_stream(dev))
    ev.eval(tg, sg, o64, out=u64)
    f1.record(torch.cuda.current_stream(dev))
    torch.cuda.synchronize()
    t64 = f0.elapsed_time(f1) * 1e-3
    log(f"fp64 evaluation: {t64 * 1e3:.1f} ms")
    rng = np.random.default_rng(11)
    sample = np.sort(rng.choice(nt, size=48, replace=False))
    Psamp = np.asfortranarray(layout.wake_targets(wake)[:, sample].T.reshape(-1, 3, 1, 1))
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    ref_s = O.bs_matlab(F1, F2, Psamp, recs[:, 6], recs[:, 7], "visc2")[:, :, 0, 0]
    got_s = u64.cpu().numpy()[:, sample].T
    fp64_vs_oracle = float(np.max(np.abs(got_s - ref_s)) / np.max(np.abs(ref_s)))
    log(f"fp64 engine vs CPU oracle on {sample.size} targets: normwise {fp64_vs_oracle:.2e}")

    # timed FP32 compensated evaluations
    for _ in range(args.warmup):
        ev.eval(tg, sg, o32, out=out)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(st)
    for _ in range(args.steps):
        ev.eval(tg, sg, o32, out=out)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1) / args.steps
    rate = inter / (ms * 1e-3) / 1e9
    log(f"fp32c: {ms:.3f} ms/evaluation (wall {wall / args.steps * 1e3:.3f}), {rate:.1f} G int/s")

    # error of the timed result
    u32 = out.cpu().numpy()
    ref = u64.cpu().numpy()
    nrm = np.linalg.norm(ref, axis=0)
    dif = np.linalg.norm(u32 - ref, axis=0)
    m = nrm >= 1e-6 * nrm.max()
    e = dif[m] / nrm[m]
    err = {"max_rel": float(e.max()), "rms_rel": float(np.sqrt(np.mean(e * e))),
           "normwise": float(dif.max() / nrm.max()), "vs": "FP64 engine (BiotSavart.m arithmetic)",
           "fp64_vs_cpu_oracle_normwise": fp64_vs_oracle, "targets": int(m.sum())}
    n_fix = ev.guard_count()
    err["guard_fp64_targets"] = n_fix
    log(f"fp32c error vs fp64: {err}")
    # the same evaluation with the cancellation guard off (its cost)
    og = winds.opts(core=winds.CORE_VATISTAS, vat_n=2, semantics=winds.SEM_MATLAB,
                    precision=winds.PREC_FP32C, co=math.inf, guard_kappa=-1.0)
    ev.eval(tg, sg, og, out=out)
    torch.cuda.synchronize()
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g0.record(st)
    for _ in range(max(3, args.steps // 4)):
        ev.eval(tg, sg, og, out=out)
    g1.record(st)
    torch.cuda.synchronize()
    ms_noguard = g0.elapsed_time(g1) / max(3, args.steps // 4)
    log(f"fp32c without the guard: {ms_noguard:.3f} ms")

    # cut-off (MEX semantics) rate, same workload
    oc = winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=winds.PREC_FP32C,
                    d=1e-4)
    ev.eval(tg, sg, oc, out=out)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nc = max(3, args.steps // 4)
    c0.record(st)
    for _ in range(nc):
        ev.eval(tg, sg, oc, out=out)
    c1.record(st)
    torch.cuda.synchronize()
    ms_c = c0.elapsed_time(c1) / nc
    rate_c = inter / (ms_c * 1e-3) / 1e9
    # its error against the FP64 MEX arithmetic (biot_p2p, bitwise the CPU oracle's)
    uc = out.cpu().numpy()
    n_fix_c = ev.guard_count()
    o64c = winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=winds.PREC_FP64,
                      d=1e-4)
    refc = ev.eval(tg, sg, o64c).cpu().numpy()
    nrm_c = np.linalg.norm(refc, axis=0)
    dif_c = np.linalg.norm(uc - refc, axis=0)
    mc_ = nrm_c >= 1e-6 * nrm_c.max()
    ec = dif_c[mc_] / nrm_c[mc_]
    err_c = {"max_rel": float(ec.max()), "rms_rel": float(np.sqrt(np.mean(ec * ec))),
             "normwise": float(dif_c.max() / nrm_c.max()), "vs": "FP64 engine (biot_p2p)",
             "guard_fp64_targets": n_fix_c}
    log(f"cut-off: {ms_c:.3f} ms, {rate_c:.1f} G int/s, error vs fp64 {err_c}")

    # the drop-in call itself: host arrays in MATLAB layout, the same segments
    # and targets, Vatistas-2 with BiotSavart.m semantics, transfers included
    gam, rcs = recs[:, 6].copy(), recs[:, 7].copy()
    call = (F1, F2, wake.domain, gam, rcs, 2, "visc", 1e300, "fulc", "true", (256, 1, 0))
    winds.BiotSavartMex(*call)
    na = 3
    a0 = time.perf_counter()
    for _ in range(na):
        winds.BiotSavartMex(*call)
    abi_ms = (time.perf_counter() - a0) / na * 1e3
    log(f"drop-in BiotSavartMex call with host arrays: {abi_ms:.1f} ms (transfers included)")

    if args.calib_copy:
        src = torch.ones(2 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cb = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(wake)
        # same core and arithmetic as the CPU path: the GPU cut-off rate
        cb["gpu_cutoff_over_cpu_all_cores"] = round(rate_c / cb["value"], 1)

    batch = None
    if not args.no_batch and budget.fits("64-wake batch", 60):
        batch = batch_bench(o32, o64, args.steps)
    march, surr_rep = None, None
    # cold start of the march: ~20 ms per step at 1,440 ages and fewer below
    if not args.no_march and budget.fits("free-wake march", args.develop_revs * 72 * 0.03 + 30):
        march, surr_rep = march_bench(args.march_revs, args.develop_revs, not args.no_surrogate,
                                      args.surrogate_points, args.surrogate_steps,
                                      args.check_every, args.fp64_revs, budget=budget)

    achieved = inter * FLOP_VATISTAS2 / (ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(rate, 2),
        "unit": "G interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"single large wake, FP32 compensated vs FP64: 3 blades x 31 stations x "
                         f"{args.ages} ages, {nt} targets x {nseg} segments (shed+trail) per "
                         f"evaluation, Vatistas n=2 core, FP64 for pairs within max(40 r_c, 0.51 |L|)"),
            "targets": nt, "segments": nseg, "interactions_per_step": inter,
            "core": "vatistas2", "precision": "fp32 compensated + fp64 near pairs",
        },
        "error": err,
        "cutoff_mex": {"value": round(rate_c, 2), "unit": "G interactions/s",
                       "ms_per_step": round(ms_c, 4),
                       "tflops": round(inter * FLOP_CUTOFF / (ms_c * 1e-3) / 1e12, 2),
                       "frac_of_fp32_peak": round(inter * FLOP_CUTOFF / (ms_c * 1e-3) / 1e12
                                                  / FP32_PEAK_TFLOPS, 4),
                       "error": err_c},
        "fp64_ms_per_step": round(t64 * 1e3, 2),
        "ms_per_step_guard_off": round(ms_noguard, 4),
        "drop_in_call_ms_host_arrays": round(abi_ms, 2),
        "batch_64_wakes": batch,
        "far_wake_surrogate": surr_rep,
        "free_wake_march": march,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": TRAFFIC_BYTES,
            "traffic_unit": "bytes per launch",
            "kernel": "bs_fp32c<2,1>",
            "flop_per_interaction": FLOP_VATISTAS2,
        },
    }
    if cb is not None:
        line["cpu_baseline"] = cb
    line["skipped"] = budget.skipped
    line["run_s"] = round(time.monotonic() - T_START, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
