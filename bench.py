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

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "winds-master_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Biot–Savart G interactions/s & free-wake step ms vs FP64 plain, at stated err"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak
FP64_PEAK_TFLOPS = 78.6
FLOP_VATISTAS2 = 59        # algorithmic FLOPs per interaction, Vatistas n=2 form
FLOP_CUTOFF = 47           # cut-off (MEX) form


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(wake, budget_s: float = 12.0) -> dict:
    """The oracle's BiotSavart_nogpu restatement on a bounded target sample."""
    import numpy as np
    from oracle import bs as O
    from winds import layout

    recs = layout.wake_segments(wake)
    F1 = np.asfortranarray(recs[:, 0:3].reshape(-1, 3, 1, 1))
    F2 = np.asfortranarray(recs[:, 3:6].reshape(-1, 3, 1, 1))
    g = recs[:, 6].copy()
    P = wake.domain
    nseg = recs.shape[0]
    nt = wake.n_targets
    threads = host_cores()
    out = {}
    for label, th, frac in (("all", threads, 0.75), ("one", 1, 0.25)):
        n = 16 * th if th > 1 else 16
        t0 = time.perf_counter()
        O.bs_mex_range(F1, F2, P, g, 1e-4, 0, n, threads=th)
        dt = time.perf_counter() - t0
        n2 = int(min(nt, max(n, n * budget_s * frac / max(dt, 1e-3))))
        t0 = time.perf_counter()
        O.bs_mex_range(F1, F2, P, g, 1e-4, 0, n2, threads=th)
        dt = time.perf_counter() - t0
        out[label] = dict(rate=n2 * nseg / dt / 1e9, targets=n2, seconds=dt, threads=th)
    a = out["all"]
    return {
        "value": round(a["rate"], 4),
        "unit": "G interactions/s",
        "cores": a["threads"],
        "kind": "port",
        "sample": (f"first {a['targets']} of {nt} targets x all {nseg} segments, cut-off core "
                   f"(BiotSavart_nogpu arithmetic, FP64), {a['seconds']:.1f} s on "
                   f"{a['threads']} OpenMP threads; 1 thread: "
                   f"{out['one']['rate']:.4f} G int/s on {out['one']['targets']} targets"),
        "value_1thread": round(out["one"]["rate"], 4),
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
    # error of four wakes against the FP64 engine
    ev = Evaluator(0)
    parts = b.split()
    worst = {"max_rel": 0.0, "rms_rel": 0.0, "normwise": 0.0}
    for i in (0, 17, 40, 63):
        w = wakes[i]
        tg = torch.from_numpy(layout.wake_targets(w)).cuda().contiguous()
        sg = torch.from_numpy(layout.wake_segments(w)).cuda().contiguous()
        ref = ev.eval(tg, sg, o64).cpu().numpy()
        u = parts[i].cpu().numpy()
        nrm = np.linalg.norm(ref, axis=0)
        dif = np.linalg.norm(u - ref, axis=0)
        m = nrm >= 1e-6 * nrm.max()
        e = dif[m] / nrm[m]
        worst["max_rel"] = max(worst["max_rel"], float(e.max()))
        worst["rms_rel"] = max(worst["rms_rel"], float(np.sqrt(np.mean(e * e))))
        worst["normwise"] = max(worst["normwise"], float(dif.max() / nrm.max()))
    log(f"batch error (4 wakes vs FP64): {worst}")
    return {"value": round(rate, 2), "unit": "G interactions/s", "ms_per_batch": round(ms, 3),
            "wakes": len(wakes), "targets": info["targets"], "segments": info["segments"],
            "interactions": info["interactions"], "work_items": info["items"],
            "tune": tune, "graph": True, "error_4_wakes_vs_fp64": worst}


def march_opts(prec):
    import winds
    # WInDS defaults: cut-off core d = 1e-4 (WInDS.m:42) through the MEX path
    return winds.opts(core=winds.CORE_CUTOFF, semantics=winds.SEM_MEX, precision=prec, d=1e-4)


def march_bench(revs: int) -> dict:
    """Free-wake time march of a 3-blade, 31-node rotor at 5 deg steps with
    PCC (2 wake evaluations per step) and an FP64 Kutta-Joukowski solve:
    (a) step time with the wake at its 1,440-age truncation length;
    (b) thrust/power over `revs` revolutions from a cold start, FP32
        compensated wake evaluations vs FP64."""
    import numpy as np
    import torch

    import winds
    from winds import gen, rotor as rmod
    from winds.march import March

    rot = rmod.synthetic_rotor(nb=3, nst=31)
    out = {"rotor": "3 blades x 31 nodes, R 63 m, TSR 7, U 11.4 m/s, 5 deg steps, PCC",
           "max_ages": 1440}
    # (a) steady length: start from a 1,440-age helical wake
    w = gen.rotor_wake(nb=3, nst=31, ages=1440, jitter=0.0, R=rot.R, hub=rot.hub, tsr=rot.tsr,
                       U=rot.U, dpsi_deg=rot.dpsi_deg)
    for name, prec, nsteps in (("fp32c", winds.PREC_FP32C, 10), ("fp64", winds.PREC_FP64, 2)):
        m = March(rot, 1440, march_opts(prec), integrator="pcc")
        m.set_state(1441, w.domain, w.gamma_shed[:, :, :-1], w.gamma_trail[:, :, :-1])
        m.step(1)
        m.sync()
        t0 = time.perf_counter()
        m.step(nsteps)
        m.sync()
        ms = (time.perf_counter() - t0) / nsteps * 1e3
        out[f"step_ms_{name}"] = round(ms, 2)
        log(f"march step at 1440 ages, {name} wake sums: {ms:.1f} ms")
        del m
    # (b) cold start, `revs` revolutions
    steps = int(round(revs * 360 / rot.dpsi_deg))
    hist = {}
    for name, prec in (("fp32c", winds.PREC_FP32C), ("fp64", winds.PREC_FP64)):
        t0 = time.perf_counter()
        m = March(rot, 1440, march_opts(prec), integrator="pcc", max_steps=steps + 8)
        m.step(steps)
        m.sync()
        hist[name] = m.perf()
        log(f"march {revs} revs ({steps} steps) {name}: {time.perf_counter() - t0:.1f} s")
        del m
    a, b = hist["fp32c"], hist["fp64"]
    last = slice(-int(360 / rot.dpsi_deg), None)   # final revolution
    rel = lambda i: float(np.max(np.abs(a[last, i] - b[last, i]) / np.abs(b[last, i])))
    out.update({"revs": revs, "steps": steps,
                "thrust_N_last_rev_fp64": float(np.mean(b[last, 1])),
                "power_W_last_rev_fp64": float(np.mean(b[last, 3])),
                "thrust_max_rel_err_fp32c_vs_fp64_last_rev": rel(1),
                "power_max_rel_err_fp32c_vs_fp64_last_rev": rel(3),
                "kj_iterations_mean": float(np.mean(a[:, 4]))})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-wake batch")
    ap.add_argument("--no-march", action="store_true", help="skip the free-wake time march")
    ap.add_argument("--march-revs", type=int, default=10)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 2 GiB device-to-device copy (reference point for "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--ages", type=int, default=1440)
    args = ap.parse_args()
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
    recs = layout.wake_segments(wake)
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

    # FP64 engine result (reference for the error), checked on a sample
    t0 = time.perf_counter()
    u64 = ev.eval(tg, sg, o64).clone()
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
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
    log(f"fp32c error vs fp64: {err}")

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

    if args.calib_copy:
        src = torch.ones(2 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    batch = None if args.no_batch else batch_bench(o32, o64, args.steps)
    march = None if args.no_march else march_bench(args.march_revs)

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
                         f"evaluation, Vatistas n=2 core, FP64 for pairs within max(40 r_c, |L|/2)"),
            "targets": nt, "segments": nseg, "interactions_per_step": inter,
            "core": "vatistas2", "precision": "fp32 compensated + fp64 near pairs",
        },
        "error": err,
        "cutoff_mex": {"value": round(rate_c, 2), "unit": "G interactions/s",
                       "ms_per_step": round(ms_c, 4),
                       "tflops": round(inter * FLOP_CUTOFF / (ms_c * 1e-3) / 1e12, 2)},
        "fp64_ms_per_step": round(t64 * 1e3, 2),
        "batch_64_wakes": batch,
        "free_wake_march": march,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": None,
            "kernel": "bs_fp32c<2,1>",
            "flop_per_interaction": FLOP_VATISTAS2,
        },
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(wake)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
