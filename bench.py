"""Throughput of the online-reconstruction chain on one MI355X.

One step = one pass of the whole chain (reduction, XZ, YZ, back-partial
cleaning, station-1 extension, output records) over one batch of synthetic
events already resident in HBM.  Default workload: the learned-pruning
configuration, 4,000,000 events with 8 % noise occupancy per channel (about
406 hits per event), 1-2 muon tracks, basic_geometry_e906 (--workload occ8);
--workload poisson300 runs 1,000,000 events with Poisson(300) hits capped at
1000 instead.

The same batch runs twice: the plain chain, then the chain with the shipped
learned YZ unit pruner (e1039-gpu_amd/data/pruner_u3v3.npz).  The value is the
pruned chain's events/s; the plain rate, the speedup and the pruned chain's
track recall (plain-chain tracks it outputs identically) go beside it, and the
same three figures for the other workload (1M events each) go in
config.other_workload.

Prints one JSON line with the throughput, the HBM roofline of the chain
kernel (algorithmic bytes per event: 12 per input hit, 1 bit per hit of keep
mask, 96 per output track) and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "e1039-gpu_amd"))

METRIC = "events/s/GPU, learned-pruned vs plain GPU chain, at ≥99.5% plain-track recall"
HBM_PEAK_GBS = 8000.0
# generator settings (include/e1039_or_gen.h) and default event counts
WORKLOADS = {
    "occ8": (dict(mult_mode=0, noise_occupancy=0.08), 4_000_000,
             "8 % noise occupancy per channel (learned-pruning configuration)"),
    "poisson300": (dict(mult_mode=1, mult_mean=300.0, mult_cap=1000), 1_000_000,
                   "hits/event Poisson(300) capped at 1000"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="occ8")
    ap.add_argument("--events", type=int, default=0, help="0: the workload's default")
    ap.add_argument("--other", action=argparse.BooleanOptionalAction, default=True,
                    help="also time plain and pruned on the other workload (1M events)")
    ap.add_argument("--seed", type=int, default=0xE1039)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the CPU baseline sample (0 = skip the baseline)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--chain", choices=["both", "plain", "pruned"], default="both",
                    help="both: time plain and pruned; plain: plain only; pruned: plain once untimed (recall), pruned timed")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer (PCIe) leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (reference for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} bytes")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); events are independent, so N GPUs would "
              "each run this same per-GPU workload on their own event range", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import orchain as oc

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    geo = oc.Geometry.load()
    t0 = time.time()
    gen, n_default, wl_text = WORKLOADS[args.workload]
    ev = oc.generate(geo, args.events or n_default, seed=args.seed, **gen)
    log(f"generated {ev.n_events} events, {ev.n_hits} hits ({ev.n_hits / max(1, ev.n_events):.1f}/event) "
        f"in {time.time() - t0:.1f}s")
    params = oc.default_params(max_hits_per_event=1024)
    sess = oc.Session(geo, params, device=0)
    dev_in = sess.upload(ev)
    torch.cuda.synchronize()
    # the autotuner (YZ candidate-list size x persistent grid size) on a slice of the batch;
    # results never depend on these settings
    n_tune = min(ev.n_events, 200_000)
    tsub = ev.subset(0, n_tune)
    tt = sess.upload(tsub)
    to = sess.alloc_outputs(tsub.n_events, tsub.n_hits, tracklets=True)
    tune_ms = sess.autotune(tt, to, reps=2)
    tuning = {"grid_percent": sess.get_tuning(oc.TUNE_GRID_PERCENT),
              "yz_list_cap": sess.get_tuning(oc.TUNE_YZ_LIST_CAP),
              "tuned_on": f"first {n_tune} events, or_session_autotune", "best_slice_ms": tune_ms}
    del tt, to
    log(f"autotune: grid {tuning['grid_percent']} %, YZ list cap {tuning['yz_list_cap']} ({tune_ms:.2f} ms on the slice)")

    def timed(label, steps=None, warmup=None, ev=ev, dev_in=dev_in):
        steps = args.steps if steps is None else steps
        warmup = args.warmup if warmup is None else warmup
        out = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
        for i in range(warmup):
            sess.run_device(dev_in, out)
            torch.cuda.synchronize()
            log(f"{label} warmup step {i} done: kernel {sess.last_kernel_ms():.1f} ms")
        kms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            sess.run_device(dev_in, out)
            kms.append(sess.last_kernel_ms())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / max(1, steps)
        km = float(np.mean(kms)) if kms else float("nan")
        log(f"{label}: timed {steps} steps: {ms:.2f} ms/step, chain kernel {km:.2f} ms")
        res = oc.Session.fetch(out, ev.n_events, params.max_tracks_out, params.max_tracklets_per_event)
        return ms, km, res

    if args.chain == "pruned":
        plain_ms, plain_kms, plain_res = timed("plain (untimed reference)", steps=1, warmup=0)
    else:
        plain_ms, plain_kms, plain_res = timed("plain")
    prune = None
    if args.chain != "plain":
        import pruner
        if not os.path.exists(oc.DEFAULT_PRUNER):
            log(f"no pruner weights at {oc.DEFAULT_PRUNER}; reporting the plain chain")
        else:
            flat, thr, meta = pruner.load()
            sess.set_pruner(flat, thr)
            pr_ms, pr_kms, pr_res = timed("pruned")
            sess.set_pruner(None)
            rec, found, total = pruner.recall(plain_res, pr_res)
            extra = int(pr_res.n_tracks.sum()) - found
            log(f"pruned chain: track recall {rec:.5f} ({found}/{total} plain tracks identical), {extra} other tracks; "
                f"{plain_ms / pr_ms:.2f}x the plain chain")
            prune = {"ms": pr_ms, "kms": pr_kms, "res": pr_res, "recall": rec, "found": found, "total": total,
                     "extra": extra, "threshold": thr}
    other = None
    if args.other and prune:
        # the other workload, 1M events: plain and pruned rates and the recall
        oname = [w for w in WORKLOADS if w != args.workload][0]
        ogen, _, otext = WORKLOADS[oname]
        oev = oc.generate(geo, 1_000_000, seed=args.seed, **ogen)
        odev = sess.upload(oev)
        torch.cuda.synchronize()
        o_plain_ms, _, o_plain = timed(f"{oname} plain", steps=3, warmup=1, ev=oev, dev_in=odev)
        sess.set_pruner(flat, thr)
        o_pr_ms, _, o_pr = timed(f"{oname} pruned", steps=3, warmup=1, ev=oev, dev_in=odev)
        sess.set_pruner(None)
        orec, ofound, ototal = pruner.recall(o_plain, o_pr)
        other = {"workload": f"{oev.n_events} events, {otext}, {oev.n_hits / oev.n_events:.0f} hits/event",
                 "plain_events_per_s": oev.n_events / (o_plain_ms * 1e-3),
                 "pruned_events_per_s": oev.n_events / (o_pr_ms * 1e-3), "speedup_vs_plain": o_plain_ms / o_pr_ms,
                 "track_recall": orec, "plain_tracks": ototal, "pruned_tracks_identical": ofound}
        log(f"{oname}: plain {other['plain_events_per_s']:.4g} events/s, pruned {other['pruned_events_per_s']:.4g} "
            f"({other['speedup_vs_plain']:.2f}x), recall {orec:.5f}")
        del odev, oev
    ms_per_step = prune["ms"] if prune else plain_ms
    kernel_ms = prune["kms"] if prune else plain_kms
    res = prune["res"] if prune else plain_res
    n_out = int(res.n_tracks.sum())
    rejected = int(((res.status & oc.EV_REJECT_MASK) != 0).sum())
    counts = np.diff(ev.offsets)
    b_alg = float(12 * counts.sum() + np.ceil(counts / 8.0).sum() + 96 * n_out)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    log(f"tracks out {n_out}, tracklets {int(res.n_tracklets.sum())}, rejected events {rejected}; "
        f"algorithmic bytes {b_alg / 1e9:.3f} GB per launch")

    # final straight-line fit of the output tracks against float64, on a slice (untimed)
    import fitcheck
    n_fit = min(ev.n_events, 200_000)
    fsub = ev.subset(0, n_fit)
    fit_rep = fitcheck.error_report(fsub, sess.reconstruct(fsub, fit=True), geo)
    log(f"final fit vs float64 on {fit_rep['tracks']} tracks: " +
        ", ".join(f"{q} max {fit_rep[q]['max_rel']:.1e}" for q in fitcheck.QUANTITIES) +
        f"; FP64 solves {fit_rep['fp64_fraction'] * 100:.2f} %")

    # R-T calibration of every hit of the batch (or_rt_calibrate, device-resident):
    # 2 B det + 4 B tdc + 4 B drift read, 4 B drift written per hit
    sess.set_calibration(oc.Calibration.load())
    rt_out = sess.rt_calibrate_device(dev_in)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        rt_out = sess.rt_calibrate_device(dev_in)
    e1.record()
    torch.cuda.synchronize()
    sess.set_calibration(None)
    rt_ms = e0.elapsed_time(e1) / reps
    rt_gbs = 14.0 * ev.n_hits / (rt_ms * 1e-3) / 1e9
    del rt_out
    log(f"R-T calibration of {ev.n_hits} hits: {rt_ms:.3f} ms, {rt_gbs:.0f} GB/s algorithmic "
        f"({rt_gbs / HBM_PEAK_GBS * 100:.1f} % of HBM peak)")

    # dimuons and display histograms (or_dimuons) after the full chain with vertexing, on a slice
    n_dim_ev = min(ev.n_events, 200_000)
    dsub = ev.subset(0, n_dim_ev)
    fparams = oc.default_params(max_hits_per_event=1024)
    fparams.full_chain = 1
    fsess = oc.Session(geo, fparams, device=0)
    ft = fsess.upload(dsub)
    fo = fsess.alloc_outputs(dsub.n_events, dsub.n_hits, vertex=True)
    fsess.run_device(ft, fo)
    dd = fsess.alloc_dimuons(dsub.n_events, histograms=True)
    fsess.run_dimuons(ft, fo, dd)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        fsess.run_dimuons(ft, fo, dd)
    e1.record()
    torch.cuda.synchronize()
    dim_ms = e0.elapsed_time(e1) / reps
    dres = oc.Session.fetch_dimuons(dd, dsub.n_events)
    n_dimuons = int(dres.n_dimuons.sum())
    log(f"dimuons on {n_dim_ev} full-chain events: {n_dimuons} dimuons, {dim_ms:.3f} ms "
        f"({n_dim_ev / (dim_ms * 1e-3):.3g} events/s)")
    del fsess, ft, fo, dd

    # end to end through host buffers (PCIe both ways) on a slice, for reference only
    e2e = None
    if not args.no_e2e:
        n_e2e = min(ev.n_events, 200_000)
        sub = ev.subset(0, n_e2e)
        sess.reconstruct_host(sub)
        t0 = time.perf_counter()
        sess.reconstruct_host(sub)
        e2e = n_e2e / (time.perf_counter() - t0)
        log(f"host-buffer path (H2D + plain chain + D2H): {e2e:.0f} events/s on {n_e2e} events")

    traffic = None
    tlabel = f"{ev.n_events // 1_000_000}M_{args.workload}"
    tfile = os.path.join(ROOT, "profiles", f"chain_kernel_{tlabel}_pruned_traffic.json" if prune else
                         f"chain_kernel_{tlabel}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        og = orc.OracleGeometry(oc.DEFAULT_GEOMETRY)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        # size the sample from a short probe so the timed run takes about cpu_seconds
        probe = ev.subset(0, min(ev.n_events, 20_000))
        t0 = time.perf_counter()
        orc.reconstruct(og, params, probe, threads=threads)
        rate = probe.n_events / max(1e-6, time.perf_counter() - t0)
        ns = int(min(ev.n_events, max(probe.n_events, rate * args.cpu_seconds)))
        sub = ev.subset(0, ns)
        t0 = time.perf_counter()
        orc.reconstruct(og, params, sub, threads=threads)
        dt = time.perf_counter() - t0
        # single-thread rate on a short prefix (about a tenth of the time budget)
        n1 = int(min(ns, max(500, rate / threads * args.cpu_seconds * 0.1)))
        t0 = time.perf_counter()
        orc.reconstruct(og, params, ev.subset(0, n1), threads=1)
        r1 = n1 / (time.perf_counter() - t0)
        cpu = {"value": ns / dt, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} events of the same workload; oracle (C++ restatement of the reference "
                         f"per-event routines, g++ -O2, OpenMP over events) took {dt:.1f}s",
               "value_1_thread": r1, "host_cpu": cpu_model()}
        log(f"cpu baseline: {ns / dt:.0f} events/s on {threads} threads ({ns} events), {r1:.0f} events/s on 1")

    line = {
        "metric": METRIC,
        "value": ev.n_events / (ms_per_step * 1e-3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"{ev.n_events} events, {wl_text}, {ev.n_hits / ev.n_events:.0f} hits/event, "
                        f"1-2 muon tracks, basic_geometry_e906, reference caps",
            "chain": "learned-pruned (YZ station-3 units, bf16 perceptron on the matrix cores)" if prune else "plain",
            "events": ev.n_events,
            "hits": int(ev.n_hits),
            "kernel_ms": kernel_ms,
            "plain_events_per_s": ev.n_events / (plain_ms * 1e-3) if args.chain != "pruned" else None,
            "plain_ms_per_step": plain_ms if args.chain != "pruned" else None,
            "speedup_vs_plain": (plain_ms / prune["ms"]) if prune and args.chain == "both" else None,
            "track_recall": prune["recall"] if prune else 1.0,
            "plain_tracks": prune["total"] if prune else int(plain_res.n_tracks.sum()),
            "pruned_tracks_identical": prune["found"] if prune else None,
            "pruned_other_tracks": prune["extra"] if prune else None,
            "prune_threshold": prune["threshold"] if prune else None,
            "tuning": tuning,
            "host_buffer_events_per_s": e2e,
            "other_workload": other,
            "final_fit_vs_fp64": dict(fit_rep, sample=f"output tracks of the first {n_fit} events, plain chain"),
            "dimuons": {"events": n_dim_ev, "dimuons": n_dimuons, "kernel_ms": dim_ms,
                        "events_per_s": n_dim_ev / (dim_ms * 1e-3),
                        "sample": "first events of the batch, full chain with vertexing, histograms on"},
            "rt_calibrate": {"hits": int(ev.n_hits), "ms": rt_ms, "algorithmic_GB_per_s": rt_gbs,
                             "frac_of_hbm_peak": rt_gbs / HBM_PEAK_GBS, "bytes_per_hit": 14},
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
