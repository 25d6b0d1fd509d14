"""Throughput of the online-reconstruction chain on one MI355X.

One step = one pass of the whole chain (reduction, XZ, YZ, back-partial
cleaning, station-1 extension, output records) over one batch of synthetic
events already resident in HBM.

Default workload (--workload c3), the learned-pruning configuration: hits per
event Poisson(1200) capped at 4000, 1-2 muon tracks, basic_geometry_e906, the
reference caps lifted (orchain.HIGH_OCCUPANCY_CAPS) so that the plain method
runs on every event.  Events are independent, so a step is one batch of
--events (default 40,000) events of that configuration; the rate in events/s
does not depend on the batch once it fills the GPU (40,000 events are about
40 per resident wave of the 2,048-hit tile class).

The same batch runs twice: the plain chain, then the chain with the shipped
learned YZ unit pruner (e1039-gpu_amd/data/pruner_u3v3.npz, per-tracklet keep
margin).  The value is the pruned chain's events/s; the plain rate, the
speedup and the track recall (plain-chain tracks the pruned chain outputs
identically) go beside it.  With --other (default) the other configurations
are timed as well, each on a batch of its own (config.other_workloads):
Poisson(300) (plain method, reference caps), the 0.8 Poisson(200) + 0.2
Poisson(1500) mixture (batched form: graph replay, autotuned; plain and
pruned), and Poisson(2000) with 1-4 tracks and the full chain (plain and
pruned, final-fit errors against float64).

Prints one JSON line with the throughput, the HBM roofline of the chain
kernels (algorithmic bytes per event: 12 per input hit, 1 bit per hit of keep
mask, 96 per output track) and the CPU oracle timed on a bounded sample.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "e1039-gpu_amd"))

METRIC = "events/s/GPU, learned-pruned vs plain GPU chain, at ≥99.5% plain-track recall"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 matrix-core peak
# name -> generator settings (include/e1039_or_gen.h), caps, default batch, description
WORKLOADS = {
    "c3": dict(gen=dict(mult_mode=1, mult_mean=1200.0, mult_cap=4000), caps="high", events=40_000,
               text="hits/event Poisson(1200) capped at 4000, 1-2 muon tracks (learned-pruning configuration)"),
    "c1": dict(gen=dict(mult_mode=1, mult_mean=300.0, mult_cap=1000), caps="reference", events=1_000_000,
               text="hits/event Poisson(300) capped at 1000, 1-2 muon tracks (plain-method configuration)"),
    "c2": dict(gen=dict(mult_mode=2, mult_mean=200.0, mult_mean2=1500.0, mult_frac2=0.2, mult_cap=4000), caps="high",
               events=100_000, text="hits/event 0.8 Poisson(200) + 0.2 Poisson(1500) capped at 4000 (batched form)"),
    "c4": dict(gen=dict(mult_mode=1, mult_mean=2000.0, mult_cap=6000, n_tracks_min=1, n_tracks_max=4), caps="high",
               events=3_000, text="hits/event Poisson(2000) capped at 6000, 1-4 muon tracks, full chain "
                                  "(high-intensity configuration)"),
    "occ8": dict(gen=dict(mult_mode=0, noise_occupancy=0.08), caps="reference", events=1_000_000,
                 text="8 % noise occupancy per channel (406 hits/event), 1-2 muon tracks"),
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


def lib_sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def kernel_sha():
    """sha256 (16 hex digits) over the chain kernel's sources, as rocprof_summary.py stamps its counter files"""
    h = hashlib.sha256()
    for n in ("or_chain.hip", "or_device.h", "or_prune.h"):
        with open(os.path.join(ROOT, "e1039-gpu_amd", "csrc", n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def params_for(oc, name):
    wl = WORKLOADS[name]
    if wl["caps"] == "reference":
        return oc.default_params(max_hits_per_event=1024)
    if name == "c4":
        return oc.high_occupancy_params(max_tracklets_per_event=4096, full_chain=1)
    return oc.high_occupancy_params()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c3")
    ap.add_argument("--events", type=int, default=0, help="0: the workload's default batch")
    ap.add_argument("--other", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the other configurations (plain-method, batched, high-intensity)")
    ap.add_argument("--seed", type=int, default=0xE1039)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the CPU baseline sample (0 = skip the baseline)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (16: the host-core share of one GPU on the GPU box)")
    ap.add_argument("--chain", choices=["both", "plain", "pruned"], default="both",
                    help="both: time plain and pruned; plain: plain only; pruned: plain once untimed (recall), pruned timed")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer (PCIe) leg")
    ap.add_argument("--waves-per-event", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="waves per event at most, each tile class up to its bound (0: the library default, 8)")
    ap.add_argument("--yz-batch", type=int, default=0, help="YZ survivor batch (0: the library default)")
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

    import fitcheck
    import orchain as oc
    import pruner

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    geo = oc.Geometry.load()
    flat, thr, pmeta = pruner.load() if os.path.exists(oc.DEFAULT_PRUNER) else (None, None, None)
    margin = pmeta["margin"] if pmeta else float("inf")

    def make(name, n):
        t0 = time.time()
        ev = oc.generate(geo, n, seed=args.seed, **WORKLOADS[name]["gen"])
        log(f"{name}: generated {ev.n_events} events, {ev.n_hits} hits ({ev.n_hits / max(1, ev.n_events):.1f}/event) "
            f"in {time.time() - t0:.1f}s")
        return ev

    def timed(sess, label, ev, dev_in, steps, warmup, fit=False, vertex=False):
        p = sess.params
        out = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True, fit=fit, vertex=vertex)
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
        ms = (time.perf_counter() - t0) * 1e3 / max(1, steps)
        km = float(np.mean(kms)) if kms else float("nan")
        log(f"{label}: timed {steps} steps: {ms:.2f} ms/step, chain kernels {km:.2f} ms")
        res = oc.Session.fetch(out, ev.n_events, p.max_tracks_out, p.max_tracklets_per_event)
        return ms, km, res

    def set_pruned(sess, on):
        if on:
            sess.set_pruner_margin(margin)
            sess.set_pruner(flat, thr)
        else:
            sess.set_pruner(None)

    # ---- the headline workload: plain, then learned-pruned ----------------------
    name = args.workload
    ev = make(name, args.events or WORKLOADS[name]["events"])
    params = params_for(oc, name)
    sess = oc.Session(geo, params, device=0)
    if args.waves_per_event:
        sess.set_tuning(oc.TUNE_WAVES_PER_EVENT, args.waves_per_event)
    if args.yz_batch:
        sess.set_tuning(oc.TUNE_YZ_LIST_CAP, args.yz_batch)
    wpe = sess.get_tuning(oc.TUNE_WAVES_PER_EVENT)
    yz_batch = sess.get_tuning(oc.TUNE_YZ_LIST_CAP)
    dev_in = sess.upload(ev)
    torch.cuda.synchronize()
    if args.chain == "pruned":
        plain_ms, plain_kms, plain_res = timed(sess, "plain (untimed reference)", ev, dev_in, 1, 0)
    else:
        plain_ms, plain_kms, plain_res = timed(sess, "plain", ev, dev_in, args.steps, args.warmup)
    prune = None
    if args.chain != "plain" and flat is not None:
        set_pruned(sess, True)
        pr_ms, pr_kms, pr_res = timed(sess, "pruned", ev, dev_in, args.steps, args.warmup)
        set_pruned(sess, False)
        rec, found, total = pruner.recall(plain_res, pr_res)
        extra = int(pr_res.n_tracks.sum()) - found
        log(f"pruned chain: track recall {rec:.5f} ({found}/{total} plain tracks identical), {extra} other tracks; "
            f"{plain_ms / pr_ms:.2f}x the plain chain")
        prune = {"ms": pr_ms, "kms": pr_kms, "res": pr_res, "recall": rec, "found": found, "total": total,
                 "extra": extra}
    ms_per_step = prune["ms"] if prune else plain_ms
    kernel_ms = prune["kms"] if prune else plain_kms
    res = prune["res"] if prune else plain_res
    n_out = int(res.n_tracks.sum())
    rejected = int(((plain_res.status & oc.EV_REJECT_MASK) != 0).sum())
    counts = np.diff(ev.offsets)
    b_alg = float(12 * counts.sum() + np.ceil(counts / 8.0).sum() + 96 * n_out)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    log(f"tracks out {n_out}, tracklets {int(res.n_tracklets.sum())}, rejected events {rejected}; "
        f"algorithmic bytes {b_alg / 1e9:.3f} GB per launch")

    # fits against float64 on a slice (untimed): the chain's own XZ fit and the final fit
    n_fit = min(ev.n_events, 5_000)
    fsub = ev.subset(0, n_fit)
    fres = sess.reconstruct(fsub, fit=True)
    fit_rep = fitcheck.error_report(fsub, fres, geo)
    xz_rep = fitcheck.xz_error_report(fsub, fres, geo)
    log("final fit vs float64 on {} tracks: {}; chain XZ fit vs float64: {}".format(
        fit_rep["tracks"], ", ".join(f"{q} {fit_rep[q]['max_rel']:.1e}" for q in fitcheck.QUANTITIES),
        ", ".join(f"{q} {xz_rep[q]['max_rel']:.1e}" for q in fitcheck.XZ_QUANTITIES)))

    # measured HBM copy bandwidth beside the vendor peak (1 GiB device-to-device, read + write)
    cbytes = 1 << 30
    csrc = torch.ones(cbytes, dtype=torch.uint8, device="cuda")
    cdst = torch.empty_like(csrc)
    cdst.copy_(csrc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        cdst.copy_(csrc)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2.0 * cbytes * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del csrc, cdst
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")

    # R-T calibration of every hit of the batch (or_rt_calibrate, device-resident):
    # 2 B det + 4 B tdc + 4 B drift read, 4 B drift written per hit
    sess.set_calibration(oc.Calibration.load())
    rt_out = sess.rt_calibrate_device(dev_in)
    torch.cuda.synchronize()
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

    # end to end through host buffers (PCIe both ways) on a slice, for reference only
    e2e = None
    if not args.no_e2e:
        n_e2e = min(ev.n_events, 10_000)
        sub = ev.subset(0, n_e2e)
        sess.reconstruct_host(sub)
        t0 = time.perf_counter()
        sess.reconstruct_host(sub)
        e2e = n_e2e / (time.perf_counter() - t0)
        log(f"host-buffer path (H2D + plain chain + D2H): {e2e:.0f} events/s on {n_e2e} events")
    del dev_in
    torch.cuda.empty_cache()

    # ---- the pruner's perceptron alone on the matrix cores (or_pruner_score) --------
    # 16-64-64-1: 2 x (16*64 + 64*64 + 64) = 10,368 FLOP per unit, bf16 operands, fp32 sums
    mlp = None
    if flat is not None:
        n_rows = 1 << 24
        frows = torch.randn((n_rows, 16), dtype=torch.float32, device="cuda")
        sess.set_pruner(flat, thr)
        sess.score_units(frows)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            sess.score_units(frows)
        e1.record()
        torch.cuda.synchronize()
        sess.set_pruner(None)
        mlp_ms = e0.elapsed_time(e1) / reps
        tflops = 10368.0 * n_rows / (mlp_ms * 1e-3) / 1e12
        mlp = {"units": n_rows, "ms": mlp_ms, "achieved": tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": tflops / BF16_PEAK_TFLOPS, "flop_per_unit": 10368,
               "note": "or_pruner_score over 16M random feature rows (64 B read, 4 B written per row)"}
        del frows
        log(f"pruner perceptron alone: {n_rows} units in {mlp_ms:.2f} ms, {tflops:.1f} TFLOP/s "
            f"({tflops / BF16_PEAK_TFLOPS * 100:.2f} % of the dense bf16 peak)")

    # ---- the other configurations -------------------------------------------------
    others = {}
    dimuons = None
    if args.other:
        for oname in [w for w in ("c1", "c2", "c4") if w != name]:
            oev = make(oname, WORKLOADS[oname]["events"])
            op = params_for(oc, oname)
            osess = oc.Session(geo, op, device=0)
            odev = osess.upload(oev)
            torch.cuda.synchronize()
            entry = {"workload": f"{oev.n_events} events, {WORKLOADS[oname]['text']}, "
                                 f"{oev.n_hits / oev.n_events:.0f} hits/event",
                     "caps": "reference" if WORKLOADS[oname]["caps"] == "reference" else "high-occupancy"}
            if oname == "c2":
                # batched form: autotuned on a slice, then the captured launch sequence replayed
                n_tune = min(oev.n_events, 20_000)
                tsub = oev.subset(0, n_tune)
                tt = osess.upload(tsub)
                to = osess.alloc_outputs(tsub.n_events, tsub.n_hits, tracklets=True)
                tune_ms = osess.autotune(tt, to, reps=1)
                entry["tuning"] = {"grid_percent": osess.get_tuning(oc.TUNE_GRID_PERCENT),
                                   "yz_batch": osess.get_tuning(oc.TUNE_YZ_LIST_CAP),
                                   "tuned_on": f"first {n_tune} events", "best_slice_ms": tune_ms}
                del tt, to
                osess.set_tuning(oc.TUNE_GRAPH, 1)
                entry["graph"] = True
            vertex = oname == "c4"
            o_ms, o_kms, o_plain = timed(osess, f"{oname} plain", oev, odev, 2, 1, vertex=vertex)
            entry["plain_events_per_s"] = oev.n_events / (o_ms * 1e-3)
            entry["plain_kernel_ms"] = o_kms
            entry["rejected_events"] = int(((o_plain.status & oc.EV_REJECT_MASK) != 0).sum())
            entry["tracks"] = int(o_plain.n_tracks.sum())
            if oname in ("c2", "c4") and flat is not None:
                # the learned-pruned chain on the same batch (for c2 also autotuned and replayed as a graph)
                set_pruned(osess, True)
                p_ms, _, o_pr = timed(osess, f"{oname} pruned", oev, odev, 2, 1, vertex=vertex)
                set_pruned(osess, False)
                orec, ofound, ototal = pruner.recall(o_plain, o_pr)
                entry.update({"pruned_events_per_s": oev.n_events / (p_ms * 1e-3), "speedup_vs_plain": o_ms / p_ms,
                              "track_recall": orec, "pruner": "trained on the Poisson(1200) configuration"})
            if oname == "c4" and flat is not None:
                osub = oev.subset(0, min(oev.n_events, 1000))
                ofr = osess.reconstruct(osub, fit=True)
                entry["final_fit_vs_fp64"] = fitcheck.error_report(osub, ofr, geo)
                entry["xz_fit_vs_fp64"] = fitcheck.xz_error_report(osub, ofr, geo)
                # dimuons and display histograms on the full-chain outputs
                dout = osess.alloc_outputs(oev.n_events, oev.n_hits, vertex=True)
                osess.run_device(odev, dout)
                dd = osess.alloc_dimuons(oev.n_events, histograms=True)
                osess.run_dimuons(odev, dout, dd)
                torch.cuda.synchronize()
                e0.record()
                for _ in range(reps):
                    osess.run_dimuons(odev, dout, dd)
                e1.record()
                torch.cuda.synchronize()
                dim_ms = e0.elapsed_time(e1) / reps
                dres = oc.Session.fetch_dimuons(dd, oev.n_events)
                dimuons = {"events": oev.n_events, "dimuons": int(dres.n_dimuons.sum()), "kernel_ms": dim_ms,
                           "events_per_s": oev.n_events / (dim_ms * 1e-3),
                           "sample": "the high-intensity batch after the full chain with vertexing, histograms on"}
                del dout, dd
            log(f"{oname}: " + json.dumps({k: v for k, v in entry.items() if not isinstance(v, dict)}))
            others[oname] = entry
            osess.close()
            del odev, oev
            torch.cuda.empty_cache()

    # ---- HBM traffic and VALU use from the committed rocprofv3 counters, when they
    # were taken on these kernel sources, settings and batch size ----
    traffic, traffic_note, valu = None, None, None
    lib_hash = lib_sha(oc.LIB_PATH)
    ksha = kernel_sha()
    mode = "pruned" if prune else "plain"
    tfile = os.path.join(ROOT, "profiles", f"chain_kernel_{name}_{mode}_counters.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if (tj.get("kernel_source_sha256_16") == ksha and tj.get("events") == ev.n_events
                and tj.get("waves_per_event") == wpe and tj.get("yz_batch") == yz_batch):
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_note = (f"{os.path.basename(tfile)}: FETCH_SIZE + WRITE_SIZE of this launch, copy-calibrated; "
                            f"{tj['launch_ms']:.1f} ms under the profiler")
        else:
            traffic_note = f"{os.path.basename(tfile)} was measured on other kernel sources, settings or batch; not used"
        pfile = os.path.join(ROOT, "profiles", f"chain_kernel_{name}_plain_counters.json")
        if os.path.exists(pfile):
            with open(pfile) as f:
                pj = json.load(f)
            if pj.get("kernel_source_sha256_16") == ksha and pj.get("derived"):
                valu = {"busy_frac": pj["derived"]["valu_busy_frac"],
                        "valu_insts_per_event": pj["derived"]["valu_insts_per_event"],
                        "lds_util_frac": pj["derived"]["lds_util_frac"],
                        "lds_bank_conflict_frac": pj["derived"]["lds_bank_conflict_frac"],
                        "source": os.path.basename(pfile) + " (plain chain, SQ counters)"}
    else:
        traffic_note = "no counter file for this workload"

    # ---- CPU baseline: the oracle on the host cores, bounded sample ----------------
    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        og = orc.OracleGeometry(oc.DEFAULT_GEOMETRY)
        threads = args.cpu_threads
        probe = ev.subset(0, min(ev.n_events, threads))
        t0 = time.perf_counter()
        orc.reconstruct(og, params, probe, threads=threads)
        rate = probe.n_events / max(1e-6, time.perf_counter() - t0)
        ns = int(min(ev.n_events, max(probe.n_events, rate * args.cpu_seconds)))
        sub = ev.subset(0, ns)
        t0 = time.perf_counter()
        orc.reconstruct(og, params, sub, threads=threads)
        dt = time.perf_counter() - t0
        n1 = int(min(ns, max(2, rate / threads * args.cpu_seconds * 0.2)))
        t0 = time.perf_counter()
        orc.reconstruct(og, params, ev.subset(0, n1), threads=1)
        r1 = n1 / (time.perf_counter() - t0)
        cpu = {"value": ns / dt, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} events of the same batch; oracle (C++ restatement of the reference "
                         f"per-event routines, g++ -O2, OpenMP over events, same caps) took {dt:.1f}s",
               "value_1_thread": r1, "host_cpu": cpu_model(), "host_nproc": os.cpu_count()}
        log(f"cpu baseline: {ns / dt:.2f} events/s on {threads} threads ({ns} events), {r1:.3f} events/s on 1")

    wl = WORKLOADS[name]
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
            "workload": f"{ev.n_events}-event batches, {wl['text']}, {ev.n_hits / ev.n_events:.0f} hits/event, "
                        f"basic_geometry_e906, "
                        + ("reference caps" if wl["caps"] == "reference" else
                           "caps lifted: " + ", ".join(f"{k}={v}" for k, v in params.as_dict().items()
                                                        if k.startswith("max_") and k != "max_tracks_out")),
            "chain": ("learned-pruned (YZ station-3 units, bf16 perceptron on the matrix cores, keep margin "
                      f"{margin:.3f} below each tracklet's best unit)") if prune else "plain",
            "events": ev.n_events,
            "hits": int(ev.n_hits),
            "waves_per_event": wpe,
            "waves_per_event_note": "at most: the 1,024- and 2,048-hit tiles run min(value, 4), the 4,096- and "
                                    "8,192-hit tiles min(value, 8)",
            "yz_batch": yz_batch,
            "kernel_ms": kernel_ms,
            "rejected_events": rejected,
            "rejected_fraction": rejected / ev.n_events,
            "plain_events_per_s": ev.n_events / (plain_ms * 1e-3) if args.chain != "pruned" else None,
            "plain_ms_per_step": plain_ms if args.chain != "pruned" else None,
            "speedup_vs_plain": (plain_ms / prune["ms"]) if prune and args.chain == "both" else None,
            "track_recall": prune["recall"] if prune else 1.0,
            "plain_tracks": prune["total"] if prune else int(plain_res.n_tracks.sum()),
            "pruned_tracks_identical": prune["found"] if prune else None,
            "pruned_other_tracks": prune["extra"] if prune else None,
            "prune_margin": margin if prune else None,
            "host_buffer_events_per_s": e2e,
            "other_workloads": others or None,
            "final_fit_vs_fp64": dict(fit_rep, sample=f"output tracks of the first {n_fit} events, plain chain"),
            "xz_fit_vs_fp64": dict(xz_rep, sample=f"output tracks of the first {n_fit} events, plain chain"),
            "dimuons": dimuons,
            "rt_calibrate": {"hits": int(ev.n_hits), "ms": rt_ms, "algorithmic_GB_per_s": rt_gbs,
                             "frac_of_hbm_peak": rt_gbs / HBM_PEAK_GBS, "bytes_per_hit": 14},
            "library_sha256_16": lib_hash,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "measured_copy_GBs": copy_gbs, "valu": valu, "pruner_mfma": mlp},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
