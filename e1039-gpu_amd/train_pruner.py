"""Trains the YZ unit pruner on the plain chain's own decisions (needs the GPU).

    python e1039-gpu_amd/train_pruner.py [--train 20000] [--out e1039-gpu_amd/data/pruner_u3v3.npz]
    python e1039-gpu_amd/train_pruner.py --refine-margin 1.9 2.125 0.025   (weights kept, margin re-picked)

Workload: hits per event Poisson(1200) capped at 4000 with the high-occupancy
caps (orchain.HIGH_OCCUPANCY_CAPS), the learned-pruning configuration.

1. Runs the plain chain on seeded synthetic events with the unit dump on
   (or_session_set_unit_dump): one row of 16 features per YZ unit, with a
   flag saying whether the tracklet's YZ winner lies in that unit.  Rows of a
   tracklet are contiguous.
2. Fits the perceptron as a ranking of each tracklet's units with its YZ
   winner as the target (pruner.train_rank).  Every tracklet's winner counts,
   not only those of tracklets that end as tracks: back-partial cleaning
   compares each surviving tracklet with the others, so a changed winner of
   any tracklet can change which tracks come out.
3. The device keeps a unit when its score is within a margin of the best
   score of its tracklet (or_session_set_pruner_margin).  Candidate margins
   are validation quantiles of (best score - winner score); each is run
   through the pruned chain on a separate event set, and the smallest whose
   track recall (tracks output identically, pruner.recall) reaches --target
   is kept.
4. Saves weights, threshold (off), margin and the measured figures.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

WORKLOAD = dict(mult_mode=1, mult_mean=1200.0, mult_cap=4000)


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20_000, help="training events")
    ap.add_argument("--valid", type=int, default=2_000, help="validation events (margin candidates)")
    ap.add_argument("--test", type=int, default=8_000, help="events for the track-recall check")
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--shard", type=int, default=1000, help="events per unit dump")
    ap.add_argument("--sample", type=int, default=1,
                    help="keep the units of one tracklet in SAMPLE (a fixed hash of event and tracklet), so that "
                         "more events fit the same number of rows")
    ap.add_argument("--train-only", action="store_true",
                    help="save the trained weights without picking a margin (then --refine-margin)")
    ap.add_argument("--target", type=float, default=0.996, help="track recall the margin must reach")
    ap.add_argument("--out", default=os.path.join(HERE, "data", "pruner_u3v3.npz"))
    ap.add_argument("--refine-margin", nargs=3, type=float, metavar=("LO", "HI", "STEP"),
                    help="keep the weights of --out and re-pick its margin on the test events from the grid "
                         "LO, LO + STEP, ... <= HI (the smallest reaching --target)")
    args = ap.parse_args()

    import numpy as np
    import torch

    import orchain as oc
    import pruner

    geo = oc.Geometry.load()
    params = oc.high_occupancy_params()
    sess = oc.Session(geo, params, device=0)
    dev = torch.device("cuda", 0)

    def dump(n_events, seed, sample=1):
        """plain chain with unit rows on n_events generated events, in shards (shard k
        generated with seed 2^20 seed + k, apart from the test events' seed); with
        sample > 1 only the units of one tracklet in sample are kept; returns
        (features, group id, winner) on the device"""
        feats, groups, wins = [], [], []
        gbase = 0
        for k, s0 in enumerate(range(0, n_events, args.shard)):
            sub = oc.generate(geo, min(n_events, s0 + args.shard) - s0, seed=(seed << 20) + k, **WORKLOAD)
            cap = sub.n_events * 200_000
            rows = torch.empty((cap, oc.UNIT_ROW), dtype=torch.float32, device=dev)
            count = torch.zeros(1, dtype=torch.int64, device=dev)
            sess.set_pruner(None)
            sess.set_unit_dump(rows, count)
            t = sess.upload(sub)
            o = sess.alloc_outputs(sub.n_events, sub.n_hits, tracklets=True)
            sess.run_device(t, o)
            torch.cuda.synchronize()
            sess.set_unit_dump(None)
            n = int(count.item())
            if n > cap:
                raise RuntimeError(f"unit rows: {n} wanted, {cap} held")
            r = rows[:n]
            r = r[r[:, 3] >= 0]          # chunks that went unit by unit carry no winner flags
            key = r[:, 0].long() * (1 << 20) + r[:, 1].long()
            if sample > 1:
                h = ((key + (s0 << 20)) * 2654435761) % 4294967296
                keep = (h >> 16) % sample == 0
                r, key = r[keep], key[keep]
            chg = torch.cat([torch.ones(1, dtype=torch.bool, device=dev), key[1:] != key[:-1]])
            g = torch.cumsum(chg.long(), 0) - 1 + gbase
            gbase = int(g[-1]) + 1 if g.numel() else gbase
            feats.append(r[:, 4:4 + pruner.NF].clone())
            groups.append(g)
            wins.append(r[:, 3] > 0.5)
            del rows, r, o, t
            if k % 50 == 49:
                log(f"  dumped {s0 + sub.n_events} of {n_events} events, {sum(f.shape[0] for f in feats)} units")
        return torch.cat(feats), torch.cat(groups), torch.cat(wins)

    def run(ev, flat=None, margin=float("inf")):
        sess.set_pruner_margin(margin)
        sess.set_pruner(None) if flat is None else sess.set_pruner(flat, pruner.THRESHOLD_OFF)
        t = sess.upload(ev)
        o = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
        sess.run_device(t, o)
        torch.cuda.synchronize()
        res = oc.Session.fetch(o, ev.n_events, params.max_tracks_out, params.max_tracklets_per_event)
        sess.run_device(t, o)
        torch.cuda.synchronize()
        return res, sess.last_kernel_ms()

    if args.refine_margin:
        flat, _, meta = pruner.load(args.out)
        te = oc.generate(geo, args.test, seed=0x7A33, **WORKLOAD)
        plain, plain_ms = run(te)
        lo, hi, step = args.refine_margin
        chosen, trials = None, []
        for k in range(int(round((hi - lo) / step)) + 1):
            mg = round(lo + k * step, 6)
            pr, ms = run(te, flat, mg)
            rec, found, total = pruner.recall(plain, pr)
            entry = {"margin": mg, "track_recall": rec, "tracks": total, "tracks_identical": found,
                     "plain_ms": plain_ms, "pruned_ms": ms, "speedup": plain_ms / ms}
            trials.append(entry)
            log(json.dumps(entry))
            if chosen is None and rec >= args.target:
                chosen = entry
        sess.set_pruner(None)
        if chosen is None:
            raise RuntimeError("no margin of the grid reaches the target recall")
        meta.update({"margin": chosen["margin"], "chosen": chosen,
                     "margin_refinement": {"grid": [lo, hi, step], "target": args.target, "trials": trials}})
        pruner.save(args.out, flat, pruner.THRESHOLD_OFF, meta)
        log(json.dumps({"saved": args.out, **chosen}))
        return

    t0 = time.time()
    x, g, w = dump(args.train, 0x7A31, args.sample)
    log(f"training: {args.train} events (one tracklet in {args.sample}), {x.shape[0]} units, "
        f"{int(g[-1]) + 1} tracklets, {int(w.sum())} winners; dumped in {time.time() - t0:.1f}s")
    t0 = time.time()
    flat = pruner.train_rank(x, g, w, epochs=args.epochs, log=log)
    log(f"trained in {time.time() - t0:.1f}s")
    del x, g, w
    torch.cuda.empty_cache()
    if args.train_only:
        meta = {"features": pruner.FEATURE_NAMES, "train_events": args.train, "train_sample": args.sample,
                "workload": WORKLOAD, "caps": oc.HIGH_OCCUPANCY_CAPS, "seeds": "train 2^20 0x7A31 + shard, test 0x7A33",
                "loss": "listwise (softmax over a tracklet's units)", "epochs": [args.epochs, 4],
                "margin": float("inf")}
        pruner.save(args.out, flat, pruner.THRESHOLD_OFF, meta)
        log(json.dumps({"saved": args.out, "trained": True}))
        return

    xv, gv, wv = dump(args.valid, 0x7A32)
    sv = pruner.score_reference(flat, xv).reshape(-1)
    gaps, gmax = pruner.margin_gaps(sv, gv, wv)
    gcpu = gv.cpu()
    scpu = sv.float().cpu()
    qs = [0.999, 0.9992, 0.9994, 0.9996, 0.9998, 0.9999, 1.0]
    gs = np.sort(gaps.numpy())
    cands = sorted({float(gs[min(len(gs) - 1, int(q * len(gs)))]) * 1.02 + 1e-3 for q in qs})
    below = (gmax[gcpu] - scpu).numpy()
    log("validation: " + ", ".join(f"margin {mg:.3f}: kept units {float((below <= mg).mean()):.4f}, "
                                   f"winners kept {float((gs <= mg).mean()):.5f}" for mg in cands))
    del xv, gv, wv, sv

    te = oc.generate(geo, args.test, seed=0x7A33, **WORKLOAD)
    plain, plain_ms = run(te)
    chosen, trials = None, []
    for mg in cands:
        pr, ms = run(te, flat, mg)
        rec, found, total = pruner.recall(plain, pr)
        entry = {"margin": mg, "track_recall": rec, "tracks": total, "tracks_identical": found,
                 "plain_ms": plain_ms, "pruned_ms": ms, "speedup": plain_ms / ms}
        trials.append(entry)
        log(f"margin {mg:.3f}: track recall {rec:.5f} ({found}/{total}), chain {ms:.1f} ms vs plain "
            f"{plain_ms:.1f} ms ({plain_ms / ms:.2f}x)")
        if chosen is None and rec >= args.target:
            chosen = entry
    sess.set_pruner(None)
    if chosen is None:
        raise RuntimeError("no margin reaches the target recall")
    meta = {"features": pruner.FEATURE_NAMES, "train_events": args.train, "train_sample": args.sample,
            "valid_events": args.valid,
            "test_events": args.test, "workload": WORKLOAD, "caps": oc.HIGH_OCCUPANCY_CAPS,
            "seeds": "train 2^20 0x7A31 + shard, valid 2^20 0x7A32 + shard, test 0x7A33", "loss": "listwise (softmax over a tracklet's units)",
            "margin": chosen["margin"], "target": args.target, "chosen": chosen, "trials": trials}
    pruner.save(args.out, flat, pruner.THRESHOLD_OFF, meta)
    log(json.dumps({"saved": args.out, **chosen}))


if __name__ == "__main__":
    main()
