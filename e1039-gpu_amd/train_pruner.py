"""Trains the YZ unit pruner on the plain chain's own decisions (needs the GPU).

    python e1039-gpu_amd/train_pruner.py [--train 100000] [--out e1039-gpu_amd/data/pruner_u3v3.npz]

1. Runs the plain chain on seeded synthetic events of each workload (hits per
   event Poisson(300) capped at 1000, and 8 % noise occupancy per channel)
   with the unit dump on (or_session_set_unit_dump): one row of 16 features
   per YZ unit, with a flag saying whether the tracklet's YZ winner lies in
   that unit.
2. Labels a unit positive when it holds the winner of a tracklet that ends as
   an output track (or_track.tracklet_id).
3. Fits one perceptron on the rows of both workloads (pruner.train), then picks
   the threshold: candidates are validation scores at several positive-unit
   recalls; each is run through the pruned chain on separate event sets of
   both workloads, and the largest one whose track recall (tracks output
   identically, pruner.recall) reaches --target on every workload is kept.
4. Saves weights, threshold and the measured figures.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# name -> generator settings (include/e1039_or_gen.h)
WORKLOADS = {
    "poisson300": dict(mult_mode=1, mult_mean=300.0, mult_cap=1000),
    "occ8": dict(mult_mode=0, noise_occupancy=0.08),
}


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=100_000, help="training events per workload")
    ap.add_argument("--valid", type=int, default=30_000, help="validation events per workload")
    ap.add_argument("--test", type=int, default=100_000, help="events per workload for the track-recall check")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--target", type=float, default=0.995, help="track recall the threshold must reach")
    ap.add_argument("--out", default=os.path.join(HERE, "data", "pruner_u3v3.npz"))
    args = ap.parse_args()

    import numpy as np
    import torch

    import orchain as oc
    import pruner

    geo = oc.Geometry.load()
    params = oc.default_params(max_hits_per_event=1024)
    sess = oc.Session(geo, params, device=0)
    dev = torch.device("cuda", 0)

    def events(name, n, seed):
        return oc.generate(geo, n, seed=seed, **WORKLOADS[name])

    def dump_rows(ev):
        """plain chain with unit rows; returns (rows, result)"""
        cap = ev.n_events * 1024
        rows = torch.empty((cap, oc.UNIT_ROW), dtype=torch.float32, device=dev)
        count = torch.zeros(1, dtype=torch.int64, device=dev)
        sess.set_pruner(None)
        sess.set_unit_dump(rows, count)
        t = sess.upload(ev)
        o = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
        sess.run_device(t, o)
        torch.cuda.synchronize()
        sess.set_unit_dump(None)
        n = int(count.item())
        if n > cap:
            raise RuntimeError(f"unit rows: {n} wanted, {cap} held")
        res = oc.Session.fetch(o, ev.n_events, params.max_tracks_out, params.max_tracklets_per_event)
        out = rows[:n].cpu()
        del rows
        return out, res

    def label(rows, res):
        r = rows.numpy()
        good = r[:, 3] >= 0
        r = r[good]
        # (event, tracklet_id) of every output track, as one int64 key
        nt = res.n_tracks.astype(np.int64)
        slot = np.arange(res.tracks.shape[1])[None, :] < nt[:, None]
        ev_of = np.broadcast_to(np.arange(len(nt))[:, None], slot.shape)[slot]
        final = ev_of.astype(np.int64) * 4096 + res.tracks["tracklet_id"][slot].astype(np.int64)
        key = r[:, 0].astype(np.int64) * 4096 + r[:, 1].astype(np.int64)
        y = (r[:, 3] > 0.5) & np.isin(key, final)
        return r[:, 4:4 + pruner.NF].astype(np.float32), y

    t0 = time.time()
    xs, ys = [], []
    for i, name in enumerate(WORKLOADS):
        rows, res = dump_rows(events(name, args.train, 0x7A11 + 16 * i))
        x, y = label(rows, res)
        log(f"{name}: training rows {len(x)} ({len(x) / args.train:.1f}/event), positive {y.mean():.4f}")
        xs.append(x)
        ys.append(y)
    xtr, ytr = np.concatenate(xs), np.concatenate(ys)
    del xs, ys
    flat = pruner.train(xtr, ytr, epochs=args.epochs, device="cuda", log=log)
    log(f"trained on {len(xtr)} rows in {time.time() - t0:.1f}s")

    pos_scores, all_scores = [], []
    for i, name in enumerate(WORKLOADS):
        rows, res = dump_rows(events(name, args.valid, 0x7A12 + 16 * i))
        xva, yva = label(rows, res)
        s_ref = pruner.score_reference(flat, torch.as_tensor(xva, device=dev)).cpu().numpy()
        pos_scores.append(s_ref[yva])
        all_scores.append(s_ref)
    pos = np.concatenate(pos_scores)
    unit_recalls = [0.995, 0.998, 0.999, 0.9995, 0.9999, 1.0]
    cands = sorted({float(np.quantile(pos, 1 - r)) if r < 1 else float(pos.min()) for r in unit_recalls},
                   reverse=True)
    cands.append(float(pos.min()) - 1.0)

    tests = {}
    for i, name in enumerate(WORKLOADS):
        te = events(name, args.test, 0x7A13 + 16 * i)
        t = sess.upload(te)
        o = sess.alloc_outputs(te.n_events, te.n_hits, tracklets=True)
        sess.set_pruner(None)
        sess.run_device(t, o)
        torch.cuda.synchronize()
        plain = oc.Session.fetch(o, te.n_events, params.max_tracks_out, params.max_tracklets_per_event)
        sess.run_device(t, o)
        tests[name] = (te, t, o, plain, sess.last_kernel_ms())
    chosen = None
    trials = []
    for thr in cands:
        sess.set_pruner(flat, thr)
        entry = {"threshold": thr, "workloads": {}}
        ok = True
        for i, name in enumerate(WORKLOADS):
            te, t, o, plain, plain_ms = tests[name]
            sess.run_device(t, o)
            torch.cuda.synchronize()
            pr = oc.Session.fetch(o, te.n_events, params.max_tracks_out, params.max_tracklets_per_event)
            sess.run_device(t, o)
            ms = sess.last_kernel_ms()
            rec, found, total = pruner.recall(plain, pr)
            kept = float((all_scores[i] >= thr).mean())
            entry["workloads"][name] = {"track_recall": rec, "tracks": total, "kept_units": kept,
                                        "plain_ms": plain_ms, "pruned_ms": ms, "speedup": plain_ms / ms}
            log(f"threshold {thr:.4f} {name}: track recall {rec:.5f} ({found}/{total}), kept units {kept:.3f}, "
                f"chain {ms:.2f} ms vs plain {plain_ms:.2f} ms ({plain_ms / ms:.2f}x)")
            ok = ok and rec >= args.target
        trials.append(entry)
        if ok:
            chosen = entry
            break
    sess.set_pruner(None)
    if chosen is None:
        raise RuntimeError("no threshold reaches the target recall")
    meta = {"features": pruner.FEATURE_NAMES, "train_events": args.train, "valid_events": args.valid,
            "test_events": args.test, "workloads": WORKLOADS, "seeds": "0x7A11/12/13 + 16 * workload index",
            "positive_fraction": float(ytr.mean()), "target": args.target, "chosen": chosen, "trials": trials}
    pruner.save(args.out, flat, chosen["threshold"], meta)
    log(json.dumps({"saved": args.out, **chosen}))


if __name__ == "__main__":
    main()
