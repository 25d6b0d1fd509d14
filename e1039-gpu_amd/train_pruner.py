"""Trains the YZ unit pruner on the plain chain's own decisions (needs the GPU).

    python e1039-gpu_amd/train_pruner.py [--train 200000] [--out e1039-gpu_amd/data/pruner_u3v3.npz]

1. Runs the plain chain on seeded synthetic events with the unit dump on
   (or_session_set_unit_dump): one row of 16 features per YZ unit, with a flag
   saying whether the tracklet's YZ winner lies in that unit.
2. Labels a unit positive when it holds the winner of a tracklet that ends as
   an output track (or_track.tracklet_id).
3. Fits the perceptron (pruner.train), then picks the threshold: candidates
   are the validation scores (bf16 restatement, pruner.score_reference) at
   several positive-unit recalls; each is run through the pruned chain on a
   separate event set and the largest one whose track recall (tracks output
   identically, pruner.recall) reaches --target is kept.
4. Saves weights, threshold and the measured figures.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200_000, help="training events")
    ap.add_argument("--valid", type=int, default=50_000, help="validation events (threshold candidates)")
    ap.add_argument("--test", type=int, default=100_000, help="events for the track-recall check")
    ap.add_argument("--mean-hits", type=float, default=300.0)
    ap.add_argument("--cap-hits", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--target", type=float, default=0.997, help="track recall the threshold must reach")
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

    def events(n, seed):
        return oc.generate(geo, n, seed=seed, mult_mode=1, mult_mean=args.mean_hits, mult_cap=args.cap_hits)

    def dump_rows(ev):
        """plain chain with unit rows; returns (rows, result)"""
        cap = ev.n_events * 160
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
        return rows[:n], res

    def label(rows, res):
        r = rows.cpu().numpy()
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
    tr_rows, tr_res = dump_rows(events(args.train, 0x7A11))
    xtr, ytr = label(tr_rows, tr_res)
    del tr_rows
    log(f"training rows {len(xtr)} ({len(xtr) / args.train:.1f}/event), positive {ytr.mean():.4f}, "
        f"{time.time() - t0:.1f}s")
    flat = pruner.train(xtr, ytr, epochs=args.epochs, device="cuda", log=log)
    log(f"trained in {time.time() - t0:.1f}s")

    va_rows, va_res = dump_rows(events(args.valid, 0x7A12))
    xva, yva = label(va_rows, va_res)
    s_ref = pruner.score_reference(flat, torch.as_tensor(xva, device=dev)).cpu().numpy()
    # device scores of the same rows must agree with the restatement
    s_dev = sess_score(sess, flat, xva, dev)
    log(f"device vs reference scores: max |diff| {np.abs(s_dev - s_ref).max():.3e}")
    pos = s_ref[yva]
    unit_recalls = [0.995, 0.998, 0.999, 0.9995, 0.9999, 1.0]
    cands = sorted({float(np.quantile(pos, 1 - r)) if r < 1 else float(pos.min()) for r in unit_recalls},
                   reverse=True)
    cands.append(float(pos.min()) - 1.0)

    te = events(args.test, 0x7A13)
    t = sess.upload(te)
    o = sess.alloc_outputs(te.n_events, te.n_hits, tracklets=True)
    sess.set_pruner(None)
    sess.run_device(t, o)
    torch.cuda.synchronize()
    plain = oc.Session.fetch(o, te.n_events, params.max_tracks_out, params.max_tracklets_per_event)
    sess.run_device(t, o)
    plain_ms = sess.last_kernel_ms()
    chosen = None
    trials = []
    for thr in cands:
        sess.set_pruner(flat, thr)
        sess.run_device(t, o)
        torch.cuda.synchronize()
        pr = oc.Session.fetch(o, te.n_events, params.max_tracks_out, params.max_tracklets_per_event)
        sess.run_device(t, o)
        ms = sess.last_kernel_ms()
        rec, found, total = pruner.recall(plain, pr)
        kept = float((s_ref >= thr).mean())
        trials.append({"threshold": thr, "track_recall": rec, "tracks": total, "kept_units": kept,
                       "plain_ms": plain_ms, "pruned_ms": ms, "speedup": plain_ms / ms})
        log(f"threshold {thr:.4f}: track recall {rec:.5f} ({found}/{total}), kept units {kept:.3f}, "
            f"chain {ms:.2f} ms vs plain {plain_ms:.2f} ms ({plain_ms / ms:.2f}x)")
        if rec >= args.target:
            chosen = trials[-1]
            break
    sess.set_pruner(None)
    if chosen is None:
        raise RuntimeError("no threshold reaches the target recall")
    meta = {"features": pruner.FEATURE_NAMES, "train_events": args.train, "valid_events": args.valid,
            "test_events": args.test, "workload": f"hits/event Poisson({args.mean_hits:g}) capped at {args.cap_hits}",
            "seeds": [0x7A11, 0x7A12, 0x7A13], "positive_fraction": float(ytr.mean()), "chosen": chosen,
            "trials": trials}
    pruner.save(args.out, flat, chosen["threshold"], meta)
    log(json.dumps({"saved": args.out, **chosen}))


def sess_score(sess, flat, x, dev):
    """device scores of rows x with weights flat (threshold irrelevant)"""
    import torch
    sess.set_pruner(flat, 0.0)
    s = sess.score_units(torch.as_tensor(x, device=dev))
    torch.cuda.synchronize()
    out = s.cpu().numpy()
    sess.set_pruner(None)
    return out


if __name__ == "__main__":
    main()
