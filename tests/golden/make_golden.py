"""Writes the golden fixtures under tests/golden/.

c0_small.npz: 200 seeded low-occupancy events (1-2 tracks, 2% noise, 5%
after-pulses) and 100 events at Poisson(300) hits, with the oracle's outputs
(status, keep bits, tracklet counts, tracklet records in slot order, global
tracks) and, from a second run with vertexing, the vertex of every global
track.  The inputs are stored too, so the fixture does not depend on the
generator staying unchanged; when the file exists its inputs are reused.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "e1039-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import orchain as oc  # noqa: E402
import oracle as orc  # noqa: E402


def concat(a, b):
    cat = lambda k: np.concatenate([getattr(a, k), getattr(b, k)])
    offs = np.concatenate([a.offsets, b.offsets[1:] + a.offsets[-1]]).astype(np.int64)
    return oc.Events(cat("det"), cat("elem"), cat("tdc"), cat("drift"), cat("flag"), offs)


def main():
    geo = oc.Geometry.load()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c0_small.npz")
    if os.path.exists(out):
        z = np.load(out)
        ev = oc.Events(z["det"], z["elem"], z["tdc"], z["drift"], z["flag"], z["offsets"])
    else:
        ev = concat(oc.generate(geo, 200, seed=0xE1039),
                    oc.generate(geo, 100, seed=0xE1040, mult_mode=1, mult_mean=300.0, mult_cap=1000))
    p = oc.default_params()
    og = orc.OracleGeometry(oc.DEFAULT_GEOMETRY)
    r = orc.reconstruct(og, p, ev, records=True)
    rv = orc.reconstruct(og, p, ev, vertex=True)
    ntk = r.n_tracklets
    rec = np.concatenate([r.records[e, :ntk[e]] for e in range(ev.n_events)]).astype(np.float32)
    vtx = rv.vertex[np.arange(p.max_tracks_out)[None, :] < rv.n_tracks[:, None]].astype(np.float32)
    np.savez_compressed(out, det=ev.det, elem=ev.elem, tdc=ev.tdc, drift=ev.drift, flag=ev.flag, offsets=ev.offsets,
                        status=r.status, keep_mask=r.keep_mask, n_tracklets=ntk, records=rec,
                        n_tracks=r.n_tracks, tracks=r.tracks.view(np.uint8).reshape(ev.n_events, -1), vertex=vtx)
    print(f"wrote {out}: {ev.n_events} events, {ev.n_hits} hits, {int(ntk.sum())} tracklets, "
          f"{int(r.n_tracks.sum())} tracks")


if __name__ == "__main__":
    main()
