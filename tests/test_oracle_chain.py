"""Whole-chain properties of the oracle on generated events.

Noise-free events with known truth: the chain must return global tracks
whose hits are the track's own hits and whose fitted parameters are close to
the truth.  Invariants: results per event do not depend on batch
composition, event order or thread count; forced caps set their status bits.
"""
import numpy as np
import pytest


def truth_hits(ev, e):
    h0, h1 = int(ev.offsets[e]), int(ev.offsets[e + 1])
    return set(range(h1 - h0))


def test_noise_free_tracks_found_with_truth_parameters(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 1500, seed=21, noise_occupancy=0.0, afterpulse_frac=0.0, n_tracks_min=1, n_tracks_max=1)
    p = oc.default_params()
    r = orc.reconstruct(ogeo, p, ev)
    assert (r.status == 0).all()
    found = np.nonzero(r.n_tracks > 0)[0]
    assert len(found) > 0.3 * ev.n_events
    dtx, dx0, dty, dy0, dinv = [], [], [], [], []
    for e in found:
        t = r.tracks[e, 0]
        x0, tx, y0, ty, invp, q = ev.truth[e, 0]
        dtx.append(t["tx"] - tx); dx0.append(t["x0"] - x0); dty.append(t["ty"] - ty); dy0.append(t["y0"] - y0)
        dinv.append(t["invp"] - invp)
        assert t["station_id"] == 6 and t["charge"] == q
        nh = int(t["nhits"])
        assert 12 <= nh <= 18
        # every hit on the track is one of the event's (single-track) hits, none repeated
        idx = t["hit_index"][:nh]
        assert len(set(idx.tolist())) == nh and idx.max() < ev.offsets[e + 1] - ev.offsets[e]
    med = [float(np.median(np.abs(v))) for v in (dtx, dx0, dty, dy0, dinv)]
    # drift smearing 0.04 cm over ~600 cm lever arms, y from stereo wires, 1/p from the KMAG kink
    assert med[0] < 5e-4 and med[1] < 0.5, med
    assert med[2] < 5e-3 and med[3] < 5.0, med
    assert med[4] < 0.02, med


def test_results_independent_of_batch_order_and_threads(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 600, seed=22, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    p = oc.default_params()
    a = orc.reconstruct(ogeo, p, ev, records=True, threads=8)
    b = orc.reconstruct(ogeo, p, ev, records=True, threads=1)
    same = np.array_equal(a.tracks.view(np.uint8), b.tracks.view(np.uint8)) and np.array_equal(a.records, b.records)
    assert same
    # reversed event order
    parts = [ev.subset(e, e + 1) for e in range(ev.n_events)][::-1]
    cat = lambda k: np.concatenate([getattr(x, k) for x in parts])
    offs = np.concatenate([[0], np.cumsum([x.n_hits for x in parts])]).astype(np.int64)
    rev = oc.Events(cat("det"), cat("elem"), cat("tdc"), cat("drift"), cat("flag"), offs)
    c = orc.reconstruct(ogeo, p, rev, records=True)
    assert np.array_equal(c.status[::-1], a.status)
    same = np.array_equal(np.ascontiguousarray(c.tracks[::-1]).view(np.uint8), a.tracks.view(np.uint8))
    assert same
    assert np.array_equal(c.records[::-1], a.records)


def test_forced_caps_set_status_bits(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 200, seed=23)
    r = orc.reconstruct(ogeo, oc.default_params(max_d2_mult=0), ev)
    assert ((r.status & oc.EV_D2_MULT) != 0).mean() > 0.9
    assert (r.n_tracks[(r.status & oc.EV_REJECT_MASK) != 0] == 0).all()
    r = orc.reconstruct(ogeo, oc.default_params(max_hits_chamber_plane=1), ev)
    assert ((r.status & oc.EV_PLANE_OVERFLOW) != 0).all()
    assert not r.keep_mask.any()
    r = orc.reconstruct(ogeo, oc.default_params(max_pairs_per_bin=0), ev)
    assert ((r.status & oc.EV_XZ_BIN_OVERFLOW) != 0).mean() > 0.5
    r = orc.reconstruct(ogeo, oc.default_params(max_tracks_out=0), ev)
    assert ((r.status & oc.EV_OUTPUT_OVERFLOW) != 0).any() and (r.n_tracks == 0).all()


def test_slot_layout_and_cleaning_outcome(oc, orc, geo, ogeo):
    # tracklets land in the 32 slots by parity round-robin; demoted ones keep stationID 3
    ev = oc.generate(geo, 400, seed=24, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    r = orc.reconstruct(ogeo, oc.default_params(), ev, records=True)
    for e in range(ev.n_events):
        n = r.n_tracklets[e]
        rec = r.records[e, :n]
        threads = rec[:, 1].astype(int)
        stid = rec[:, 0].astype(int)
        assert set(stid.tolist()) <= {3, 5, 6}
        # nhits consistent with the stage reached
        nh = rec[:, 2].astype(int)
        assert ((stid != 6) | (nh >= 9)).all() and ((stid != 5) | (nh >= 6)).all()
        assert (nh >= 2).all()
        # D3p tracklets come from even threads, D3m from odd ones; hit detectors agree
        for k in range(n):
            d3 = [d for d in rec[k, 22:22 + nh[k]].astype(int) if 19 <= d <= 30]
            if threads[k] % 2 == 0:
                assert all(19 <= d <= 24 for d in d3)
            else:
                assert all(25 <= d <= 30 for d in d3)


def test_yz_sign_state_model(oc, orc, geo, ogeo):
    # The device evaluates YZ candidates in parallel, deriving each one's
    # starting station-3 signs (unit seed + last writer of the last sign) instead
    # of carrying them; the oracle carries them and compares at every step.
    import ctypes as C
    lib = orc.load()
    cnt = (C.c_longlong * 3)()
    lib.orc_yz_state_check(1, None)
    try:
        for kw, n in ((dict(mult_mode=1, mult_mean=300.0, mult_cap=1000), 4000), ({}, 3000),
                      (dict(mult_mode=1, mult_mean=500.0, mult_cap=2000), 600)):
            ev = oc.generate(geo, n, seed=31, **kw)
            p = oc.default_params(max_hits_per_event=2048, max_d0_mult=2000, max_d2_mult=2000, max_d3_mult=2000)
            orc.reconstruct(ogeo, p, ev)
    finally:
        lib.orc_yz_state_check(0, C.cast(cnt, C.c_void_p))
    checks, bad, cross = list(cnt)
    assert checks > 1_000_000 and cross > 100_000, (checks, cross)
    assert bad == 0, f"{bad} of {checks} sign states differ from the model"
