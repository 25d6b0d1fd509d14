"""HIP chain against the CPU oracle, bit for bit, through the C ABI.

Integer/index outputs (status, keep bits, tracklet counts, hit indices, signs)
and every float of the tracklet records are compared as bit patterns: the
device build and the oracle both round every float operation once
(contraction off, IEEE division and square root)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_same(got, ref, ev, what=""):
    bad = np.nonzero(got.status != ref.status)[0]
    assert bad.size == 0, f"{what} status differs at events {bad[:10]}: {got.status[bad[:5]]} vs {ref.status[bad[:5]]}"
    assert np.array_equal(got.keep_mask, ref.keep_mask), f"{what} keep bits differ"
    bad = np.nonzero(got.n_tracklets != ref.n_tracklets)[0]
    assert bad.size == 0, f"{what} n_tracklets differ at {bad[:10]}"
    if ref.records is not None:
        g, r = got.records.view(np.uint32), ref.records.view(np.uint32)
        diff = np.nonzero((g != r).any(axis=2))
        assert diff[0].size == 0, (f"{what} records differ at (event, tracklet) {list(zip(diff[0][:5], diff[1][:5]))}; "
                                   f"fields {np.nonzero(g[diff[0][0], diff[1][0]] != r[diff[0][0], diff[1][0]])[0] if diff[0].size else ''}")
    bad = np.nonzero(got.n_tracks != ref.n_tracks)[0]
    assert bad.size == 0, f"{what} n_tracks differ at {bad[:10]}"
    same = same_tracks(got, ref)
    assert same, f"{what} track records differ"


def same_tracks(a, b):
    # slots beyond n_tracks are not written by the library; compare the valid ones
    return bool(np.array_equal(a.n_tracks, b.n_tracks) and
                np.array_equal(a.valid_tracks().view(np.uint8), b.valid_tracks().view(np.uint8)))


def run_both(oc, orc, session, ogeo, ev, records=True):
    got = session.reconstruct(ev, records=records)
    ref = orc.reconstruct(ogeo, session.params, ev, records=records)
    return got, ref


def test_c0_low_occupancy(oc, orc, geo, ogeo, session):
    ev = oc.generate(geo, 3000, seed=0xE1039)
    got, ref = run_both(oc, orc, session, ogeo, ev)
    assert ref.n_tracks.sum() > 500
    assert_same(got, ref, ev, "c0")


def test_c1_poisson300(oc, orc, geo, ogeo, session):
    ev = oc.generate(geo, 2000, seed=11, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    got, ref = run_both(oc, orc, session, ogeo, ev)
    assert_same(got, ref, ev, "c1")


def test_high_occupancy_reference_caps(oc, orc, geo, ogeo, session):
    # most events breach the station multiplicity cuts; status bits must agree
    ev = oc.generate(geo, 300, seed=12, mult_mode=1, mult_mean=1200.0, mult_cap=2000, noise_occupancy=0.08)
    got, ref = run_both(oc, orc, session, ogeo, ev)
    assert (ref.status != 0).mean() > 0.5
    assert_same(got, ref, ev, "high occupancy")


def test_yz_unit_path(oc, orc, geo, ogeo):
    # a YZ candidate list of one entry sends almost every chunk through the
    # unit-by-unit path; both paths must give the oracle's result
    sess = oc.Session(geo, oc.default_params(), device=0)
    sess.set_tuning(oc.TUNE_YZ_LIST_CAP, 1)
    ev = oc.generate(geo, 1000, seed=15, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    got, ref = run_both(oc, orc, sess, ogeo, ev)
    assert_same(got, ref, ev, "yz unit path")
    sess.set_tuning(oc.TUNE_YZ_LIST_CAP, 40)   # mixed: small chunks parallel, large ones by unit
    got = sess.reconstruct(ev, records=True)
    assert_same(got, ref, ev, "yz mixed paths")


def test_lifted_multiplicity_caps(oc, orc, geo, ogeo):
    params = oc.default_params(max_hits_per_event=2048, max_d0_mult=2000, max_d2_mult=2000, max_d3_mult=2000)
    sess = oc.Session(geo, params, device=0)
    ev = oc.generate(geo, 200, seed=13, mult_mode=1, mult_mean=700.0, mult_cap=2000)
    got, ref = run_both(oc, orc, sess, ogeo, ev)
    assert_same(got, ref, ev, "lifted caps")


def test_edge_cases(oc, orc, geo, ogeo, session):
    rng = np.random.default_rng(5)
    base = oc.generate(geo, 40, seed=14)
    evs = []
    for e in range(base.n_events):
        sub = base.subset(e, e + 1)
        k = e % 8
        det, elem, tdc, drift, flag = sub.det.copy(), sub.elem.copy(), sub.tdc.copy(), sub.drift.copy(), sub.flag.copy()
        if k == 1:     # empty event
            det, elem, tdc, drift, flag = det[:0], elem[:0], tdc[:0], drift[:0], flag[:0]
        elif k == 2:   # dark-photon planes beyond 54 are skipped
            det = np.concatenate([det, np.array([55, 60], np.int16)]); elem = np.concatenate([elem, np.array([1, 2], np.int16)])
            tdc = np.concatenate([tdc, np.zeros(2, np.float32)]); drift = np.concatenate([drift, np.zeros(2, np.float32)])
            flag = np.concatenate([flag, np.full(2, 2, np.uint8)])
        elif k == 3:   # element outside the plane
            elem[len(elem) // 2] = 999
        elif k == 4:   # detector 0
            det[0] = 0
        elif k == 5:   # out-of-time hits
            flag[rng.random(len(flag)) < 0.3] = 1
        elif k == 6:   # negative drift values are used as |drift|
            drift = -drift
        elif k == 7:   # unsorted hits within the event
            perm = rng.permutation(len(det)); det, elem, tdc, drift, flag = det[perm], elem[perm], tdc[perm], drift[perm], flag[perm]
        evs.append((det, elem, tdc, drift, flag))
    cat = lambda i: np.concatenate([x[i] for x in evs])
    offs = np.concatenate([[0], np.cumsum([len(x[0]) for x in evs])]).astype(np.int64)
    ev = oc.Events(cat(0), cat(1), cat(2), cat(3), cat(4), offs)
    got, ref = run_both(oc, orc, session, ogeo, ev)
    assert (ref.status & oc.EV_BAD_HIT).sum() >= 8
    assert_same(got, ref, ev, "edge cases")


def test_large_tiles(oc, orc, geo, ogeo):
    # 2048-, 4096- and 8192-hit tiles and the too-large flag above 8192; with the
    # reference caps these events stop at the station multiplicity cuts, after
    # the reducer (whose keep bits are compared)
    params = oc.default_params(max_hits_per_event=8192)
    sess = oc.Session(geo, params, device=0)
    parts = [oc.generate(geo, 20, seed=18 + k, mult_mode=1, mult_mean=m, mult_cap=9000)
             for k, m in enumerate((1800.0, 3500.0, 8190.0))]
    cat = lambda f: np.concatenate([getattr(x, f) for x in parts])
    base = np.cumsum([0] + [x.n_hits for x in parts])
    offs = np.concatenate([[0]] + [x.offsets[1:] + b for x, b in zip(parts, base)]).astype(np.int64)
    ev = oc.Events(cat("det"), cat("elem"), cat("tdc"), cat("drift"), cat("flag"), offs)
    counts = np.diff(ev.offsets)
    for lo, hi in ((1024, 2048), (2048, 4096), (4096, 8192), (8192, 10**9)):
        assert ((counts > lo) & (counts <= hi)).sum() > 3
    got, ref = run_both(oc, orc, sess, ogeo, ev)
    assert ((ref.status & oc.EV_EVENT_TOO_LARGE) != 0).sum() == (counts > 8192).sum()
    assert_same(got, ref, ev, "large tiles")


def test_event_too_large(oc, orc, geo, ogeo):
    params = oc.default_params(max_hits_per_event=512)
    sess = oc.Session(geo, params, device=0)
    ev = oc.generate(geo, 200, seed=15, mult_mode=2, mult_mean=300.0, mult_mean2=1500.0, mult_frac2=0.2, mult_cap=2000)
    got, ref = run_both(oc, orc, sess, ogeo, ev)
    assert (ref.status & oc.EV_EVENT_TOO_LARGE).sum() > 10
    assert_same(got, ref, ev, "too large")


def test_deterministic_and_batch_independent(oc, geo, session):
    ev = oc.generate(geo, 4000, seed=16, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    a = session.reconstruct(ev)
    b = session.reconstruct(ev)
    same = same_tracks(a, b) and np.array_equal(a.status, b.status)
    assert same
    # a sub-batch gives the same per-event results as the full batch
    i0, i1 = 1234, 2345
    s = session.reconstruct(ev.subset(i0, i1))
    assert np.array_equal(s.status, a.status[i0:i1])
    sub = oc.Result(a.status[i0:i1], a.keep_mask, a.n_tracks[i0:i1], a.tracks[i0:i1])
    same = same_tracks(s, sub)
    assert same
    h0, h1 = int(ev.offsets[i0]), int(ev.offsets[i1])
    assert np.array_equal(s.keep_bits(h1 - h0), a.keep_bits(ev.n_hits)[h0:h1])


def test_host_buffer_path(oc, orc, geo, ogeo, session):
    ev = oc.generate(geo, 500, seed=17)
    got = session.reconstruct_host(ev, records=True)
    ref = orc.reconstruct(ogeo, session.params, ev, records=True)
    assert_same(got, ref, ev, "host path")


def test_graph_replay_and_tuning(oc, orc, geo, ogeo):
    # captured launch sequence replayed several times, with eager calls in between;
    # every replay must equal the oracle (and so the eager path); tuning knobs and
    # the autotuner leave results unchanged
    import torch
    sess = oc.Session(geo, oc.default_params(), device=0)
    ev = oc.generate(geo, 3000, seed=19, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    ref = orc.reconstruct(ogeo, sess.params, ev, records=True)
    din = sess.upload(ev)
    dout = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True, records=True)
    sess.set_tuning(oc.TUNE_GRAPH, 1)
    other = oc.generate(geo, 500, seed=20)
    for rep in range(3):
        for k in ("status", "n_tracks", "tracks", "n_tracklets", "records", "keep_mask"):
            dout[k].fill_(0x5A if k != "records" else 7)
        sess.run_device(din, dout)
        torch.cuda.synchronize()
        got = oc.Session.fetch(dout, ev.n_events, sess.params.max_tracks_out, sess.params.max_tracklets_per_event)
        assert_same(got, ref, ev, f"graph replay {rep}")
        sess.set_tuning(oc.TUNE_GRAPH, 0)
        sess.reconstruct(other)            # eager work between replays
        sess.set_tuning(oc.TUNE_GRAPH, 1)
    sess.set_tuning(oc.TUNE_GRAPH, 0)
    sess.autotune(din, dout, reps=1)
    torch.cuda.synchronize()
    assert sess.get_tuning(oc.TUNE_GRID_PERCENT) in (100, 75, 50)
    assert sess.get_tuning(oc.TUNE_YZ_LIST_CAP) in (4096, 2048, 1024, 512)
    assert sess.get_tuning(oc.TUNE_WAVES_PER_EVENT) in (1, 2, 4, 8)
    got = oc.Session.fetch(dout, ev.n_events, sess.params.max_tracks_out, sess.params.max_tracklets_per_event)
    assert_same(got, ref, ev, "after autotune")
    for pct in (10, 60):
        sess.set_tuning(oc.TUNE_GRID_PERCENT, pct)
        got = sess.reconstruct(ev, records=True)
        assert_same(got, ref, ev, f"grid {pct}%")


def test_vertexing(oc, orc, geo, ogeo, session):
    # FP64 swim of every global track, bit-exact with the oracle, on both paths
    ev = oc.generate(geo, 2000, seed=21, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    ref = orc.reconstruct(ogeo, session.params, ev, records=True, vertex=True)
    got = session.reconstruct(ev, records=True, vertex=True)
    assert ref.n_tracks.sum() > 500
    assert_same(got, ref, ev, "vertexing")
    k = np.arange(ref.tracks.shape[1])[None, :] < ref.n_tracks[:, None]
    same = np.array_equal(got.vertex[k].view(np.uint32), ref.vertex[k].view(np.uint32))
    assert same, "vertex differs"
    host = session.reconstruct_host(ev.subset(0, 500), vertex=True)
    k5 = k[:500]
    same = np.array_equal(host.vertex[k5].view(np.uint32), ref.vertex[:500][k5].view(np.uint32))
    assert same, "vertex differs on the host-buffer path"


def test_full_chain(oc, orc, geo, ogeo):
    # prop-tube segment matching, global cleaning and vertexing after global
    # tracking; outputs are the stationID 7 tracks
    sess = oc.Session(geo, oc.default_params(full_chain=1), device=0)
    for kw, n in ((dict(mult_mode=1, mult_mean=300.0, mult_cap=1000), 2000), ({}, 2000)):
        ev = oc.generate(geo, n, seed=22, **kw)
        ref = orc.reconstruct(ogeo, sess.params, ev, records=True, vertex=True)
        got = sess.reconstruct(ev, records=True, vertex=True)
        assert ref.n_tracks.sum() > 300
        assert_same(got, ref, ev, "full chain")
        k = np.arange(ref.tracks.shape[1])[None, :] < ref.n_tracks[:, None]
        same = np.array_equal(got.vertex[k].view(np.uint32), ref.vertex[k].view(np.uint32))
        assert same, "vertex differs"



def test_dense_prop_planes(oc, orc, geo, ogeo, session):
    # The YZ prop-tube Y tests search each plane's sorted positions when the
    # plane holds at most psort_cap hits (32 in the 512-hit tile) and scan it
    # hit by hit otherwise: events of the 512-hit tile with 0 to 60 extra hits on
    # each of planes 47, 48, 53, 54 take both forms, in one batch
    rng = np.random.default_rng(21)
    base = oc.generate(geo, 600, seed=22, mult_mode=1, mult_mean=250.0, mult_cap=270)
    evs = []
    for e in range(base.n_events):
        sub = base.subset(e, e + 1)
        parts = [(sub.det, sub.elem, sub.tdc, sub.drift, sub.flag)]
        extra = (e % 4) * 20   # 0, 20, 40, 60 per plane
        for d in (47, 48, 53, 54):
            if extra:
                parts.append((np.full(extra, d, np.int16), rng.integers(1, 73, extra).astype(np.int16),
                              rng.uniform(0, 100, extra).astype(np.float32), rng.uniform(0, 2.5, extra).astype(np.float32),
                              np.full(extra, 2, np.uint8)))
        evs.append(tuple(np.concatenate([p[i] for p in parts]) for i in range(5)))
    offs = np.concatenate([[0], np.cumsum([len(x[0]) for x in evs])]).astype(np.int64)
    assert np.diff(offs).max() <= 512
    cat = lambda i: np.concatenate([x[i] for x in evs])
    ev = oc.Events(cat(0), cat(1), cat(2), cat(3), cat(4), offs)
    got, ref = run_both(oc, orc, session, ogeo, ev)
    assert ref.n_tracks.sum() > 100
    assert_same(got, ref, ev, "dense prop planes")
