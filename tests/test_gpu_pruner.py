"""Learned YZ unit pruning on the device, through the C ABI.

* The matrix-core perceptron (or_pruner_score) against the torch fp32
  restatement of the same arithmetic (pruner.score_reference), and each of its
  layers (or_pruner_layers) against float64 within the fp32 summation bound.
* The unit-list path with every unit kept (threshold -inf) against the CPU
  oracle, bit for bit: enumerating kept units is the plain chain when nothing
  is pruned.
* The unit rows (or_session_set_unit_dump): one winner row per output track.
* The pruned chain is the plain chain over the units it searched: those units
  (from its unit rows) fed to the oracle's unit filter give its outputs bit
  for bit.
* The shipped pruner keeps >= 99.5 % of the plain chain's tracks identical.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def psess(oc, geo):
    s = oc.Session(geo, oc.default_params(max_hits_per_event=2048), device=0)
    yield s
    s.close()


def random_weights(seed):
    import pruner
    rng = np.random.default_rng(seed)
    return pruner.pack(rng.normal(0, 0.5, (64, 16)), rng.normal(0, 0.2, 64), rng.normal(0, 0.2, (64, 64)),
                       rng.normal(0, 0.2, 64), rng.normal(0, 0.3, 64), 0.1)


def test_scores_match_reference(psess):
    import torch

    import pruner
    flat = random_weights(1)
    rng = np.random.default_rng(2)
    feats = rng.normal(0, 1, (1000, 16)).astype(np.float32)   # not a multiple of 64 rows
    psess.set_pruner(flat, 0.0)
    try:
        dev = psess.score_units(torch.as_tensor(feats, device="cuda"))
        torch.cuda.synchronize()
        dev = dev.cpu().numpy()
    finally:
        psess.set_pruner(None)
    ref = pruner.score_reference(flat, torch.as_tensor(feats)).numpy()
    # fp32 sums in another order can move a bf16 rounding of a hidden unit by one step
    err = np.abs(dev - ref) / (1.0 + np.abs(ref))
    assert np.quantile(err, 0.999) < 1e-2, f"scores differ: {np.quantile(err, 0.999):.3e}"
    assert err.max() < 5e-2, f"worst score differs by {err.max():.3e}"
    assert np.std(ref) > 0.5   # the scores do depend on the features


def bf16(x):
    """float32 -> bfloat16 (round to nearest even) -> float64, as the device rounds operands"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_hidden_layers_exact(psess):
    # Each layer of the matrix-core perceptron against float64 on the same
    # bf16 operands, fed with the device's own previous layer: bf16 x bf16
    # products are exact in fp32, so the only error is the fp32 summation
    # (at most 64 terms), bounded by 64 * 2^-24 of the sum of |terms|.  A
    # lane-layout slip (a transposed tile, a wrong row group) moves values by
    # the order of the terms themselves.
    import torch
    flat = random_weights(4).astype(np.float64)
    rng = np.random.default_rng(5)
    n = 1000                                                     # not a multiple of 64 rows
    feats = rng.normal(0, 1, (n, 16)).astype(np.float32)
    psess.set_pruner(flat.astype(np.float32), 0.0)
    try:
        sc, p1, p2 = psess.score_layers(torch.as_tensor(feats, device="cuda"))
        torch.cuda.synchronize()
        sc, p1, p2 = sc.cpu().numpy().astype(np.float64), p1.cpu().numpy().astype(np.float64), p2.cpu().numpy().astype(np.float64)
    finally:
        psess.set_pruner(None)
    o = 0
    w1 = flat[o:o + 1024].reshape(64, 16); o += 1024
    b1 = flat[o:o + 64]; o += 64
    w2 = flat[o:o + 4096].reshape(64, 64); o += 4096
    b2 = flat[o:o + 64]; o += 64
    w3 = flat[o:o + 64]; o += 64
    b3 = flat[o]
    f32 = lambda a: a.astype(np.float32).astype(np.float64)   # biases and the output layer are fp32 on the device
    x, W1, W2 = bf16(feats), bf16(w1), bf16(w2)
    ref1 = x @ W1.T + f32(b1)
    bound1 = 64 * 2.0 ** -24 * (np.abs(x) @ np.abs(W1.T) + np.abs(f32(b1))) + 1e-30
    assert (np.abs(p1 - ref1) <= bound1).all(), f"layer 1: worst {np.max(np.abs(p1 - ref1) / bound1):.3g} of the bound"
    h = bf16(np.maximum(p1, 0.0))
    ref2 = h @ W2.T + f32(b2)
    bound2 = 64 * 2.0 ** -24 * (np.abs(h) @ np.abs(W2.T) + np.abs(f32(b2))) + 1e-30
    assert (np.abs(p2 - ref2) <= bound2).all(), f"layer 2: worst {np.max(np.abs(p2 - ref2) / bound2):.3g} of the bound"
    r = np.maximum(p2, 0.0)
    ref3 = r @ f32(w3) + f32(np.array(b3))
    bound3 = 70 * 2.0 ** -24 * (r @ np.abs(f32(w3)) + abs(f32(np.array(b3)))) + 1e-30
    assert (np.abs(sc - ref3) <= bound3).all(), f"output: worst {np.max(np.abs(sc - ref3) / bound3):.3g} of the bound"
    # the layers are not degenerate: both signs before ReLU, scores spread
    assert (p1 > 0).mean() > 0.2 and (p1 < 0).mean() > 0.2 and (p2 > 0).mean() > 0.1 and np.std(sc) > 0.5
    # and or_pruner_score gives the same scores
    psess.set_pruner(flat.astype(np.float32), 0.0)
    try:
        only = psess.score_units(torch.as_tensor(feats, device="cuda")).cpu().numpy()
    finally:
        psess.set_pruner(None)
    assert np.array_equal(only.view(np.uint32), sc.astype(np.float32).view(np.uint32))


def same_outputs(got, ref):
    assert np.array_equal(got.status, ref.status)
    assert np.array_equal(got.keep_mask, ref.keep_mask)
    assert np.array_equal(got.n_tracklets, ref.n_tracklets)
    assert np.array_equal(got.records.view(np.uint32), ref.records.view(np.uint32))
    assert np.array_equal(got.n_tracks, ref.n_tracks)
    assert np.array_equal(got.valid_tracks().view(np.uint8), ref.valid_tracks().view(np.uint8))


def test_keep_all_units_is_the_plain_chain(oc, orc, geo, ogeo, psess):
    import torch
    ev = oc.generate(geo, 1500, seed=21, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    rows = torch.empty((ev.n_events * 400, oc.UNIT_ROW), dtype=torch.float32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    psess.set_pruner(random_weights(3), float("-inf"))
    psess.set_unit_dump(rows, count)
    try:
        got = psess.reconstruct(ev, records=True)
    finally:
        psess.set_unit_dump(None)
        psess.set_pruner(None)
    ref = orc.reconstruct(ogeo, psess.params, ev, records=True)
    assert ref.n_tracks.sum() > 300
    same_outputs(got, ref)
    n = int(count.item())
    assert 0 < n <= rows.shape[0]
    r = rows[:n].cpu().numpy()
    assert set(np.unique(r[:, 3])) <= {-1.0, 0.5, 1.0}   # every unit searched
    assert (r[:, 0] >= 0).all() and (r[:, 0] < ev.n_events).all()
    # every output track's tracklet has exactly one winner unit
    win = r[r[:, 3] == 1.0]
    wkeys, wcount = np.unique(win[:, 0].astype(np.int64) * 4096 + win[:, 1].astype(np.int64), return_counts=True)
    assert (wcount == 1).all()
    unknown = set((r[r[:, 3] < 0][:, 0].astype(np.int64) * 4096 + r[r[:, 3] < 0][:, 1].astype(np.int64)).tolist())
    for e in np.nonzero(got.n_tracks)[0]:
        for k in range(got.n_tracks[e]):
            key = int(e) * 4096 + int(got.tracks[e, k]["tracklet_id"])
            assert key in set(wkeys.tolist()) or key in unknown


def test_margin_keeping_everything_is_the_plain_chain(oc, orc, geo, ogeo):
    # relative pruning with a margin wider than any score spread keeps every unit
    import pruner
    sess = oc.Session(geo, oc.high_occupancy_params(), device=0)
    ev = oc.generate(geo, 60, seed=23, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    sess.set_pruner_margin(1.0e30)
    sess.set_pruner(random_weights(4), pruner.THRESHOLD_OFF)
    try:
        got = sess.reconstruct(ev, records=True)
    finally:
        sess.set_pruner(None)
        sess.set_pruner_margin()
    ref = orc.reconstruct(ogeo, sess.params, ev, records=True)
    same_outputs(got, ref)
    sess.close()


PRUNED_WORKLOADS = {
    # mean 1,200 hits (the learned-pruning configuration)
    "mean1200": (120, dict(mult_mode=1, mult_mean=1200.0, mult_cap=4000), {}, 1000),
    # the batched mixture: Poisson(200) events run in the one-wave 512-hit tile
    "mixture": (300, dict(mult_mode=2, mult_mean=200.0, mult_mean2=1500.0, mult_frac2=0.2, mult_cap=4000), {}, 300),
    # high intensity with the full chain (prop matching, global cleaning,
    # vertexing; only stationID >= 7 tracks come out)
    "mean2000_full": (60, dict(mult_mode=1, mult_mean=2000.0, mult_cap=6000, n_tracks_min=1, n_tracks_max=4),
                      dict(max_tracklets_per_event=4096, full_chain=1), 60),
}


@pytest.mark.parametrize("workload", sorted(PRUNED_WORKLOADS))
def test_pruned_chain_is_the_plain_chain_over_its_searched_units(oc, orc, geo, ogeo, workload):
    # the shipped pruner; tracklets without unit rows (chunks beyond the unit
    # list) and chunks that went unit by unit (flag -1) searched every unit, and
    # the oracle's filter leaves them unrestricted
    import torch

    import pruner
    if not os.path.exists(oc.DEFAULT_PRUNER):
        pytest.skip("no shipped pruner weights")
    n_events, gen, caps, min_tracks = PRUNED_WORKLOADS[workload]
    flat, thr, meta = pruner.load()
    sess = oc.Session(geo, oc.high_occupancy_params(**caps), device=0)
    ev = oc.generate(geo, n_events, seed=0xF17, **gen)
    rows = torch.empty((12_000_000, oc.UNIT_ROW), dtype=torch.float32, device="cuda")   # 960 MB
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    sess.set_pruner_margin(meta["margin"])
    sess.set_pruner(flat, thr)
    sess.set_unit_dump(rows, count)
    try:
        got = sess.reconstruct(ev, records=True, vertex=True)
    finally:
        sess.set_unit_dump(None)
        sess.set_pruner(None)
        sess.set_pruner_margin()
    n = int(count.item())
    assert 0 < n <= rows.shape[0]
    r = rows[:n].cpu().numpy()
    del rows
    flag = r[:, 3]
    assert set(np.unique(flag)) <= {-1.0, 0.0, 0.5, 1.0}
    assert (flag == 0.0).mean() > 0.5, "the pruner kept most units"
    s = flag >= 0.5
    keys = orc.unit_keys(r[s, 0].astype(np.int64), r[s, 1].astype(np.int64), r[s, 2].astype(np.int64))
    ref = orc.reconstruct(ogeo, sess.params, ev, records=True, vertex=True, units=keys)
    assert ref.n_tracks.sum() > min_tracks
    same_outputs(got, ref)
    k = np.arange(ref.tracks.shape[1])[None, :] < ref.n_tracks[:, None]
    assert np.array_equal(got.vertex[k].view(np.uint32), ref.vertex[k].view(np.uint32)), "vertex differs"
    # the pruning changed some tracklet's YZ winner (the comparison is not the plain chain's)
    plain = orc.reconstruct(ogeo, sess.params, ev, records=True, vertex=True)
    assert not np.array_equal(plain.records.view(np.uint32), ref.records.view(np.uint32))
    sess.close()


def test_pruned_graph_replay(oc, geo):
    # the batched form with pruning: the captured launch sequence replayed
    # several times, with eager calls in between, equals the eager pruned chain
    import torch

    import pruner
    if not os.path.exists(oc.DEFAULT_PRUNER):
        pytest.skip("no shipped pruner weights")
    flat, thr, meta = pruner.load()
    sess = oc.Session(geo, oc.high_occupancy_params(), device=0)
    ev = oc.generate(geo, 400, seed=0xB7, mult_mode=2, mult_mean=200.0, mult_mean2=1500.0, mult_frac2=0.2,
                     mult_cap=4000)
    other = oc.generate(geo, 100, seed=0xB8, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    sess.set_pruner_margin(meta["margin"])
    sess.set_pruner(flat, thr)
    try:
        ref = sess.reconstruct(ev, records=True)
        din = sess.upload(ev)
        dout = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True, records=True)
        for rep in range(3):
            for k in ("status", "n_tracks", "tracks", "n_tracklets", "records", "keep_mask"):
                dout[k].fill_(0x5A if k != "records" else 7)
            sess.set_tuning(oc.TUNE_GRAPH, 1)
            sess.run_device(din, dout)
            torch.cuda.synchronize()
            got = oc.Session.fetch(dout, ev.n_events, sess.params.max_tracks_out, sess.params.max_tracklets_per_event)
            same_outputs(got, ref)
            sess.set_tuning(oc.TUNE_GRAPH, 0)
            sess.reconstruct(other)            # eager work between replays
    finally:
        sess.set_tuning(oc.TUNE_GRAPH, 0)
        sess.set_pruner(None)
        sess.set_pruner_margin()
    assert ref.n_tracks.sum() > 500
    sess.close()


def test_shipped_pruner_recall(oc, geo):
    # the learned-pruning configuration: hits Poisson(1200), high-occupancy caps,
    # events of a seed the training never saw
    import pruner
    if not os.path.exists(oc.DEFAULT_PRUNER):
        pytest.skip("no shipped pruner weights")
    flat, thr, meta = pruner.load()
    sess = oc.Session(geo, oc.high_occupancy_params(), device=0)
    ev = oc.generate(geo, 6000, seed=0x5EED, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    plain = sess.reconstruct(ev)
    sess.set_pruner_margin(meta["margin"])
    sess.set_pruner(flat, thr)
    try:
        pr = sess.reconstruct(ev)
    finally:
        sess.set_pruner(None)
        sess.set_pruner_margin()
    rec, found, total = pruner.recall(plain, pr)
    assert total > 50000
    assert rec >= 0.995, f"track recall {rec:.5f} ({found}/{total})"
    sess.close()
