"""The oracle's YZ unit filter (test aid for the pruned chain).

reconstruct(units=keys) restricts the YZ search of each listed tracklet to its
listed (u3, v3) units and leaves unlisted tracklets alone.  Pinned here by
what the filter must leave unchanged (keys that name no tracklet of the batch)
and by what it must change (a listed tracklet whose only listed unit does not
exist has no YZ winner, so no global track can come from it).  That the unit
numbering is the device's is pinned on the GPU, where the pruned chain's
searched units fed through the filter give the device's outputs bit for bit
(tests/test_gpu_pruner.py).
"""
import numpy as np


def same(a, b):
    return (np.array_equal(a.status, b.status) and np.array_equal(a.n_tracklets, b.n_tracklets)
            and np.array_equal(a.records.view(np.uint32), b.records.view(np.uint32))
            and np.array_equal(a.n_tracks, b.n_tracks)
            and np.array_equal(a.valid_tracks().view(np.uint8), b.valid_tracks().view(np.uint8)))


def test_filter_naming_no_tracklet_changes_nothing(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 120, seed=31, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    p = oc.default_params(max_hits_per_event=1024)
    ref = orc.reconstruct(ogeo, p, ev, records=True)
    # event indices past the batch: the filter is on but names none of its tracklets
    keys = orc.unit_keys(np.arange(500, 600), np.zeros(100, np.int64), np.zeros(100, np.int64))
    got = orc.reconstruct(ogeo, p, ev, records=True, units=keys)
    assert ref.n_tracks.sum() > 40
    assert same(got, ref)
    # and the filter is off again afterwards
    assert same(orc.reconstruct(ogeo, p, ev, records=True), ref)


def test_filter_without_existing_unit_removes_the_tracklets_winners(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 120, seed=32, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    p = oc.default_params(max_hits_per_event=1024)
    ref = orc.reconstruct(ogeo, p, ev, records=True)
    with_tracks = np.nonzero(ref.n_tracks > 0)[0]
    assert len(with_tracks) > 20
    cut = with_tracks[::2]
    # every tracklet of the chosen events (tracklet ids 0 .. n - 1) listed with
    # one unit index that no tracklet has
    ev_idx, tk_idx = [], []
    for e in cut:
        for t in range(int(ref.n_tracklets[e])):
            ev_idx.append(e)
            tk_idx.append(t)
    keys = orc.unit_keys(ev_idx, tk_idx, np.full(len(ev_idx), 0xFFFFF, np.int64))
    got = orc.reconstruct(ogeo, p, ev, records=True, units=keys)
    # the XZ stage is untouched, YZ finds nothing in the chosen events, the others are as before
    assert np.array_equal(got.n_tracklets, ref.n_tracklets)
    assert (got.n_tracks[cut] == 0).all()
    rest = np.setdiff1d(np.arange(ev.n_events), cut)
    assert np.array_equal(got.n_tracks[rest], ref.n_tracks[rest])
    assert np.array_equal(got.tracks[rest].view(np.uint8), ref.tracks[rest].view(np.uint8))
