"""Dimuons and display histograms on the device (or_dimuons) against the oracle
(reconstruction_kernels.cuh:2820-3195 restated in oracle/or_oracle_dimuon.cpp),
on the full chain's vertexed tracks: the dimuon records bit-exact (phi within
one ulp: both sides round an FP64 atan2 to float, and the two FP64 libraries
may differ in their last place), the per-event counts and the histograms equal.
Events are generated with two opposite-charge muons so that most have a pair."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full(oc, geo):
    p = oc.default_params(max_hits_per_event=2048)
    p.full_chain = 1
    return oc.Session(geo, p, device=0)


def run_both(oc, orc, geo, ogeo, sess, ev, trig):
    import torch
    t = sess.upload(ev)
    o = sess.alloc_outputs(ev.n_events, ev.n_hits, vertex=True)
    sess.run_device(t, o)
    d = sess.alloc_dimuons(ev.n_events, histograms=True)
    sess.run_dimuons(t, o, d, torch.from_numpy(trig).to(t["det"].device))
    torch.cuda.synchronize()
    chain = oc.Session.fetch(o, ev.n_events, sess.params.max_tracks_out, sess.params.max_tracklets_per_event)
    got = oc.Session.fetch_dimuons(d, ev.n_events)
    ref_chain = orc.reconstruct(ogeo, sess.params, ev, vertex=True)
    return chain, ref_chain, got


def check(got, ref):
    assert np.array_equal(got.n_dimuons, ref.n_dimuons)
    n = got.n_dimuons
    rows = np.arange(got.dimuons.shape[1])[None, :] < n[:, None]
    g, r = got.dimuons[rows], ref.dimuons[rows]
    other = np.ones(g.shape[1], bool)
    other[6] = False
    assert np.array_equal(g[:, other].view(np.uint32), r[:, other].view(np.uint32))
    dphi = np.abs(g[:, 6].view(np.int32).astype(np.int64) - r[:, 6].view(np.int32).astype(np.int64))
    assert dphi.max(initial=0) <= 1
    assert np.array_equal(got.histograms, ref.histograms)
    assert got.lost == ref.lost


def test_dimuons_against_oracle(oc, orc, geo, ogeo, full):
    ev = oc.generate(geo, 4000, seed=77, mult_mode=1, mult_mean=150.0, mult_cap=600, n_tracks_min=2, n_tracks_max=3)
    trig = np.random.default_rng(3).integers(0, 1024, ev.n_events).astype(np.int32)
    chain, ref_chain, got = run_both(oc, orc, geo, ogeo, full, ev, trig)
    assert np.array_equal(chain.valid_tracks().view(np.uint8), ref_chain.valid_tracks().view(np.uint8))
    ref = orc.dimuons(ref_chain, ev, geo.export()[0][0], full.params.max_tracks_out, trig)
    check(got, ref)
    assert got.n_dimuons.sum() > 200
    assert got.histograms.sum() > 0


def test_dimuon_histograms_accumulate(oc, orc, geo, ogeo, full):
    import torch
    ev = oc.generate(geo, 500, seed=78, mult_mode=1, mult_mean=100.0, mult_cap=400, n_tracks_min=2, n_tracks_max=3)
    t = full.upload(ev)
    o = full.alloc_outputs(ev.n_events, ev.n_hits, vertex=True)
    full.run_device(t, o)
    d = full.alloc_dimuons(ev.n_events, histograms=True)
    full.run_dimuons(t, o, d, None)
    full.run_dimuons(t, o, d, None)
    torch.cuda.synchronize()
    two = oc.Session.fetch_dimuons(d, ev.n_events)
    d1 = full.alloc_dimuons(ev.n_events, histograms=True)
    full.run_dimuons(t, o, d1, None)
    torch.cuda.synchronize()
    one = oc.Session.fetch_dimuons(d1, ev.n_events)
    assert np.array_equal(two.histograms, 2 * one.histograms)
    # trigger bits NULL: only the last ("all") trigger row counts
    h = one.histograms.reshape(11, 18, 256)
    assert h[:10].sum() == 0 and h[10].sum() > 0
