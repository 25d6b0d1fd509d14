"""Golden fixture (tests/golden/c0_small.npz, written by make_golden.py): the
oracle must keep reproducing it, and the HIP chain must match it bit for bit."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden(oc):
    z = np.load(os.path.join(HERE, "golden", "c0_small.npz"))
    ev = oc.Events(z["det"], z["elem"], z["tdc"], z["drift"], z["flag"], z["offsets"])
    return ev, z


def check(oc, res, ev, z):
    assert np.array_equal(res.status, z["status"])
    assert np.array_equal(res.keep_mask, z["keep_mask"])
    assert np.array_equal(res.n_tracklets, z["n_tracklets"])
    ntk = z["n_tracklets"]
    rec = np.concatenate([res.records[e, :ntk[e]] for e in range(ev.n_events)])
    assert np.array_equal(rec.view(np.uint32), z["records"].view(np.uint32))
    assert np.array_equal(res.n_tracks, z["n_tracks"])
    # slots beyond n_tracks are not written by the library
    zt = z["tracks"].reshape(ev.n_events, res.tracks.shape[1], -1)
    k = np.arange(res.tracks.shape[1])[None, :] < z["n_tracks"][:, None]
    got = np.ascontiguousarray(res.tracks).view(np.uint8).reshape(ev.n_events, res.tracks.shape[1], -1)
    same = np.array_equal(got[k], zt[k])
    assert same


def test_oracle_reproduces_golden(oc, orc, ogeo, golden):
    ev, z = golden
    res = orc.reconstruct(ogeo, oc.default_params(), ev, records=True)
    check(oc, res, ev, z)


@pytest.mark.gpu
def test_hip_chain_matches_golden(oc, geo, golden):
    ev, z = golden
    sess = oc.Session(geo, oc.default_params(), device=0)
    res = sess.reconstruct(ev, records=True)
    check(oc, res, ev, z)


def check_vertex(res, z):
    k = np.arange(res.tracks.shape[1])[None, :] < res.n_tracks[:, None]
    assert (res.tracks[k]["station_id"] == 7).all()
    same = np.array_equal(res.vertex[k].view(np.uint32), z["vertex"].view(np.uint32))
    assert same


def test_oracle_reproduces_golden_vertex(oc, orc, ogeo, golden):
    ev, z = golden
    check_vertex(orc.reconstruct(ogeo, oc.default_params(), ev, vertex=True), z)


@pytest.mark.gpu
def test_hip_vertex_matches_golden(oc, geo, golden):
    ev, z = golden
    sess = oc.Session(geo, oc.default_params(), device=0)
    check_vertex(sess.reconstruct(ev, vertex=True), z)

