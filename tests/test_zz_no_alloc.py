"""or_reconstruct_noalloc (the entry point that allocates nothing) against
or_reconstruct, bit for bit, through the C ABI: two calls in a row, the second
with a smaller batch, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _outputs(sess, oc, t, o, n_events):
    import torch
    torch.cuda.synchronize()
    return oc.Session.fetch(o, n_events, sess.params.max_tracks_out, sess.params.max_tracklets_per_event)


def _same(a, b):
    assert np.array_equal(a.status, b.status)
    assert np.array_equal(a.keep_mask, b.keep_mask)
    assert np.array_equal(a.n_tracklets, b.n_tracklets)
    assert np.array_equal(a.n_tracks, b.n_tracks)
    assert np.array_equal(a.valid_tracks().view(np.uint8), b.valid_tracks().view(np.uint8))
    if a.records is not None:
        assert np.array_equal(a.records.view(np.uint32), b.records.view(np.uint32))


def test_noalloc_matches_reconstruct(oc, geo):
    print("case: or_reconstruct_noalloc vs or_reconstruct, 6000 then 700 events (Poisson(300)), then 300 at mean 1200",
          flush=True)
    sess = oc.Session(geo, oc.default_params(max_hits_per_event=2048), device=0)
    for n, seed in [(6000, 31), (700, 32)]:   # the second call is smaller and reuses the array
        ev = oc.generate(geo, n, seed=seed, mult_mode=1, mult_mean=300.0, mult_cap=1000)
        t = sess.upload(ev)
        o_ref = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True, records=True)
        sess.run_device(t, o_ref)
        ref = _outputs(sess, oc, t, o_ref, ev.n_events)
        o = sess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True, records=True)
        sess.run_device_noalloc(t, o)
        got = _outputs(sess, oc, t, o, ev.n_events)
        assert ref.n_tracks.sum() > 0
        _same(got, ref)
    sess.close()
    # high occupancy: larger per-wave scratch, fewer waves in the array
    hsess = oc.Session(geo, oc.high_occupancy_params(), device=0)
    ev = oc.generate(geo, 300, seed=33, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    t = hsess.upload(ev)
    o_ref = hsess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
    hsess.run_device(t, o_ref)
    ref = _outputs(hsess, oc, t, o_ref, ev.n_events)
    o = hsess.alloc_outputs(ev.n_events, ev.n_hits, tracklets=True)
    hsess.run_device_noalloc(t, o)
    _same(_outputs(hsess, oc, t, o, ev.n_events), ref)
    hsess.close()


def test_noalloc_limits(oc, geo):
    sess = oc.Session(geo, oc.default_params(), device=0)
    n = oc.NOALLOC_MAX_EVENTS + 1
    lib = oc.load_library()
    offs = np.zeros(n + 1, np.int64)
    import torch
    d_offs = torch.as_tensor(offs, device="cuda")
    dummy = torch.zeros(16, dtype=torch.int16, device="cuda")
    f = torch.zeros(16, dtype=torch.float32, device="cuda")
    hits = oc.OrHits(n, 0, dummy.data_ptr(), dummy.data_ptr(), f.data_ptr(), f.data_ptr(), None, d_offs.data_ptr(), 1)
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    res = oc.OrResult(st.data_ptr(), st.data_ptr(), st.data_ptr(), st.data_ptr(), None, None, None, None, 1)
    assert lib.or_reconstruct_noalloc(sess.handle, C.byref(hits), C.byref(res)) == 7   # OR_ERR_CAPACITY
    hits.on_device = 0
    assert lib.or_reconstruct_noalloc(sess.handle, C.byref(hits), C.byref(res)) == 1   # host pointers refused
    sess.close()
