"""HIP chain against the CPU oracle at high occupancy, with the reference caps
lifted (orchain.HIGH_OCCUPANCY_CAPS), bit for bit through the C ABI.

Workloads: hits per event Poisson(1200) (the learned-pruning configuration),
the 0.8 Poisson(200) + 0.2 Poisson(1500) mixture capped at 4000 (the batched
configuration) and Poisson(2000) capped at 6000 with 1-4 tracks and the full
chain (the high-intensity configuration).  At these occupancies the YZ
survivor lists of a chunk run in several batches; a small batch size forces
many batch cuts and the unit-by-unit path."""
import numpy as np
import pytest

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def _run(oc, orc, geo, ogeo, ev, params, records=True, list_cap=None, vertex=False):
    sess = oc.Session(geo, params, device=0)
    if list_cap:
        sess.set_tuning(oc.TUNE_YZ_LIST_CAP, list_cap)
    got = sess.reconstruct(ev, records=records, vertex=vertex)
    ref = orc.reconstruct(ogeo, params, ev, records=records, vertex=vertex)
    sess.close()
    return got, ref


def test_c3_mean1200_lifted_caps(oc, orc, geo, ogeo):
    print("case: 200 events, hits Poisson(1200), high-occupancy caps", flush=True)
    ev = oc.generate(geo, 200, seed=0xC3, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    params = oc.high_occupancy_params()
    got, ref = _run(oc, orc, geo, ogeo, ev, params)
    rejected = ((ref.status & oc.EV_REJECT_MASK) != 0).mean()
    assert rejected < 0.10, f"{rejected:.2%} of events rejected"
    assert ref.n_tracklets.mean() > 300 and ref.n_tracks.sum() > 1000
    assert_same(got, ref, ev, "c3 mean 1200")


@pytest.mark.parametrize("cap", [64, 333])
def test_yz_batches(oc, orc, geo, ogeo, cap):
    # survivor batches far smaller than a chunk's survivors: many cuts at unit
    # boundaries, and units larger than the batch (unit-by-unit path)
    print(f"case: 40 events, hits Poisson(1200), YZ batch {cap}", flush=True)
    ev = oc.generate(geo, 40, seed=0xBA7C + cap, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    got, ref = _run(oc, orc, geo, ogeo, ev, oc.high_occupancy_params(), list_cap=cap)
    assert_same(got, ref, ev, f"yz batch {cap}")


def test_c2_mixture(oc, orc, geo, ogeo):
    print("case: 200 events, 0.8 Poisson(200) + 0.2 Poisson(1500), cap 4000", flush=True)
    ev = oc.generate(geo, 200, seed=0xC2, mult_mode=2, mult_mean=200.0, mult_mean2=1500.0, mult_frac2=0.2,
                     mult_cap=4000)
    counts = np.diff(ev.offsets)
    assert (counts > 1024).sum() >= 20 and (counts <= 512).sum() >= 100   # both tile classes populated
    got, ref = _run(oc, orc, geo, ogeo, ev, oc.high_occupancy_params())
    assert ((ref.status & oc.EV_REJECT_MASK) != 0).mean() < 0.10
    assert_same(got, ref, ev, "c2 mixture")


def test_c4_mean2000_full_chain(oc, orc, geo, ogeo):
    print("case: 16 events, hits Poisson(2000) cap 6000, 1-4 tracks, full chain", flush=True)
    ev = oc.generate(geo, 16, seed=0xC4, mult_mode=1, mult_mean=2000.0, mult_cap=6000, n_tracks_min=1,
                     n_tracks_max=4)
    params = oc.high_occupancy_params(max_tracklets_per_event=4096, full_chain=1)
    got, ref = _run(oc, orc, geo, ogeo, ev, params, vertex=True)
    assert ((ref.status & oc.EV_REJECT_MASK) != 0).mean() < 0.10
    assert ref.n_tracklets.mean() > 1000
    assert_same(got, ref, ev, "c4 mean 2000")
    for e in range(ev.n_events):
        n = ref.n_tracks[e]
        assert np.array_equal(got.vertex[e, :n].view(np.uint32), ref.vertex[e, :n].view(np.uint32)), f"vertex {e}"


def test_geometry_conflict(oc, geo, tmp_path):
    # the plane table is per device: while a session is live, a session on a
    # different geometry is refused and the live session's results do not change
    sess = oc.Session(geo, oc.default_params(), device=0)
    ev = oc.generate(geo, 200, seed=21)
    before = sess.reconstruct(ev)
    with open(oc.DEFAULT_GEOMETRY) as f:
        lines = f.read().splitlines()
    for i, ln in enumerate(lines):
        tok = ln.split()
        if tok and tok[0] == "16":
            tok[1] = repr(float(tok[1]) + 1.0)
            lines[i] = "\t".join(tok)
            break
    path = tmp_path / "geometry_shifted.txt"
    path.write_text("\n".join(lines) + "\n")
    other = oc.Geometry.load(str(path))
    with pytest.raises(oc.OrError):
        oc.Session(other, oc.default_params(), device=0)
    after = sess.reconstruct(ev)
    assert np.array_equal(before.status, after.status)
    assert np.array_equal(before.n_tracks, after.n_tracks)
    assert np.array_equal(before.valid_tracks().view(np.uint8), after.valid_tracks().view(np.uint8))
    sess.close()


@pytest.mark.parametrize("mean,cap,n", [(800.0, 1000, 30), (1200.0, 4000, 30), (2500.0, 6000, 8), (4300.0, 8000, 2)])
def test_waves_per_event(oc, orc, geo, ogeo, mean, cap, n):
    # an event's YZ chunks split over 1, 2, 4 or 8 waves (tile classes 1024,
    # 2048 at most 4, 4096 and 8192 hits at most 8), with full and small
    # survivor batches: the oracle's result at each
    print(f"case: {n} events, hits Poisson({mean:g}) cap {cap}, waves per event 1/2/4/8", flush=True)
    ev = oc.generate(geo, n, seed=0x3E + int(mean), mult_mode=1, mult_mean=mean, mult_cap=cap)
    params = oc.high_occupancy_params(max_tracklets_per_event=4096)
    ref = orc.reconstruct(ogeo, params, ev, records=True)
    assert ref.n_tracklets.sum() > 100 * n
    sess = oc.Session(geo, params, device=0)
    for wpe in (1, 2, 4, 8):
        sess.set_tuning(oc.TUNE_WAVES_PER_EVENT, wpe)
        assert sess.get_tuning(oc.TUNE_WAVES_PER_EVENT) == wpe
        for batch in (4096, 150):
            sess.set_tuning(oc.TUNE_YZ_LIST_CAP, batch)
            got = sess.reconstruct(ev, records=True)
            assert_same(got, ref, ev, f"{wpe} waves per event, batch {batch}")
    sess.close()


def test_waves_per_event_pruned(oc, geo):
    # learned pruning gives the same tracks whichever wave takes a chunk
    import pruner
    flat, thr, meta = pruner.load()
    print("case: 300 events, hits Poisson(1200), pruned, waves per event 1 vs 4", flush=True)
    ev = oc.generate(geo, 300, seed=0x3F, mult_mode=1, mult_mean=1200.0, mult_cap=4000)
    sess = oc.Session(geo, oc.high_occupancy_params(), device=0)
    sess.set_pruner_margin(meta["margin"])
    sess.set_pruner(flat, thr)
    outs = []
    for wpe in (1, 4, 2):
        sess.set_tuning(oc.TUNE_WAVES_PER_EVENT, wpe)
        outs.append(sess.reconstruct(ev, records=True))
    for o in outs[1:]:
        assert_same(o, outs[0], ev, "pruned, waves per event")
    sess.close()
