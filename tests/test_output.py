"""The columnar event writer (or_output_*, the reference's OROutput tree,
OROutput.cxx:24-102, filled as online_reconstruction.cu:1063-1222): a round
trip from chain results through the .npz archive, branch by branch, against
the results and inputs it was given.  Host code only; the results come from
the oracle (full chain with vertexing, and its dimuons)."""
import zipfile

import numpy as np
import pytest

# OROutput.cxx:28-101, in order
REF_BRANCHES = (
    ["evtid", "nhits", "hit.detid", "hit.chan", "hit.pos", "hit.tdc", "hit.drift", "ntracks", "track.stid",
     "track.nhits", "track.chi2", "track.x0", "track.y0", "track.tx", "track.ty", "track.invp", "track.err_x0",
     "track.err_y0", "track.err_tx", "track.err_ty", "track.err_invp", "track.charge", "track.vx", "track.vy",
     "track.vz", "track.px", "track.py", "track.pz", "track.hitdetid", "track.hitchan", "track.hitpos",
     "track.hitdrift", "track.hitsign", "track.hittdc", "track.hitresid", "track.invp_tgt", "ndimuons"]
    + ["dimuon." + n for n in ("mass", "pT", "xF", "x1", "x2", "costheta", "phi", "mass_single", "chi2_single",
                               "vx", "vy", "vz", "px", "py", "pz", "pos_vx", "pos_vy", "pos_vz", "neg_vx",
                               "neg_vy", "neg_vz", "pos_e", "pos_px", "pos_py", "pos_pz", "neg_e", "neg_px",
                               "neg_py", "neg_pz", "pos_single_e", "pos_single_px", "pos_single_py",
                               "pos_single_pz", "neg_single_e", "neg_single_px", "neg_single_py",
                               "neg_single_pz", "chi2_vtx", "chi2_kf")])
INT_BRANCHES = {"evtid", "nhits", "hit.detid", "hit.chan", "ntracks", "track.stid", "track.nhits", "track.charge",
                "track.hitdetid", "track.hitchan", "track.hitsign", "ndimuons"}
MULT_BITS = 0xF   # OR_EV_D0_MULT .. OR_EV_D3M_MULT


@pytest.fixture(scope="module")
def case(oc, orc, geo, ogeo):
    p = oc.default_params(max_hits_per_event=2048)
    p.full_chain = 1
    ev = oc.generate(geo, 400, seed=0x0A7, mult_mode=1, mult_mean=150.0, mult_cap=600, n_tracks_min=2,
                     n_tracks_max=3)
    # two events without hits (not written), and a station-1 multiplicity cut
    # that some events breach (not written either)
    off = np.insert(ev.offsets, [5, 200], [ev.offsets[5], ev.offsets[200]])
    ev = oc.Events(ev.det, ev.elem, ev.tdc, ev.drift, ev.flag, off)
    d0 = np.add.reduceat((ev.det >= 1) & (ev.det <= 6), ev.offsets[:-1])
    p.max_d0_mult = int(np.quantile(d0, 0.8))
    chain = orc.reconstruct(ogeo, p, ev, records=True, vertex=True, threads=8)
    assert ((chain.status & MULT_BITS) != 0).sum() >= 10
    d = orc.dimuons(chain, ev, geo.export()[0][0], p.max_tracks_out, None)
    return p, ev, chain, d


def write(oc, geo, p, ev, chain, d, path, split=None, **kw):
    with oc.OutputWriter(path) as w:
        if split is None:
            w.fill(geo, ev, chain, p.max_tracklets_per_event, d, **kw)
        else:
            for a, b in ((0, split), (split, ev.n_events)):
                sub = ev.subset(a, b)
                sl = slice(a, b)
                res = oc.Result(chain.status[sl], sub_keep(chain, ev, a, b), chain.n_tracks[sl], chain.tracks[sl],
                                chain.n_tracklets[sl], chain.records[sl], chain.vertex[sl])
                dd = oc.DimuonResult(d.n_dimuons[sl], d.dimuons[sl], None, 0)
                eid = None if kw.get("event_id") is None else kw["event_id"][sl]
                w.fill(geo, sub, res, p.max_tracklets_per_event, dd, event_id=eid,
                       min_station_id=kw.get("min_station_id", 7))
    return np.load(path)


def sub_keep(chain, ev, a, b):
    bits = np.unpackbits(chain.keep_mask.view(np.uint8), bitorder="little")[ev.offsets[a]:ev.offsets[b]]
    return np.packbits(np.concatenate([bits, np.zeros((-len(bits)) % 32, np.uint8)]), bitorder="little").view(np.uint32)


def expected(oc, geo, p, ev, chain, d, event_id, min_stid=7):
    _, wire = geo.export()
    keep = chain.keep_bits(ev.n_hits)
    E = {k: [] for k in REF_BRANCHES + ["status"]}
    for e in range(ev.n_events):
        h0, h1 = ev.offsets[e], ev.offsets[e + 1]
        if h1 == h0 or (chain.status[e] & MULT_BITS):
            continue
        E["evtid"].append(event_id[e])
        idx = [h for h in range(h0, h1) if keep[h] and 1 <= ev.det[h] <= 54]
        idx.sort(key=lambda h: (ev.det[h], h))   # detector by detector, staging order inside
        E["nhits"].append(len(idx))
        for h in idx:
            E["hit.detid"].append(ev.det[h]); E["hit.chan"].append(ev.elem[h])
            E["hit.pos"].append(wire[ev.det[h], ev.elem[h]]); E["hit.tdc"].append(ev.tdc[h])
            E["hit.drift"].append(abs(ev.drift[h]))
        nt = 0
        for k in range(min(chain.n_tracklets[e], p.max_tracklets_per_event)):
            r = chain.records[e, k]
            if int(r[0]) < min_stid:
                continue
            nt += 1
            nh = int(r[2])
            for name, f in (("track.stid", 0), ("track.nhits", 2), ("track.charge", 15)):
                E[name].append(int(r[f]))
            for name, f in (("track.chi2", 3), ("track.x0", 7), ("track.y0", 8), ("track.tx", 5), ("track.ty", 6),
                            ("track.invp", 9), ("track.err_x0", 12), ("track.err_y0", 13), ("track.err_tx", 10),
                            ("track.err_ty", 11), ("track.err_invp", 14), ("track.vx", 16), ("track.vy", 17),
                            ("track.vz", 18), ("track.px", 19), ("track.py", 20), ("track.pz", 21)):
                E[name].append(r[f])
            for name, f in (("track.hitdetid", 22), ("track.hitchan", 40), ("track.hitsign", 94)):
                E[name] += [int(x) for x in r[f:f + nh]]
            for name, f in (("track.hitpos", 58), ("track.hitdrift", 76), ("track.hittdc", 112),
                            ("track.hitresid", 130)):
                E[name] += list(r[f:f + nh])
        E["ntracks"].append(nt)
        nd = d.n_dimuons[e]
        E["ndimuons"].append(nd)
        for k in range(nd):
            r = d.dimuons[e, k]
            for i, n in enumerate(["mass", "pT", "xF", "x1", "x2", "costheta", "phi", "mass_single", "chi2_single",
                                   "vx", "vy", "vz"]):
                E["dimuon." + n].append(r[i])
            E["dimuon.px"].append(np.float32(r[18] + r[22]))
            E["dimuon.py"].append(np.float32(r[19] + r[23]))
            E["dimuon.pz"].append(np.float32(r[20] + r[24]))
            for n, f in (("pos_vx", 12), ("pos_vy", 13), ("pos_vz", 14), ("neg_vx", 15), ("neg_vy", 16),
                         ("neg_vz", 17), ("pos_px", 18), ("pos_py", 19), ("pos_pz", 20), ("pos_e", 21),
                         ("neg_px", 22), ("neg_py", 23), ("neg_pz", 24), ("neg_e", 25), ("pos_single_px", 26),
                         ("pos_single_py", 27), ("pos_single_pz", 28), ("pos_single_e", 29),
                         ("neg_single_px", 30), ("neg_single_py", 31), ("neg_single_pz", 32),
                         ("neg_single_e", 33), ("chi2_vtx", 34), ("chi2_kf", 35)):
                E["dimuon." + n].append(r[f])
        E["status"].append(chain.status[e])
    return E


def check(got, E):
    assert list(got.files) == REF_BRANCHES + ["status"]
    for name in REF_BRANCHES + ["status"]:
        a = got[name]
        want = np.asarray(E[name], np.uint32 if name == "status" else (np.int32 if name in INT_BRANCHES else np.float32))
        assert a.dtype == want.dtype, name
        assert a.shape == want.shape, f"{name}: {a.shape} vs {want.shape}"
        assert np.array_equal(a.view(np.uint32), want.view(np.uint32)), name


def test_round_trip(oc, geo, case, tmp_path):
    p, ev, chain, d = case
    eid = (np.arange(ev.n_events) * 7 + 1000).astype(np.int32)
    got = write(oc, geo, p, ev, chain, d, tmp_path / "out.npz", event_id=eid)
    E = expected(oc, geo, p, ev, chain, d, eid)
    # the case holds what the rules are about
    assert len(E["evtid"]) < ev.n_events and sum(E["ntracks"]) > 300 and sum(E["ndimuons"]) > 20
    check(got, E)
    # counters and vector lengths agree, as a reader of the tree relies on
    assert got["nhits"].sum() == got["hit.detid"].size
    assert got["ntracks"].sum() == got["track.stid"].size
    assert got["track.nhits"].sum() == got["track.hitdetid"].size
    assert got["ndimuons"].sum() == got["dimuon.mass"].size
    assert got["track.invp_tgt"].size == 0
    assert (got["track.stid"] >= 7).all()
    with zipfile.ZipFile(tmp_path / "out.npz") as z:
        assert z.testzip() is None   # every member's CRC-32


def test_batches_append_and_station_threshold(oc, geo, case, tmp_path):
    p, ev, chain, d = case
    eid = np.arange(ev.n_events, dtype=np.int32)
    one = write(oc, geo, p, ev, chain, d, tmp_path / "one.npz")                      # event_id NULL: index
    two = write(oc, geo, p, ev, chain, d, tmp_path / "two.npz", split=157, event_id=eid)
    for name in one.files:
        assert np.array_equal(one[name], two[name]), name
    low = write(oc, geo, p, ev, chain, d, tmp_path / "low.npz", min_station_id=5, event_id=eid)
    check(low, expected(oc, geo, p, ev, chain, d, eid, min_stid=5))
    assert low["ntracks"].sum() > one["ntracks"].sum()


def test_argument_checks(oc, geo, case, tmp_path):
    p, ev, chain, d = case
    w = oc.OutputWriter(tmp_path / "x.npz")
    bare = oc.Result(chain.status, chain.keep_mask, chain.n_tracks, chain.tracks)
    with pytest.raises(ValueError):
        w.fill(geo, ev, bare, p.max_tracklets_per_event)
    with pytest.raises(oc.OrError):
        w.fill(geo, ev, chain, 0)
    w.close()
    assert np.load(tmp_path / "x.npz")["evtid"].size == 0
    with pytest.raises(oc.OrError):
        oc.OutputWriter(tmp_path / "no_such_dir" / "x.npz")
