"""Prop-tube segment matching and global track cleaning in the oracle
(reconstruction_helper.cuh:1916-2257, 2265-2354), by hand-worked cases and a
brute-force restatement of the segment test, plus whole-chain properties of
the full_chain mode."""
import ctypes as C
import itertools

import numpy as np
import pytest

Z1A, Z1B, Z2A, Z2B = 1900.0, 1910.0, 2300.0, 2310.0


def seg(orc, p1a, p1b, p2a, p2b, slope, icpt, win2=7.62, smax=0.15, imax=150.0):
    arrs = [np.asarray(x, np.float32) for x in (p1a, p1b, p2a, p2b)]
    lib = orc.load()
    return lib.orc_prop_segment(arrs[0].ctypes.data, len(arrs[0]), Z1A, arrs[1].ctypes.data, len(arrs[1]), Z1B,
                                arrs[2].ctypes.data, len(arrs[2]), Z2A, arrs[3].ctypes.data, len(arrs[3]), Z2B,
                                slope, icpt, 5.08, win2, smax, imax) == 1


def brute(p1a, p1b, p2a, p2b, slope, icpt, win2=7.62, smax=0.15, imax=150.0):
    f = np.float32
    for (pa, za), (pb, zb) in itertools.product([(p, Z1A) for p in p1a] + [(p, Z1B) for p in p1b],
                                                [(p, Z2A) for p in p2a] + [(p, Z2B) for p in p2b]):
        pa, pb, za, zb = f(pa), f(pb), f(za), f(zb)
        if not abs(pa - (f(slope) * za + f(icpt))) < f(5.08):
            continue
        if abs(pb - (f(slope) * zb + f(icpt))) > f(win2):
            continue
        a = f((pb - pa) / f(zb - za))
        if abs(a) > f(smax) or abs(a - f(slope)) > f(0.03):
            continue
        b = f(pa - a * za)
        if abs(b) > f(imax):
            continue
        if abs((f(slope) * f(2028.19) + f(icpt)) - (a * f(2028.19) + b)) > f(3.0):
            continue
        return True
    return False


def test_segment_hand_cases(orc):
    s, b = 0.02, 10.0
    x = lambda z: s * z + b
    assert seg(orc, [x(Z1A)], [], [x(Z2A)], [], s, b)                  # hits on the track
    assert seg(orc, [], [x(Z1B)], [], [x(Z2B)], s, b)                  # second planes only
    assert not seg(orc, [x(Z1A) + 5.1], [], [x(Z2A)], [], s, b)        # first hit outside 5.08
    assert not seg(orc, [x(Z1A)], [], [x(Z2A) + 7.7], [], s, b)        # second hit outside 7.62
    assert seg(orc, [x(Z1A)], [], [x(Z2A) + 7.7], [], s, b, win2=10.16)
    # segment slope 0.0315 off the track: |a - slope| > 0.03
    assert not seg(orc, [x(Z1A) - 5.0], [], [x(Z2A) + 7.6], [], s, b)
    assert seg(orc, [x(Z1A) - 5.0], [], [x(Z2A) + 6.0], [], s, b)
    assert not seg(orc, [], [], [x(Z2A)], [], s, b)                    # no first-station hit
    # extrapolation to z = 2028.19 more than 3 cm away while slope and windows pass
    d1, d2 = -5.0, -5.0
    a = ((x(Z2A) + d2) - (x(Z1A) + d1)) / (Z2A - Z1A)
    bb = (x(Z1A) + d1) - a * Z1A
    assert abs(a - s) <= 0.03 and abs((s * 2028.19 + b) - (a * 2028.19 + bb)) > 3.0
    assert not seg(orc, [x(Z1A) + d1], [], [x(Z2A) + d2], [], s, b)


def test_segment_matches_brute_force(orc):
    rng = np.random.default_rng(3)
    agree = passes = 0
    for _ in range(400):
        s, b = rng.uniform(-0.1, 0.1), rng.uniform(-40, 40)
        lists = [list(s * z + b + rng.normal(0, 4, rng.integers(0, 4))) for z in (Z1A, Z1B, Z2A, Z2B)]
        r1 = seg(orc, *lists, s, b)
        r2 = brute(*lists, s, b)
        agree += r1 == r2
        passes += r1
    assert agree == 400 and 50 < passes < 350


def clean(orc, stid, chi2, nhits, invp, charge, slot):
    stid = np.asarray(stid, np.float32).copy()
    arrs = [np.asarray(x, np.float32) for x in (chi2, nhits, invp, charge)]
    sl = np.asarray(slot, np.int32)
    orc.load().orc_global_cleaning(len(stid), stid.ctypes.data, *[a.ctypes.data for a in arrs], sl.ctypes.data)
    return stid.astype(int).tolist()


def test_global_cleaning_hand_cases(orc):
    # one slot; track 0 better (chi2/ndf 1 vs 2); 1/p equal within 1.5 %, same charge: 1 demoted
    assert clean(orc, [6, 6], [10, 20], [15, 15], [0.05, 0.0505], [1, 1], [0, 0]) == [6, 5]
    # opposite charges: both kept
    assert clean(orc, [6, 6], [10, 20], [15, 15], [0.05, 0.0505], [1, -1], [0, 0]) == [6, 6]
    # signed difference: j with much larger 1/p still merged, with much smaller 1/p kept
    assert clean(orc, [6, 6], [10, 20], [15, 15], [0.05, 0.10], [1, 1], [0, 0]) == [6, 5]
    assert clean(orc, [6, 6], [10, 20], [15, 15], [0.05, 0.02], [1, 1], [0, 0]) == [6, 6]
    # track 0 worse: the within-slot else branch compares 1/p(i) with itself -> i demoted
    assert clean(orc, [6, 6], [20, 10], [15, 15], [0.05, 0.02], [1, 1], [0, 0]) == [5, 6]
    # chi2/ndf above 250 demotes, lower stations untouched
    assert clean(orc, [6, 5, 3], [2600, 1, 1], [15, 15, 15], [0.05, 0.05, 0.05], [1, 1, 1], [0, 0, 0]) == [5, 5, 3]
    # across slots: slot 0 against slot 1 keeps the worse track 0 ((0.05 - 0.02) / 0.02 is
    # not below 0.015), but slot 1 against slot 0 then demotes it (signed (0.02 - 0.05) / 0.02 is)
    assert clean(orc, [6, 6], [20, 10], [15, 15], [0.05, 0.02], [1, 1], [0, 1]) == [5, 6]
    assert clean(orc, [6, 6], [20, 10], [15, 15], [0.05, 0.02], [1, -1], [0, 1]) == [6, 6]
    # across slots, better j in another slot demotes i: slot 0's pass runs first
    assert clean(orc, [6, 6], [20, 10], [15, 15], [0.05, 0.0501], [1, 1], [0, 1]) == [5, 6]


def test_full_chain_outputs_subset(oc, orc, geo, ogeo):
    ev = oc.generate(geo, 1500, seed=43, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    a = orc.reconstruct(ogeo, oc.default_params(), ev, records=True)
    b = orc.reconstruct(ogeo, oc.default_params(full_chain=1), ev, records=True, vertex=True)
    assert (b.n_tracks <= a.n_tracks).all() and b.n_tracks.sum() > 0.5 * a.n_tracks.sum()
    ta = a.tracks.copy()
    ta["station_id"] = 7
    for e in np.nonzero(b.n_tracks)[0]:
        ids_a = list(a.tracks[e, : a.n_tracks[e]]["tracklet_id"])
        for t in b.tracks[e, : b.n_tracks[e]]:
            assert t["station_id"] == 7 and t["tracklet_id"] in ids_a
            ref = ta[e, ids_a.index(t["tracklet_id"])]
            assert ref.tobytes() == t.tobytes()
        # every record reaching stationID 7 is an output track
        assert (b.records[e, : b.n_tracklets[e], 0] == 7).sum() == b.n_tracks[e]
