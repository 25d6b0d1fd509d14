"""Pins the oracle's event reducer (reconstruction_helper.cuh:168-361).

Two kinds of checks:
 * hand-worked cases, one per rule of the reference reducer, with the kept
   set written down from reading the reference;
 * many random small planes against a second restatement written here in
   plain Python from the same reference lines, so a slip in either shows up.
"""
import ctypes as C

import numpy as np
import pytest


def run(orc, det, chan, pos=None, tdc=None, drift=None, flag=None):
    n = len(chan)
    chan = np.asarray(chan, np.float32)
    pos = np.asarray(pos if pos is not None else chan * 2.0, np.float32)
    tdc = np.asarray(tdc if tdc is not None else np.arange(n) * 50.0, np.float32)
    drift = np.asarray(drift if drift is not None else np.full(n, 0.1), np.float32)
    flag = np.asarray(flag if flag is not None else np.full(n, 2.0), np.float32)
    keep = np.zeros(max(n, 1), np.int8)
    r = orc.load().orc_reduce_plane(det, n, chan.ctypes.data, pos.ctypes.data, tdc.ctypes.data, flag.ctypes.data,
                                    drift.ctypes.data, keep.ctypes.data)
    return r, [i for i in range(n) if keep[i] > 0]


# ---- hand-worked cases -----------------------------------------------------------

def test_out_of_time_dropped(orc):
    # flag bit 1 (value 2) means in time (online_reconstruction.cu:599); flag 1 is out of time
    assert run(orc, 40, [3, 7], flag=[1, 2]) == (1, [1])
    assert run(orc, 40, [3, 7], flag=[3, 2]) == (2, [0, 1])


def test_first_hit_per_channel_wins_hodoscope(orc):
    # no declustering outside chambers; second pass drops repeated channels
    assert run(orc, 35, [4, 4, 5, 4]) == (2, [0, 2])


def test_isolated_chamber_hits_kept(orc):
    assert run(orc, 16, [3, 10, 20, 40]) == (4, [0, 1, 2, 3])


def test_two_hit_cluster_drops_larger_drift(orc):
    # pos step 2.0 -> w_max = 0.9, w_min = 0.4 (reconstruction_helper.cuh:267-276)
    # then a far hit closes the cluster
    assert run(orc, 16, [5, 6, 30], pos=[10, 12, 60], drift=[1.0, 0.5, 0.1]) == (2, [1, 2])
    assert run(orc, 16, [5, 6, 30], pos=[10, 12, 60], drift=[0.5, 1.0, 0.1]) == (2, [0, 2])
    # equal drifts above both thresholds: the second (last) is dropped
    assert run(orc, 16, [5, 6, 30], pos=[10, 12, 60], drift=[1.0, 1.0, 0.1]) == (2, [0, 2])


def test_two_hit_cluster_time_rule_only_d3p(orc):
    # small drifts: neither drift condition; |dt| < 8 drops both, for detectors 19-24 only
    kw = dict(pos=[10, 12, 60], drift=[0.3, 0.3, 0.1], tdc=[100.0, 105.0, 500.0])
    assert run(orc, 20, [5, 6, 30], **kw) == (1, [2])
    assert run(orc, 16, [5, 6, 30], **kw) == (3, [0, 1, 2])
    kw["tdc"] = [100.0, 110.0, 500.0]
    assert run(orc, 20, [5, 6, 30], **kw) == (3, [0, 1, 2])


def test_three_hit_cluster(orc):
    # mean |dt| < 10 drops all; otherwise all but the first (reconstruction_helper.cuh:300-318)
    assert run(orc, 16, [5, 6, 7, 30], tdc=[100, 104, 108, 900]) == (1, [3])
    assert run(orc, 16, [5, 6, 7, 30], tdc=[100, 130, 160, 900]) == (2, [0, 3])


def test_closing_hit_not_in_next_cluster(orc):
    # 10 closes {5,6} and is not added to a new cluster; 11 then starts one and
    # is the last hit, so {10, 11} is never examined: drifts that would drop
    # one of them leave both in place
    r = run(orc, 16, [5, 6, 10, 11], pos=[10, 12, 20, 22], drift=[0.1, 0.1, 1.0, 0.5], tdc=[0, 100, 200, 300])
    assert r == (4, [0, 1, 2, 3])


def test_last_hit_joins_cluster(orc):
    # the last hit is appended when adjacent, then the cluster is processed
    assert run(orc, 16, [5, 6], pos=[10, 12], drift=[1.0, 0.5]) == (1, [1])


def test_after_pulse_pair(orc):
    # same channel twice: a 2-hit cluster with w_max = 0, the larger drift goes
    assert run(orc, 22, [8, 8, 40], pos=[16, 16, 80], drift=[0.2, 0.6, 0.1], tdc=[100, 180, 500]) == (2, [0, 2])


def test_out_of_time_hit_inside_cluster_is_skipped(orc):
    # the out-of-time hit neither joins nor closes the cluster
    r = run(orc, 16, [5, 6, 6, 30], pos=[10, 12, 12, 60], drift=[1.0, 0.1, 0.5, 0.1], flag=[2, 1, 2, 2])
    assert r == (2, [2, 3])


def test_unsorted_channels_follow_array_order(orc):
    # clusters follow array order: 3 closes {7} and is dropped from clustering,
    # 4 starts {4, 5}, which 50 closes; the 2-hit rule then drops 4 (larger drift)
    assert run(orc, 16, [7, 3, 4, 5, 50], pos=[14, 6, 8, 10, 100], drift=[0.1, 1.0, 1.0, 0.5, 0.1]) == (4, [0, 1, 3, 4])


def test_cluster_overflow_flagged(orc):
    # a 200-hit run of adjacent channels exceeds the 192-entry cluster array
    n = 200
    r, _ = run(orc, 16, list(range(1, n + 1)), tdc=np.arange(n) * 30.0)
    assert r == -1


# ---- second restatement, exhaustive-ish comparison ------------------------------

def reduce_ref(det, chan, pos, tdc, drift, flag):
    """Plain-Python reading of reconstruction_helper.cuh:168-361 (float32 where the reference is float)."""
    f32 = np.float32
    n = len(chan)
    keep = [1] * n
    cl = []
    for i in range(n):
        keep[i] = 1
        if (int(flag[i]) & 2) == 0:
            keep[i] = -1
            continue
        if det > 30:
            continue
        if not cl:
            cl.append(i)
            continue
        far = abs(f32(chan[i]) - f32(chan[cl[-1]])) > 1
        if far or i == n - 1:
            if i == n - 1 and not far:
                cl.append(i)
            if len(cl) == 2:
                a, b = cl
                wmax = f32(0.9 * 0.5 * float(f32(pos[b]) - f32(pos[a])))
                wmin = f32(4.0 / 9.0 * float(wmax))
                if (drift[a] > wmax and drift[b] > wmin) or (drift[a] > wmin and drift[b] > wmax):
                    keep[a if drift[a] > drift[b] else b] = -1
                else:
                    dt = f32(f32(tdc[a]) - f32(tdc[b]))
                    if ((0 <= dt < 8) or (-8 < dt <= 0)) and 19 <= det <= 24:
                        keep[a] = keep[b] = -1
            if len(cl) >= 3:
                s = f32(0)
                for c in range(1, len(cl)):
                    s = f32(s + abs(f32(f32(tdc[cl[c]]) - f32(tdc[cl[c - 1]]))))
                s = f32(s / f32(len(cl) - 1))
                for c in range(len(cl)):
                    if s < 10 or c > 0:
                        keep[cl[c]] = -1
            cl = []
        else:
            cl.append(i)
    seen = set()
    for i in range(n):
        if keep[i] > 0:
            if int(chan[i]) in seen:
                keep[i] = -1
            else:
                seen.add(int(chan[i]))
    return [i for i in range(n) if keep[i] > 0]


@pytest.mark.parametrize("det", [3, 16, 20, 24, 28, 40, 50])
def test_random_planes_match_second_restatement(orc, det):
    rng = np.random.default_rng(det)
    for _ in range(1500):
        n = int(rng.integers(1, 8))
        chan = np.sort(rng.integers(1, 9, n)).astype(np.float32) if rng.random() < 0.7 else \
            rng.integers(1, 9, n).astype(np.float32)
        pos = (chan * 2.0 + rng.normal(0, 0.01, n)).astype(np.float32)
        tdc = rng.choice([100.0, 103.0, 109.0, 120.0, 160.0], n).astype(np.float32) + rng.integers(0, 3, n)
        drift = rng.uniform(0, 1.2, n).astype(np.float32)
        flag = rng.choice([2.0, 2.0, 2.0, 1.0], n).astype(np.float32)
        r, kept = run(orc, det, chan, pos, tdc, drift, flag)
        assert kept == reduce_ref(det, chan, pos, tdc.astype(np.float32), drift, flag), (det, chan, pos, tdc, drift, flag)
        assert r == len(kept)
