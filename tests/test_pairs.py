"""Binned X/X' pair lists (make_hitpairs_in_station_bins,
reconstruction_helper.cuh:368-460) against a brute-force enumeration."""
import numpy as np
import pytest

LO = np.array([1 + 4 * b for b in range(28)])
HI = np.array([4 + 4 * b for b in range(28)])


def brute(stid, chan1, pos1, chan2, pos2, spacing, cap):
    hi = HI.copy()
    if stid > 2:
        hi[27] = 116
    lists = [[] for _ in range(28)]
    used1, used2 = set(), set()

    def put(c, a, b):
        for k in range(28):
            if LO[k] <= c <= hi[k]:
                lists[k].append((a, b))

    for i in range(len(chan1)):
        for j in range(len(chan2)):
            if abs(np.float32(pos1[i]) - np.float32(pos2[j])) <= np.float32(spacing):
                put(chan1[i], i, j)
                used1.add(i)
                used2.add(j)
    for i in range(len(chan1)):
        if i not in used1:
            put(chan1[i], i, -1)
    for j in range(len(chan2)):
        if j not in used2:
            put(chan2[j], -1, j)
    return [l[:cap] for l in lists], [len(l) for l in lists]


@pytest.mark.parametrize("stid,spacing", [(2, 1.3), (3, 1.2), (4, 1.2)])
def test_binned_pairs_brute_force(orc, stid, spacing):
    rng = np.random.default_rng(stid)
    cap = 20
    lib = orc.load()
    for _ in range(300):
        n1, n2 = int(rng.integers(0, 9)), int(rng.integers(0, 9))
        c1 = rng.integers(1, 117 if stid > 2 else 113, n1).astype(np.float32)
        c2 = rng.integers(1, 117 if stid > 2 else 113, n2).astype(np.float32)
        if rng.random() < 0.5:  # crowd a few channels so bins overflow the cap sometimes
            c1 = rng.integers(40, 44, n1).astype(np.float32)
            c2 = rng.integers(40, 44, n2).astype(np.float32)
        p1 = (c1 * 2.0 + rng.normal(0, 0.3, n1)).astype(np.float32)
        p2 = (c2 * 2.0 - 1.0 + rng.normal(0, 0.3, n2)).astype(np.float32)
        out = np.zeros(28 * cap * 2, np.int32)
        cnt = np.zeros(28, np.int32)
        lib.orc_pairs_binned(stid, n1, c1.ctypes.data, p1.ctypes.data, n2, c2.ctypes.data, p2.ctypes.data, cap,
                             out.ctypes.data, cnt.ctypes.data)
        want, wcnt = brute(stid, c1, p1, c2, p2, spacing, cap)
        assert list(cnt) == wcnt
        got = out.reshape(28, cap, 2)
        for b in range(28):
            assert [tuple(x) for x in got[b, :min(cnt[b], cap)]] == want[b]
