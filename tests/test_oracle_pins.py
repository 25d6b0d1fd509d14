"""Pins of the oracle's quirk-laden functions, independent of its code:

* left/right resolution (reconstruction_helper.cuh:714-892): hand-worked hit
  pairs from a known line (hypothesis order, indexmin > 0, the threshold, the
  step back by one hit, signs already set) and a float64 enumeration of the
  four hypotheses on random pairs, for both the x-only variant (no slope or
  intercept limits) and the new-hits variant (limits applied); the single-hit
  rule;
* the order in which YZ and station-1 candidates are enumerated and the winner
  rule, from a trace of every candidate the oracle evaluates
  (reconstruction_kernels.cuh:1320-1322, 1461-1463, 1778; 2107, 2216-2218, 2470);
* the accumulation of shared hits over j in back-partial cleaning
  (reconstruction_helper.cuh:1812-1838, 1879-1904), hand-worked;
* the station-1 sagitta windows (reconstruction_helper.cuh:556-613) against a
  closed form derived for lines through the target or the dump point.
"""
import ctypes as C

import numpy as np
import pytest

KLR = [(1, 1), (1, -1), (-1, 1), (-1, -1)]   # lrpossibility (reconstruction_constants.h:122)
F_Z, F_COSTHETA, F_SINTHETA, F_SLOPE_MAX, F_INTER_MAX = 0, 9, 14, 25, 26


@pytest.fixture(scope="module")
def lib(orc):
    return orc.load()


@pytest.fixture(scope="module")
def planes(ogeo):
    p, _, _ = ogeo.export()
    return p.astype(np.float64)


def lr_x(lib, ogeo, x0, tx, ex0, etx, det, pos, drift, sign, thr=150.0):
    det = np.asarray(det, np.int16)
    pos = np.asarray(pos, np.float32)
    drift = np.asarray(drift, np.float32)
    s = np.asarray(sign, np.int16).copy()
    lib.orc_lr_xhits(ogeo.handle, x0, tx, ex0, etx, len(det), det.ctypes.data, pos.ctypes.data, drift.ctypes.data,
                     s.ctypes.data, thr)
    return s.tolist()


def lr_new(lib, ogeo, x0, tx, y0, ty, ex0, etx, ey0, ety, det, pos, drift, sign, thr=150.0):
    det = np.asarray(det, np.int16)
    pos = np.asarray(pos, np.float32)
    drift = np.asarray(drift, np.float32)
    s = np.asarray(sign, np.int16).copy()
    lib.orc_lr_newhits(ogeo.handle, x0, tx, y0, ty, ex0, etx, ey0, ety, len(det), det.ctypes.data, pos.ctypes.data,
                       drift.ctypes.data, s.ctypes.data, thr)
    return s.tolist()


# ---- left/right, x-only variant: hand-worked pairs on D2 X (16) and X' (15) ----

def pair_on_line(planes, x0, tx, si, sj, di=0.3, dj=0.2, i=16, j=15):
    """wire positions such that the line x0 + tx z passes at drift di (dj) on
    side si (sj) of wire i (j): the line's x = pos + s * drift"""
    zi, zj = planes[F_Z][i], planes[F_Z][j]
    return [x0 + tx * zi - si * di, x0 + tx * zj - sj * dj], [di, dj], [i, j]


@pytest.mark.parametrize("k", [1, 2, 3])
def test_lr_true_hypothesis_is_taken(lib, ogeo, planes, k):
    si, sj = KLR[k]
    pos, drift, det = pair_on_line(planes, 10.0, 0.05, si, sj)
    assert lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det, pos, drift, [0, 0]) == [si, sj]


def test_lr_first_hypothesis_is_never_taken(lib, ogeo, planes):
    # indexmin > 0: the (+1, +1) hypothesis, even when it fits best, leaves the signs unset
    pos, drift, det = pair_on_line(planes, 10.0, 0.05, 1, 1)
    assert lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det, pos, drift, [0, 0]) == [0, 0]


def test_lr_threshold(lib, ogeo, planes):
    # the line misses the expected one by 2 cm with err_x0 0.01: best pull about 200
    pos, drift, det = pair_on_line(planes, 12.0, 0.05, 1, -1)
    assert lr_x(lib, ogeo, 10.0, 0.05, 0.01, 0.01, det, pos, drift, [0, 0]) == [0, 0]            # thr 150
    assert lr_x(lib, ogeo, 10.0, 0.05, 0.01, 0.01, det, pos, drift, [0, 0], thr=1.0e4) == [1, -1]
    # the threshold is strict: pull < thr (float64 pull of the best hypothesis, then +-0.1 %)
    xi, xj = pos[0] + drift[0], pos[1] - drift[1]
    zi, zj = planes[F_Z][16], planes[F_Z][15]
    sl = (xi - xj) / (zi - zj)
    it = xi - sl * zi
    pull = np.sqrt((0.05 - sl) ** 2 / 0.01 ** 2 + (10.0 - it) ** 2 / 0.01 ** 2)
    assert lr_x(lib, ogeo, 10.0, 0.05, 0.01, 0.01, det, pos, drift, [0, 0], thr=pull * 1.001) == [1, -1]
    assert lr_x(lib, ogeo, 10.0, 0.05, 0.01, 0.01, det, pos, drift, [0, 0], thr=pull * 0.999) == [0, 0]


def test_lr_steps_back_over_a_non_adjacent_hit(lib, ogeo, planes):
    pos, drift, det = pair_on_line(planes, 10.0, 0.05, -1, 1)
    # (22, 16) are not adjacent detectors: the walk steps back one hit and pairs (16, 15)
    got = lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, [22] + det, [0.0] + pos, [0.4] + drift, [0, 0, 0])
    assert got == [0, -1, 1]
    # pairs (16, 15) then (22, 21): both resolved
    p2, d2, det2 = pair_on_line(planes, 10.0, 0.05, 1, -1, i=22, j=21)
    got = lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det + det2, pos + p2, drift + d2, [0, 0, 0, 0])
    assert got == [-1, 1, 1, -1]
    # a lone last hit has no partner (its partner index is past the list)
    got = lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det + [22], pos + [p2[0]], drift + [d2[0]], [0, 0, 0])
    assert got == [-1, 1, 0]


def test_lr_pairs_with_both_signs_set_are_kept(lib, ogeo, planes):
    pos, drift, det = pair_on_line(planes, 10.0, 0.05, 1, -1)
    assert lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det, pos, drift, [-1, 1]) == [-1, 1]
    assert lr_x(lib, ogeo, 10.0, 0.05, 1.0, 0.01, det, pos, drift, [-1, 0]) == [1, -1]   # one unset: resolved


def hypotheses64(planes, det_i, det_j, pos, drift, sexp, esl, iexp, ein, limits):
    zi, zj = planes[F_Z][det_i], planes[F_Z][det_j]
    out = []
    for si, sj in KLR:
        xi, xj = pos[0] + si * drift[0], pos[1] + sj * drift[1]
        sl = (xi - xj) / (zi - zj)
        it = xi - sl * zi
        ok = not limits or (abs(sl) <= planes[F_SLOPE_MAX][det_i] and abs(it) <= planes[F_INTER_MAX][det_i])
        pull = np.sqrt((sexp - sl) ** 2 / esl ** 2 + (iexp - it) ** 2 / ein ** 2)
        out.append((pull if ok else np.inf, sl, it))
    return out


@pytest.mark.parametrize("variant", ["x", "new"])
def test_lr_enumeration_of_the_four_hypotheses(lib, ogeo, planes, variant):
    rng = np.random.default_rng(7)
    # x pairs (D2 X/X') and stereo pairs (D2 U/U', D3p V/V')
    pairs = [(16, 15), (22, 21)] if variant == "x" else [(14, 13), (24, 23), (18, 17)]
    checked = resolved = 0
    for trial in range(4000):
        di, dj = pairs[trial % len(pairs)]
        x0, tx = rng.uniform(-60, 60), rng.uniform(-0.15, 0.15)
        y0, ty = (rng.uniform(-40, 40), rng.uniform(-0.1, 0.1)) if variant == "new" else (0.0, 0.0)
        ex0, etx = rng.uniform(0.05, 3.0), rng.uniform(0.001, 0.02)
        ey0, ety = rng.uniform(0.05, 3.0), rng.uniform(0.001, 0.02)
        pos = list(np.float32(rng.uniform(-100, 100, 2)))
        if rng.random() < 0.7:   # near the line, so that the pull is small and the threshold matters
            c, s = planes[F_COSTHETA], planes[F_SINTHETA]
            for q, d in enumerate((di, dj)):
                pos[q] = float(np.float32((x0 + tx * planes[F_Z][d]) * c[d] + (y0 + ty * planes[F_Z][d]) * s[d]
                                          + rng.normal(0, 0.5)))
        drift = list(np.float32(rng.uniform(0, 0.6, 2)))
        c, s = planes[F_COSTHETA][di], planes[F_SINTHETA][di]
        if variant == "x":
            sexp, esl, iexp, ein = tx, etx, x0, ex0
        else:
            sexp, esl = c * tx + s * ty, abs(c * etx) + abs(s * ety)
            iexp, ein = c * x0 + s * y0, abs(c * ex0) + abs(s * ey0)
        h = hypotheses64(planes, di, dj, pos, drift, sexp, esl, iexp, ein, limits=variant == "new")
        pulls = np.array([p for p, _, _ in h])
        order = np.argsort(pulls, kind="stable")
        best = int(order[0])
        # skip near-ties and pulls or limits near their cut (float32 vs float64)
        fin = pulls[np.isfinite(pulls)]
        if fin.size == 0:
            expect = [0, 0]
        else:
            if fin.size > 1 and abs(pulls[order[1]] - pulls[best]) < 1e-3 * (1 + pulls[best]):
                continue
            if abs(pulls[best] - 150.0) < 0.5:
                continue
            if variant == "new" and any(abs(abs(sl) - planes[F_SLOPE_MAX][di]) < 1e-4 or
                                        abs(abs(it) - planes[F_INTER_MAX][di]) < 1e-2 for _, sl, it in h):
                continue
            expect = list(KLR[best]) if (best > 0 and pulls[best] < 150.0) else [0, 0]
        if variant == "x":
            got = lr_x(lib, ogeo, x0, tx, ex0, etx, [di, dj], pos, drift, [0, 0])
        else:
            got = lr_new(lib, ogeo, x0, tx, y0, ty, ex0, etx, ey0, ety, [di, dj], pos, drift, [0, 0])
        assert got == expect, (trial, pulls, got, expect)
        checked += 1
        resolved += expect != [0, 0]
    assert checked > 3000 and resolved > 500


def test_lr_slope_limit_only_in_the_new_hits_variant(lib, ogeo, planes):
    # a steep pair: the line's own hypothesis breaks slope_max of the plane
    d_i, d_j = 14, 13
    smax = planes[F_SLOPE_MAX][d_i]
    steep = 3.0 * smax
    zi, zj = planes[F_Z][d_i], planes[F_Z][d_j]
    pos = [float(np.float32(5.0 + steep * zi - 0.3)), float(np.float32(5.0 + steep * zj + 0.2))]
    drift = [0.3, 0.2]
    c, s = planes[F_COSTHETA][d_i], planes[F_SINTHETA][d_i]
    # expectation of the new-hits variant equal to the steep line in the plane's projection
    got_x = lr_x(lib, ogeo, 5.0, steep, 1.0, 0.01, [d_i, d_j], pos, drift, [0, 0], thr=1e6)
    assert got_x == [1, -1]
    h = hypotheses64(planes, d_i, d_j, pos, drift, steep, 0.01, 5.0, 1.0, limits=True)
    assert not np.isfinite(h[1][0])   # the true hypothesis is over the limits
    got_new = lr_new(lib, ogeo, 5.0 * c, steep * c, 5.0 * s, steep * s, 1.0, 0.01, 1.0, 0.01, [d_i, d_j], pos, drift,
                     [0, 0], thr=1e6)
    assert got_new != [1, -1]


def test_lr_single_hit_rule(lib, ogeo, planes):
    det = np.array([16, 15, 22], np.int16)
    x0, tx = 3.0, 0.02
    exp = [x0 + tx * planes[F_Z][d] for d in det]
    pos = np.array([exp[0] - 0.5, exp[1] + 0.5, exp[2]], np.float32)   # left, right, exactly on it
    s = np.array([0, 0, 0], np.int16)
    lib.orc_lr_single_xhits(ogeo.handle, x0, tx, 3, det.ctypes.data, pos.ctypes.data, s.ctypes.data)
    assert s.tolist() == [1, -1, -1]       # expected > pos -> +1, else -1 (ties -> -1)
    s = np.array([-1, 0, 1], np.int16)
    lib.orc_lr_single_xhits(ogeo.handle, x0, tx, 3, det.ctypes.data, pos.ctypes.data, s.ctypes.data)
    assert s.tolist() == [-1, -1, 1]       # signs already set are kept
    det2 = np.array([14, -1], np.int16)
    c, sn = planes[F_COSTHETA][14], planes[F_SINTHETA][14]
    y0, ty = 4.0, -0.03
    e14 = (planes[F_Z][14] * tx + x0) * c + (planes[F_Z][14] * ty + y0) * sn
    pos2 = np.array([e14 - 0.2, 0.0], np.float32)
    s2 = np.array([0, 0], np.int16)
    lib.orc_lr_single_newhits(ogeo.handle, x0, tx, y0, ty, 2, det2.ctypes.data, pos2.ctypes.data, s2.ctypes.data)
    assert s2.tolist() == [1, 0]           # negative detector ids are skipped


# ---- enumeration order and the winner rule --------------------------------------

def trace_events(oc, orc, lib, ogeo, ev, params):
    lib.orc_trace(1)
    try:
        orc.reconstruct(ogeo, params, ev, threads=1)
        cap = 4_000_000
        kind = np.zeros(cap, np.int32); evi = np.zeros(cap, np.int64); canon = np.zeros(cap, np.int32)
        idx = np.zeros((cap, 4), np.int32); chi2 = np.zeros(cap, np.float32)
        n = lib.orc_trace_fetch(cap, kind.ctypes.data, evi.ctypes.data, canon.ctypes.data, idx.ctypes.data,
                                chi2.ctypes.data)
    finally:
        lib.orc_trace(0)
    assert n <= cap
    return kind[:n], evi[:n], canon[:n], idx[:n], chi2[:n]


def test_yz_and_global_enumeration_order_and_winner(oc, orc, lib, geo, ogeo):
    ev = oc.generate(geo, 60, seed=41, mult_mode=1, mult_mean=500.0, mult_cap=1000)
    params = oc.default_params(max_hits_per_event=1024, max_d2_mult=2000, max_d3_mult=2000, max_d0_mult=2000)
    kind, evi, canon, idx, chi2 = trace_events(oc, orc, lib, ogeo, ev, params)
    # group rows per (event, tracklet) and stage, in recorded order
    n_yz = n_gl = winners_yz = winners_gl = 0
    r = 0
    while r < len(kind):
        k = kind[r]
        if k not in (4, 5):
            r += 1
            continue
        key = (evi[r], canon[r])
        sizes = idx[r]
        cand_kind, win_kind = (0, 1) if k == 4 else (2, 3)
        r += 1
        flats, vals, win = [], [], None
        while r < len(kind) and (evi[r], canon[r]) == key and kind[r] in (cand_kind, win_kind):
            i = idx[r]
            if kind[r] == cand_kind:
                if k == 4:   # YZ: (u3, v3) pairs outer with u3 fastest, (u2, v2) inner with u2 fastest
                    nu3, nv3, nu2, nv2 = sizes
                    flats.append(((i[1] * nu3 + i[0]) * (nu2 * nv2)) + i[3] * nu2 + i[2])
                else:        # station 1: x entries outer, (u, v) pairs inner with u fastest
                    nx1, nu1, nv1 = sizes[:3]
                    flats.append(i[0] * (nu1 * nv1) + i[2] * nu1 + i[1])
                vals.append(float(chi2[r]))
            else:
                win = (tuple(i), float(chi2[r]))
            r += 1
        assert all(a < b for a, b in zip(flats, flats[1:])), f"enumeration order at {key}"
        # the winner: strict < against a start of 10000.1 (station 1 also drops chi2 > 10000),
        # so the first of equal minima is kept
        best, best_pos = 10000.1, None
        for p, v in enumerate(vals):
            if k == 5 and v > 10000.0:
                continue
            if v < best:
                best, best_pos = v, p
        if best_pos is None:
            assert win is None
        else:
            assert win is not None and win[1] == np.float32(best)
            w = win[0]
            if k == 4:
                nu3, nv3, nu2, nv2 = sizes
                wflat = ((w[1] * nu3 + w[0]) * (nu2 * nv2)) + w[3] * nu2 + w[2]
            else:
                nx1, nu1, nv1 = sizes[:3]
                wflat = w[0] * (nu1 * nv1) + w[2] * nu1 + w[1]
            assert wflat == flats[best_pos], f"winner at {key} is not the first minimum"
        if k == 4:
            n_yz += 1
            winners_yz += win is not None
        else:
            n_gl += 1
            winners_gl += win is not None
    assert n_yz > 300 and winners_yz > 100 and n_gl > 20 and winners_gl > 10


# ---- back-partial cleaning: shared hits accumulate over j ----------------------

def record(stid, chi2, hits):
    """148-float tracklet record: stationID, nhits, chi2, (det, elem) per hit"""
    r = np.zeros(148, np.float32)
    r[0], r[2], r[3] = stid, len(hits), chi2
    for h, (d, e) in enumerate(hits):
        r[22 + h], r[40 + h] = d, e
    return r


def clean(lib, recs, slots):
    recs = np.ascontiguousarray(np.stack(recs), np.float32)
    slots = np.asarray(slots, np.int32)
    out = np.zeros(len(slots), np.float32)
    lib.orc_back_cleaning(len(slots), recs.ctypes.data, slots.ctypes.data, out.ctypes.data)
    return out.tolist()


def six(base):
    return [(13 + k, base + k) for k in range(6)]


def test_cleaning_accumulates_shared_hits_within_a_slot(lib):
    a = six(10)
    b = six(40)
    b[0] = a[0]                    # B shares one hit with A
    c = six(70)
    c[1] = a[1]                    # C shares one (other) hit with A
    # A is better (chi2/(n - 4) = 0.5); a threshold of 6 * 0.3333 = 2.0: one shared
    # hit never exceeds it, but the count carried from B makes C's two
    got = clean(lib, [record(5, 1.0, a), record(5, 100.0, b), record(5, 100.0, c)], [0, 0, 0])
    assert got == [5.0, 5.0, 3.0]
    # without B the count stays at one and C survives
    assert clean(lib, [record(5, 1.0, a), record(5, 100.0, c)], [0, 0]) == [5.0, 5.0]


def test_cleaning_accumulates_per_other_slot(lib):
    a, b, c = six(10), six(40), six(70)
    b[0] = a[0]
    c[1] = a[1]
    # A in slot 0, B and C in slot 3: the cross pass carries the count over B to C
    assert clean(lib, [record(5, 1.0, a), record(5, 100.0, b), record(5, 100.0, c)], [0, 3, 3]) == [5.0, 5.0, 3.0]
    # B and C in different slots: the count restarts for each other slot
    assert clean(lib, [record(5, 1.0, a), record(5, 100.0, b), record(5, 100.0, c)], [0, 3, 5]) == [5.0, 5.0, 5.0]


def test_cleaning_demotes_the_worse_of_a_pair(lib):
    a, b = six(10), six(40)
    b[0], b[1], b[2] = a[0], a[1], a[2]      # three shared hits > 2.0
    assert clean(lib, [record(5, 100.0, a), record(5, 1.0, b)], [0, 0]) == [3.0, 5.0]   # the earlier one is worse
    assert clean(lib, [record(5, 1.0, a), record(5, 100.0, b)], [0, 0]) == [5.0, 3.0]
    assert clean(lib, [record(3, 1.0, a), record(5, 100.0, b)], [0, 0]) == [3.0, 5.0]   # stationID < 5 is skipped


# ---- station-1 sagitta windows -------------------------------------------------

SAG_T_C, SAG_T_W, SAG_D_C, SAG_D_W = 1.85, 0.25, 1.5, 0.3   # reconstruction_constants.h:75-81 (E1039)
Z_TARGET, Z_DUMP = -300.0, 42.0
VIEW_D2 = {0: 16, 1: 18, 2: 14}      # station-2 plane whose angle projects each view (detsuperid[2][view] * 2)
VIEW_D1 = {0: 4, 1: 2, 2: 6}         # station-1 plane of each view


def sagitta(lib, ogeo, x0, tx, y0, ty, dlast):
    pexp = np.zeros(3, np.float32)
    win = np.zeros(3, np.float32)
    lib.orc_sagitta(ogeo.handle, x0, tx, y0, ty, dlast, pexp.ctypes.data, win.ctypes.data)
    return pexp.astype(np.float64), win.astype(np.float64)


@pytest.mark.parametrize("through", ["target", "dump"])
def test_sagitta_windows_closed_form(lib, ogeo, planes, through):
    # a line through the beam axis at z0: its projection on each view is p(z) = k (z - z0);
    # the window is the hull of the two hypotheses' intervals [c +- w], with
    #   through the target: s2t = 0 (pet = p(z1), wt = 0), and
    #     s2d = k (zd - zt)(z3 - z2) / (z3 - zd),  ped = 1.5 s2d + k (z3 - zt)(z1 - zd)/(z3 - zd),  wd = 0.3 |s2d|
    #   through the dump: s2d = 0 (ped = p(z1), wd = 0), and
    #     s2t = k (zt - zd)(z3 - z2) / (z3 - zt), pet = 1.85 s2t + k (z3 - zd)(z1 - zt)/(z3 - zt), wt = 0.25 |s2t|
    rng = np.random.default_rng(3)
    z0 = Z_TARGET if through == "target" else Z_DUMP
    zt, zd = Z_TARGET, Z_DUMP
    for _ in range(50):
        tx, ty = rng.uniform(-0.15, 0.15), rng.uniform(-0.1, 0.1)
        x0, y0 = -tx * z0, -ty * z0
        dlast = int(rng.choice([21, 22, 27, 28, 23, 24]))
        z3 = planes[F_Z][dlast]
        pexp, win = sagitta(lib, ogeo, np.float32(x0), np.float32(tx), np.float32(y0), np.float32(ty), dlast)
        for v in range(3):
            d2 = VIEW_D2[v]
            z1, z2 = planes[F_Z][VIEW_D1[v]], planes[F_Z][d2]
            k = tx * planes[F_COSTHETA][d2] + ty * planes[F_SINTHETA][d2]
            if through == "target":
                pet, wt = k * (z1 - zt), 0.0
                s2d = k * (zd - zt) * (z3 - z2) / (z3 - zd)
                ped, wd = SAG_D_C * s2d + k * (z3 - zt) * (z1 - zd) / (z3 - zd), abs(SAG_D_W * s2d)
            else:
                ped, wd = k * (z1 - zd), 0.0
                s2t = k * (zt - zd) * (z3 - z2) / (z3 - zt)
                pet, wt = SAG_T_C * s2t + k * (z3 - zd) * (z1 - zt) / (z3 - zt), abs(SAG_T_W * s2t)
            lo, hi = min(pet - wt, ped - wd), max(pet + wt, ped + wd)
            scale = 1.0 + abs(k) * 2000.0
            assert pexp[v] == pytest.approx(0.5 * (lo + hi), abs=2e-5 * scale)
            assert win[v] == pytest.approx(0.5 * (hi - lo), abs=2e-5 * scale)
