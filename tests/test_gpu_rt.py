"""R-T calibration on the device against the oracle (online_reconstruction.cu:
387-397 with TSpline3 "b1e1" splines): or_rt_calibrate per hit, its host-buffer
path, and the chain with the calibration fused into its event load, which must
equal the chain run on the calibrated drift, and the oracle chain on it.

The spline is evaluated in FP64 on both sides from coefficients built two
different ways (second derivatives here, slopes in the oracle), so a result
may in principle differ in the last place of the float rounding; the test
allows one ulp and requires all but 1e-6 of the hits identical, and where the
spline is 0 (its last knots) an absolute difference of 1e-15 cm."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cal(oc):
    return oc.Calibration.load()


@pytest.fixture(scope="module")
def ocal(oc, orc):
    return orc.OracleCalibration.load(oc.DEFAULT_CALIBRATION)


def sweep_events(geo, ocal, n=200_000, seed=9):
    """Hits on every detector with tdc spread over [tmin - 30, tmax + 30]."""
    import orchain as oc
    rng = np.random.default_rng(seed)
    det = rng.integers(1, 55, n).astype(np.int16)
    lo = np.array([ocal.tables[d][0] if d in ocal.tables else 900.0 for d in range(55)])
    hi = np.array([ocal.tables[d][1] if d in ocal.tables else 1100.0 for d in range(55)])
    tdc = rng.uniform(lo[det] - 30.0, hi[det] + 30.0).astype(np.float32)
    # exact knot times and the table bounds themselves
    k = rng.integers(0, n, 2000)
    for j, i in enumerate(k):
        d = int(det[i])
        if d in ocal.tables:
            t0, t1, T, _ = ocal.tables[d]
            tdc[i] = np.float32([t0, t1, T[0], T[-1], T[j % len(T)]][j % 5])
    drift = rng.uniform(0.0, 1.0, n).astype(np.float32)
    elem = np.ones(n, np.int16)
    return oc.Events(det, elem, tdc, drift, None, np.array([0, n], np.int64))


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def test_rt_calibrate_against_oracle(oc, orc, geo, session, cal, ocal):
    import torch
    ev = sweep_events(geo, ocal)
    planes, _ = geo.export()
    ref = orc.rt_calibrate(ocal, planes[2], ev.det, ev.tdc, ev.drift)
    session.set_calibration(cal)
    try:
        t = session.upload(ev)
        dev = session.rt_calibrate_device(t)
        torch.cuda.synchronize()
        got = dev.cpu().numpy()
        host = session.rt_calibrate(ev)
    finally:
        session.set_calibration(None)
    assert np.array_equal(got.view(np.uint32), host.view(np.uint32))
    # at knots where R is 0 both sides leave a rounding residue of ~1e-19 cm
    big = np.abs(ref) >= 1e-6
    assert np.abs(got[~big] - ref[~big]).max() <= 1e-15
    assert ulp_diff(got[big], ref[big]).max() <= 1
    assert (got[big] != ref[big]).sum() <= max(1, int(big.sum()) // 1_000_000)
    hodo = (ev.det >= 31) & (ev.det <= 46)
    assert np.array_equal(got[hodo], ev.drift[hodo])
    # every branch of the rule is exercised
    t0 = np.array([ocal.tables[d][0] if d in ocal.tables else 0 for d in range(55)], np.float32)[ev.det]
    t1 = np.array([ocal.tables[d][1] if d in ocal.tables else 0 for d in range(55)], np.float32)[ev.det]
    calib = ~hodo
    assert (calib & (ev.tdc < t0)).sum() > 1000 and (calib & (ev.tdc > t1)).sum() > 1000


def test_chain_with_fused_calibration(oc, orc, geo, ogeo, session, cal):
    ev = oc.generate(geo, 600, seed=41, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    session.set_calibration(cal)
    try:
        fused = session.reconstruct(ev, records=True)
        drift = session.rt_calibrate(ev)
    finally:
        session.set_calibration(None)
    ev_cal = dataclasses.replace(ev, drift=drift)
    two_step = session.reconstruct(ev_cal, records=True)
    for f in ("status", "keep_mask", "n_tracks", "n_tracklets"):
        assert np.array_equal(getattr(fused, f), getattr(two_step, f)), f
    assert np.array_equal(fused.valid_tracks().view(np.uint8), two_step.valid_tracks().view(np.uint8))
    assert np.array_equal(fused.records.view(np.uint32), two_step.records.view(np.uint32))
    ref = orc.reconstruct(ogeo, session.params, ev_cal, records=True)
    assert np.array_equal(ref.n_tracks, fused.n_tracks)
    assert np.array_equal(ref.valid_tracks().view(np.uint8), fused.valid_tracks().view(np.uint8))
    assert fused.n_tracks.sum() > 0
    # calibration off again: the input drift is used
    plain = session.reconstruct(ev)
    again = orc.reconstruct(ogeo, session.params, ev)
    assert np.array_equal(plain.valid_tracks().view(np.uint8), again.valid_tracks().view(np.uint8))


def test_rt_calibrate_irregular_knots(oc, orc, geo, session):
    """Uneven knot spacing, 2 to 60 knots, ranges wider and narrower than the
    knots: the interval guess from the mean spacing must still land on the
    interval the binary search defines."""
    import torch
    rng = np.random.default_rng(21)
    tables = {}
    for d in list(range(1, 31)) + list(range(47, 55)):
        n = int(rng.integers(2, 61))
        T = np.cumsum(rng.exponential(1.0, n) ** 3 + 1e-3) * 10.0 + 500.0
        R = np.sort(rng.uniform(0.0, 2.0, n))[::-1].copy()
        tables[d] = (np.float32(T[0] - rng.uniform(-5, 20)), np.float32(T[-1] + rng.uniform(-5, 20)), T, R)
    cal = oc.Calibration.create(tables)
    ocal = orc.OracleCalibration(tables)
    n = 300_000
    det = rng.integers(1, 55, n).astype(np.int16)
    lo = np.array([tables[d][0] if d in tables else 0.0 for d in range(55)])
    hi = np.array([tables[d][1] if d in tables else 1.0 for d in range(55)])
    tdc = rng.uniform(lo[det] - 10.0, hi[det] + 10.0).astype(np.float32)
    k = rng.integers(0, n, 5000)
    for i in k:                                   # exact knot times
        if int(det[i]) in tables:
            T = tables[int(det[i])][2]
            tdc[i] = np.float32(T[int(rng.integers(0, len(T)))])
    ev = oc.Events(det, np.ones(n, np.int16), tdc, rng.uniform(0, 1, n).astype(np.float32), None,
                   np.array([0, n], np.int64))
    planes, _ = geo.export()
    session.set_calibration(cal)
    try:
        t = session.upload(ev)
        got = session.rt_calibrate_device(t)
        torch.cuda.synchronize()
        got = got.cpu().numpy()
    finally:
        session.set_calibration(None)
    ref = orc.rt_calibrate(ocal, planes[2], ev.det, ev.tdc, ev.drift)
    big = np.abs(ref) >= 1e-6
    assert np.abs(got[~big] - ref[~big]).max(initial=0.0) <= 1e-12
    assert ulp_diff(got[big], ref[big]).max() <= 1
    assert (got[big] != ref[big]).sum() <= max(1, int(big.sum()) // 100_000)
