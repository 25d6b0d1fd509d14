"""R-T calibration (online_reconstruction.cu:220-242 parse, :387-397 per-hit
rule; the spline is ROOT TSpline3 with option "b1e1").

The oracle restates TSpline3's algorithm (de Boor's CUBSPL with prescribed end
slopes, binary-search interval choice).  It is pinned here by an independent
implementation (scipy's clamped CubicSpline, which solves for the spline a
different way), by a cubic with zero end slopes that any clamped spline must
reproduce exactly, by the interpolation and C2 conditions read off its
coefficients, and by the shipped calibration file's own numbers.  The C-ABI
argument checks run without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ocal(oc, orc):
    return orc.OracleCalibration.load(oc.DEFAULT_CALIBRATION)


def test_calibration_file_contents(ocal):
    t = ocal.tables
    assert sorted(t) == list(range(1, 31)) + list(range(47, 55))
    t0, t1, T, R = t[1]
    assert (t0, t1, len(T)) == (1650.0, 1750.0, 41) and (T[0], T[-1], R[0]) == (1650.0, 1750.0, 0.3175)
    t0, t1, T, R = t[13]
    assert (t0, t1, len(T)) == (1450.0, 1710.0, 105) and R[0] == 1.0105
    t0, t1, T, R = t[47]
    assert (t0, t1, len(T)) == (560.0, 1200.0, 128) and (T[0], T[-1], R[0]) == (562.5, 1197.5, 2.38125)
    for d, (t0, t1, T, R) in t.items():
        assert np.all(np.diff(T) > 0), d


def test_oracle_spline_against_scipy(ocal, orc):
    from scipy.interpolate import CubicSpline
    for d, (t0, t1, T, R) in ocal.tables.items():
        x = np.concatenate([np.linspace(min(t0, T[0]) - 5.0, max(t1, T[-1]) + 5.0, 4001), T])
        ref = CubicSpline(T, R, bc_type=((1, 0.0), (1, 0.0)), extrapolate=True)(x)
        got = orc.rt_eval(T, R, x)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(R).max()), d


def test_oracle_spline_reproduces_clamped_cubic(orc):
    # p(t) = a + k (3 s^2 - 2 s^3), s = (t - t0) / (t1 - t0): p'(t0) = p'(t1) = 0,
    # so the clamped spline through any knots of p is p itself
    rng = np.random.default_rng(5)
    for n in (2, 3, 7, 40):
        T = np.sort(rng.uniform(0.0, 100.0, n))
        T[0], T[-1] = 0.0, 100.0
        if n > 2:
            T[1:-1] = np.sort(rng.uniform(1.0, 99.0, n - 2))
        p = lambda t: 0.3 - 0.25 * (3 * (t / 100.0) ** 2 - 2 * (t / 100.0) ** 3)
        x = np.linspace(-10.0, 110.0, 2001)
        assert np.abs(orc.rt_eval(T, p(T), x) - p(x)).max() < 1e-12


def test_oracle_spline_coefficients_conditions(ocal, orc):
    for d in (1, 13, 47):
        _, _, T, R = ocal.tables[d]
        b, c, dd = orc.rt_spline(T, R)
        h = np.diff(T)
        end_val = R[:-1] + h * (b + h * (c + h * dd))
        end_slope = b + 2 * c * h + 3 * dd * h * h
        end_curv = 2 * c + 6 * dd * h
        assert np.allclose(end_val, R[1:], rtol=0, atol=1e-13)          # interpolation
        assert np.allclose(end_slope[:-1], b[1:], rtol=0, atol=1e-12)   # C1
        assert np.allclose(end_curv[:-1], 2 * c[1:], rtol=0, atol=1e-12)  # C2
        assert b[0] == 0.0 and abs(end_slope[-1]) < 1e-12              # "b1e1", zero slopes


def test_oracle_rule(ocal, orc, oc, geo):
    planes, _ = geo.export()
    cw = planes[2]                                   # cellwidth row
    det = np.array([1, 1, 1, 35, 47, 47, 47, 13], np.int16)
    tdc = np.array([1649.0, 1751.0, 1700.0, 1700.0, 1199.0, 561.0, 900.0, 1710.0], np.float32)
    drift = np.full(len(det), 0.777, np.float32)
    out = orc.rt_calibrate(ocal, cw, det, tdc, drift)
    assert out[0] == np.float32(cw[1] * 0.5)         # tdc < tmin: half cell
    assert out[1] == 0.0                             # tdc > tmax
    assert out[3] == np.float32(0.777)               # hodoscope keeps its drift
    for i in (2, 4, 5, 6, 7):                        # spline, also beyond the prop-tube knots
        _, _, T, R = ocal.tables[int(det[i])]
        assert out[i] == np.float32(orc.rt_eval(T, R, np.array([float(tdc[i])]))[0])
    # a detector without knots keeps its drift
    part = orc.OracleCalibration({k: v for k, v in ocal.tables.items() if k != 1})
    assert orc.rt_calibrate(part, cw, det[:1], tdc[:1], drift[:1])[0] == np.float32(0.777)


def test_abi_argument_checks(oc):
    lib = oc.load_library()
    h = C.c_void_p()
    assert lib.or_calibration_load(oc.DEFAULT_CALIBRATION.encode(), C.byref(h)) == 0 and h.value
    lib.or_calibration_free(h)
    assert lib.or_calibration_load(b"/nonexistent/calibration.txt", C.byref(h)) == 2

    def create(ids, nk, t, r):
        ids = np.array(ids, np.int32)
        nk = np.array(nk, np.int32)
        lo = np.zeros(len(ids), np.float32)
        hi = np.full(len(ids), 1e4, np.float32)
        t = np.array(t, np.float64)
        r = np.array(r, np.float64)
        hh = C.c_void_p()
        st = lib.or_calibration_create(len(ids), ids.ctypes.data, nk.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                                       t.ctypes.data, r.ctypes.data, C.byref(hh))
        if hh.value:
            lib.or_calibration_free(hh)
        return st

    assert create([1], [3], [0.0, 1.0, 2.0], [1.0, 0.5, 0.0]) == 0
    assert create([1], [1], [0.0], [1.0]) == 1                     # one knot
    assert create([1], [3], [0.0, 2.0, 1.0], [1.0, 0.5, 0.0]) == 1  # not increasing
    assert create([0], [2], [0.0, 1.0], [1.0, 0.0]) == 1            # detector id 0
    assert create([1], [513], np.arange(513.0), np.zeros(513)) == 7  # > OR_RT_MAX_KNOTS
    hits = oc.OrHits(0, 0, None, None, None, None, None, None, 0)
    assert lib.or_rt_calibrate(None, C.byref(hits), None) == 1
    assert lib.or_session_set_calibration(None, None) == 1


def test_product_parser_reads_the_file(oc, tmp_path):
    lib = oc.load_library()
    src = open(oc.DEFAULT_CALIBRATION).read()
    bad = tmp_path / "bad.txt"
    bad.write_text(src.split("\n", 1)[0] + "\n0 1650.0\n")          # truncated knot line
    h = C.c_void_p()
    assert lib.or_calibration_load(str(bad).encode(), C.byref(h)) == 3
    ok = tmp_path / "ok.txt"
    ok.write_text("\n" + src + "\n\n")                               # blank lines around
    assert lib.or_calibration_load(str(ok).encode(), C.byref(h)) == 0
    lib.or_calibration_free(h)
