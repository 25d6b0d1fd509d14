"""Pins the oracle's fits and geometry helpers against textbook closed forms.

trackanalyticalminimizer.cuh:32-119 (weighted straight-line fit, weight = the
resolution itself), trackfittingtools.cuh:3-82 (pivoted 2x2 inverse),
reconstruction_helper.cuh:75-158 (wire y at the track x, wire x at y) and
:1149-1204 (distance between track and wire lines).
"""
import ctypes as C

import numpy as np
import pytest


def fit(orc, x, z, w):
    x, z, w = (np.asarray(a, np.float32) for a in (x, z, w))
    out = np.zeros(4, np.float32)
    orc.load().orc_fit2d(len(x), x.ctypes.data, z.ctypes.data, w.ctypes.data, out.ctypes.data)
    return out


def normal_eq(x, z, w):
    x, z, w = (np.asarray(a, np.float64) for a in (x, z, w))
    A = np.array([[w.sum(), (w * z).sum()], [(w * z).sum(), (w * z * z).sum()]])
    B = np.array([(w * x).sum(), (w * x * z).sum()])
    Ai = np.linalg.inv(A)
    return Ai @ B, np.sqrt(np.abs(np.diag(Ai)))


def test_fit_recovers_exact_line(orc):
    z = [1315.3, 1322.2, 1923.0, 1925.0]
    x = [12.0 + 0.05 * zz for zz in z]
    p = fit(orc, x, z, [0.1, 0.1, 0.08, 0.08])
    assert p[0] == pytest.approx(12.0, abs=2e-3) and p[1] == pytest.approx(0.05, abs=2e-6)


@pytest.mark.parametrize("n", [2, 3, 4, 8, 12])
def test_fit_matches_normal_equations_weight_is_sigma(orc, n):
    rng = np.random.default_rng(n)
    for _ in range(200):
        z = np.sort(rng.uniform(600, 2400, n))
        if n == 2:
            z[1] = z[0] + rng.uniform(5, 1000)
        x = rng.uniform(-150, 150) + rng.uniform(-0.15, 0.15) * z + rng.normal(0, 0.1, n)
        w = rng.uniform(0.05, 0.5, n)
        p = fit(orc, x, z, w)
        x32, z32, w32 = x.astype(np.float32), z.astype(np.float32), w.astype(np.float32)
        par, err = normal_eq(x32, z32, w32)
        # FP32 normal equations: error grows with the condition number of A
        A = np.array([[w32.sum(), (w32 * z32).sum()], [(w32 * z32).sum(), (w32 * z32 * z32).sum()]], np.float64)
        tol = 1e-4 + 4e-7 * np.linalg.cond(A)
        assert p[1] == pytest.approx(par[1], rel=tol, abs=2e-6)
        assert p[0] == pytest.approx(par[0], rel=tol, abs=5e-3)
        assert p[2] == pytest.approx(err[0], rel=tol)
        assert p[3] == pytest.approx(err[1], rel=tol)


def test_fit_weighting_is_not_inverse_variance(orc):
    z = np.array([600.0, 1300.0, 1900.0])
    x = np.array([0.0, 10.0, 0.0])
    w = np.array([0.1, 1.0, 0.1])
    p = fit(orc, x, z, w)
    par_sigma, _ = normal_eq(x, z, w)
    par_ivar, _ = normal_eq(x, z, 1.0 / w ** 2)
    assert p[0] == pytest.approx(par_sigma[0], rel=1e-4)
    assert abs(p[0] - par_ivar[0]) > 1.0


def test_matinv2_with_and_without_pivot(orc):
    rng = np.random.default_rng(3)
    for _ in range(500):
        A = rng.normal(size=(2, 2)).astype(np.float32)
        if abs(np.linalg.det(A)) < 1e-2:
            continue
        a = np.array([A[0, 0], A[1, 0], A[0, 1], A[1, 1]], np.float32)  # read column-major
        ai = np.zeros(4, np.float32)
        orc.load().orc_matinv2(a.ctypes.data, ai.ctypes.data)
        ref = np.linalg.inv(A.astype(np.float64))
        got = ai.reshape(2, 2)
        assert np.allclose(got, ref, rtol=2e-4, atol=2e-4 * np.abs(ref).max())


def dca(orc, *a):
    return orc.load().orc_dca(*[float(v) for v in a])


def textbook_dca(p1, dp, x0, y0, tx, ty):
    t = np.array([tx, ty, 1.0])
    n = np.cross(t, dp)
    return np.dot(np.asarray(p1) - np.array([x0, y0, 0.0]), n) / np.linalg.norm(n)


def test_dca_equals_line_distance_for_planar_wires(orc):
    # dpz = 0: the reference denominator equals |t x dp| exactly
    rng = np.random.default_rng(4)
    for _ in range(500):
        p1 = [rng.uniform(-150, 150), rng.uniform(-150, 0), rng.uniform(600, 2400)]
        ang = rng.uniform(-0.3, 0.3)
        dp = np.array([-264 * np.sin(ang), 264 * np.cos(ang), 0.0])
        x0, y0, tx, ty = rng.uniform(-100, 100), rng.uniform(-40, 40), rng.uniform(-.15, .15), rng.uniform(-.1, .1)
        got = dca(orc, *p1, *dp, x0, y0, tx, ty)
        assert got == pytest.approx(textbook_dca(p1, dp, x0, y0, tx, ty), rel=1e-4, abs=2e-3)


def test_dca_denominator_quirk(orc):
    # the reference denominator carries -2*ty*dpx*dpz where |t x dp|^2 has
    # -2*tx*dpx*dpz, and omits the dpz^2 terms (reconstruction_helper.cuh:1157);
    # the oracle must follow the reference, not the textbook
    p1, dp = [10.0, -100.0, 1300.0], np.array([60.0, 250.0, 40.0])
    x0, y0, tx, ty = 5.0, 3.0, 0.12, -0.08
    num = np.dot(np.array(p1) - [x0, y0, 0], np.cross([tx, ty, 1.0], dp))
    den_ref = dp[1] ** 2 * (1 + tx * tx) + dp[0] ** 2 * (1 + ty * ty) - 2 * (
        ty * dp[0] * dp[2] + ty * dp[1] * dp[2] + tx * ty * dp[0] * dp[1])
    got = dca(orc, *p1, *dp, x0, y0, tx, ty)
    assert got == pytest.approx(num / np.sqrt(den_ref), rel=1e-5)
    assert abs(got - textbook_dca(p1, dp, x0, y0, tx, ty)) > 1e-3


def planes(oc, geo):
    pl, _ = geo.export()
    names = ["z", "nelem", "cellwidth", "spacing", "xoffset", "scalex", "x0", "x1", "x2", "costheta", "scaley",
             "y0", "y1", "y2", "sintheta", "resolution", "p1x", "p1y", "p1z", "deltapx", "deltapy", "deltapz",
             "dp1x", "dp1y", "dp1z", "slope_max", "inter_max"]
    return {n: pl[i] for i, n in enumerate(names)}


@pytest.mark.parametrize("d", [1, 2, 5, 6, 13, 14, 17, 18, 19, 20, 23, 24, 25, 26, 29, 30])
def test_uv_y_lies_on_wire_at_track_x(oc, orc, geo, ogeo, d):
    P = planes(oc, geo)
    lib = orc.load()
    rng = np.random.default_rng(d)
    for _ in range(100):
        e = int(rng.integers(1, int(P["nelem"][d]) + 1))
        p1x = P["p1x"][d] + P["dp1x"][d] * (e - 1)
        p1y = P["p1y"][d] + P["dp1y"][d] * (e - 1)
        dpx, dpy = P["deltapx"][d], P["deltapy"][d]
        # a track x inside the wire's x span at this z
        xt = p1x + rng.uniform(0.05, 0.95) * dpx
        tx = 0.01
        x0 = xt - P["z"][d] * tx
        y, ey = C.c_float(), C.c_float()
        ok = lib.orc_y_uvhit(ogeo.handle, d, e, 0.0, 0, x0, tx, C.byref(y), C.byref(ey))
        assert ok == 1
        # (xt, y) is on the wire line (within float rounding)
        assert y.value == pytest.approx(p1y + (xt - p1x) * dpy / dpx, abs=2e-3)
        assert ey.value == pytest.approx(P["resolution"][d] * abs(dpy / dpx), rel=1e-5)
        # sign term: +copysign(1, sin theta) * 2 * (drift - 0.5 * sign), then clamp to the wire's y span
        for s in (1, -1):
            y2 = C.c_float()
            lib.orc_y_uvhit(ogeo.handle, d, e, 0.3, s, x0, tx, C.byref(y2), C.byref(ey))
            want = min(max(y.value + np.copysign(1.0, P["sintheta"][d]) * 2 * (0.3 - 0.5 * s), p1y), p1y + dpy)
            assert y2.value == pytest.approx(want, abs=2e-3)
        # outside the span by more than the drift: rejected
        xo = p1x + (1.5 if dpx > 0 else -0.5) * dpx
        assert lib.orc_y_uvhit(ogeo.handle, d, e, 0.1, 0, xo - P["z"][d] * tx, tx, C.byref(y), C.byref(ey)) == 0


@pytest.mark.parametrize("d", [3, 4, 15, 16, 21, 22, 27, 28])
def test_xz_at_y_on_wire(oc, orc, geo, ogeo, d):
    P = planes(oc, geo)
    lib = orc.load()
    for e in (1, 30, int(P["nelem"][d])):
        p1 = np.array([P["p1x"][d] + P["dp1x"][d] * (e - 1), P["p1y"][d] + P["dp1y"][d] * (e - 1),
                       P["p1z"][d] + P["dp1z"][d] * (e - 1)], np.float64)
        dp = np.array([P["deltapx"][d], P["deltapy"][d], P["deltapz"][d]], np.float64)
        for y0, ty in ((0.0, 0.0), (10.0, 0.02), (-30.0, -0.05)):
            x, z = C.c_float(), C.c_float()
            lib.orc_xz_fy(ogeo.handle, d, e, 0.5, 0, y0, ty, C.byref(x), C.byref(z))
            # the returned point is on the wire where the wire meets y = y0 + ty*z(at y=0 crossing)
            s = (y0 + ty * (p1[2] - p1[1] * dp[2] / dp[1]) - p1[1]) / dp[1]
            assert x.value == pytest.approx(p1[0] + s * dp[0], abs=2e-3)
            assert z.value == pytest.approx(p1[2] + s * dp[2], abs=2e-3)
            xs, zs = C.c_float(), C.c_float()
            lib.orc_xz_fy(ogeo.handle, d, e, 0.5, -1, y0, ty, C.byref(xs), C.byref(zs))
            assert xs.value == pytest.approx(x.value - 0.5, abs=1e-5)


def test_hodo_window_widens_with_hit_order(orc, ogeo):
    # H2 (cell 13.0, fudge 0.2), x_hodo = 50, err_x0 = 0.1, err_tx = 1e-5 at z = 1421.11.
    # per hit: [p - 6.5, p + 6.5] -> xmin -= 13*0.2 -> [p - 9.1, p + 6.5]; xmax += 15.6*0.2 -> p + 9.62
    # (asymmetric); err_x = 3*(0.0142 + 0.1) + 18.72*0.15 = 3.151 after hit 1, 5.959 after hit 2
    # (accumulates, reconstruction_helper.cuh:1127).  Match needs x_hodo <= p + 9.62 + err_x:
    # p >= 37.23 as the first hit, p >= 34.42 as the second.
    lib = orc.load()
    x0, tx, ex0, etx = 50.0, 0.0, 0.1, 1e-5
    one = np.array([36.0], np.float32)
    assert lib.orc_hodo_xz(ogeo.handle, 1, 37, 1, one.ctypes.data, x0, tx, ex0, etx) == 0
    two = np.array([-100.0, 36.0], np.float32)
    assert lib.orc_hodo_xz(ogeo.handle, 1, 37, 2, two.ctypes.data, x0, tx, ex0, etx) == 1
    edge = np.array([37.3], np.float32)
    assert lib.orc_hodo_xz(ogeo.handle, 1, 37, 1, edge.ctypes.data, x0, tx, ex0, etx) == 1
    # low edge: x_hodo >= p - 12.25 for the first hit
    assert lib.orc_hodo_xz(ogeo.handle, 1, 37, 1, np.array([62.2], np.float32).ctypes.data, x0, tx, ex0, etx) == 1
    assert lib.orc_hodo_xz(ogeo.handle, 1, 37, 1, np.array([62.4], np.float32).ctypes.data, x0, tx, ex0, etx) == 0
