"""FP64 reference of the final track fit and the per-quantity error report.

The device fits every output track (or_result.fit): measurement u = wire
position + sign * drift of each hit, model u = cos(theta) (x0 + tx z) +
sin(theta) (y0 + ty z) of its plane, weight 1/resolution^2.  This module
solves the same least-squares problem in float64 with numpy from the inputs,
the geometry and the track hit lists (independent of the device arithmetic),
and reports, per fitted quantity, how far the device result is from it.
"""
from __future__ import annotations

import numpy as np

# plane table rows (or_geometry_export order)
F_Z, F_COSTHETA, F_SINTHETA, F_RESOLUTION = 0, 9, 14, 15
QUANTITIES = ["x0", "tx", "y0", "ty", "chi2"]
# floors of the relative error: |device - fp64| / max(|fp64|, floor)
FLOORS = {"x0": 1.0, "tx": 1e-3, "y0": 1.0, "ty": 1e-3, "chi2": 1.0}


def track_hits(ev, res, planes, wire):
    """Per output track (valid slots, event order): padded hit arrays (T, 18)
    of u, z, cos, sin, weight and a mask."""
    nt = res.n_tracks.astype(np.int64)
    valid = np.arange(res.tracks.shape[1])[None, :] < nt[:, None]
    ev_of = np.broadcast_to(np.arange(len(nt), dtype=np.int64)[:, None], valid.shape)[valid]
    tr = res.tracks[valid]
    nh = tr["nhits"].astype(np.int64)
    k = np.arange(18)[None, :]
    mask = k < nh[:, None]
    gidx = ev.offsets[ev_of][:, None] + tr["hit_index"].astype(np.int64)
    gidx = np.where(mask, gidx, 0)
    det = ev.det[gidx].astype(np.int64)
    elem = ev.elem[gidx].astype(np.int64)
    drift = np.abs(ev.drift[gidx].astype(np.float64))
    sp = (tr["sign_plus"][:, None] >> k) & 1
    sm = (tr["sign_minus"][:, None] >> k) & 1
    sign = sp.astype(np.float64) - sm.astype(np.float64)
    pos = wire[det, elem].astype(np.float64)
    u = pos + sign * drift
    z = planes[F_Z][det].astype(np.float64)
    c = planes[F_COSTHETA][det].astype(np.float64)
    s = planes[F_SINTHETA][det].astype(np.float64)
    r = planes[F_RESOLUTION][det].astype(np.float64)
    w = np.where(mask, 1.0 / (r * r), 0.0)
    return ev_of, u, z, c, s, w, mask


def fp64_fit(u, z, c, s, w, mask):
    """Weighted least squares in float64 per row: (x0, tx, y0, ty), chi2."""
    wn = np.where(mask, w, 0.0)
    zr = (wn * z).sum(1) / np.maximum(wn.sum(1), 1e-300)
    zc = z - zr[:, None]
    H = np.stack([c, c * zc, s, s * zc], axis=2)                 # (T, 18, 4)
    A = np.einsum("tk,tki,tkj->tij", wn, H, H)
    b = np.einsum("tk,tki,tk->ti", wn, H, u)
    p = np.linalg.solve(A, b[..., None])[..., 0]
    m = np.einsum("tki,ti->tk", H, p)
    chi2 = (wn * (u - m) ** 2).sum(1)
    x0 = p[:, 0] - p[:, 1] * zr
    y0 = p[:, 2] - p[:, 3] * zr
    return np.stack([x0, p[:, 1], y0, p[:, 3], chi2], axis=1)


def error_report(ev, res, geo, max_tracks: int | None = None, cond_bound: float = 1.0e4) -> dict:
    """Per-quantity error of the device fit (res.fit) against the float64 fit;
    cond_bound is the FP32/FP64 switch the device ran with (OR_FIT_COND_FP32)."""
    planes, wire = geo.export()
    ev_of, u, z, c, s, w, mask = track_hits(ev, res, planes, wire)
    nt = res.n_tracks.astype(np.int64)
    valid = np.arange(res.tracks.shape[1])[None, :] < nt[:, None]
    dev = res.fit[valid].astype(np.float64)
    if max_tracks is not None:
        dev, u, z, c, s, w, mask = (a[:max_tracks] for a in (dev, u, z, c, s, w, mask))
    ref = fp64_fit(u, z, c, s, w, mask)
    out = {"tracks": int(len(dev)), "fp64_fraction": float((dev[:, 7] > 0.5).mean()) if len(dev) else 0.0,
           "cond_median": float(np.median(dev[:, 6])) if len(dev) else 0.0,
           "cond_bound": cond_bound, "floors": FLOORS}
    for i, q in enumerate(QUANTITIES):
        err = np.abs(dev[:, i] - ref[:, i]) / np.maximum(np.abs(ref[:, i]), FLOORS[q])
        out[q] = {"max_rel": float(err.max()) if len(err) else 0.0,
                  "p99_rel": float(np.quantile(err, 0.99)) if len(err) else 0.0,
                  "median_rel": float(np.median(err)) if len(err) else 0.0}
        fp32 = dev[:, 7] < 0.5
        out[q]["max_rel_fp32_path"] = float(err[fp32].max()) if fp32.any() else 0.0
    return out


# ---- the chain's own 2-parameter XZ fit -------------------------------------
# fit_2D_track_drift (trackanalyticalminimizer.cuh:78-119): points x at y = 0
# from the wire end points (calculate_xz_fy with y0 = ty = 0,
# reconstruction_helper.cuh:140-158) plus drift * sign, weights w = resolution
# (the reference weights by sigma, not 1/sigma^2), normal equations
# A = [sum w, sum w z; sum w z, sum w z^2], B = [sum w x, sum w x z],
# parameters A^-1 B, errors sqrt(|A^-1_jj|).  Output tracks keep these
# (x0, tx, err_x0, err_tx) from the XZ stage; their X hits lead the hit list.
F_P1X, F_P1Y, F_P1Z, F_DELTAPX, F_DELTAPY, F_DELTAPZ, F_DP1X, F_DP1Y, F_DP1Z = range(16, 25)
X_DETS = (15, 16, 21, 22, 27, 28)
XZ_QUANTITIES = ["x0", "tx", "err_x0", "err_tx"]
XZ_FLOORS = {"x0": 1.0, "tx": 1e-3, "err_x0": 1e-3, "err_tx": 1e-6}


def xz_fit_fp64(det, elem, drift, sign, mask, planes):
    """float64 solve of the reference's XZ fit per row (arrays (T, 4))."""
    e1 = elem.astype(np.float64) - 1.0
    p1x = planes[F_P1X][det] + planes[F_DP1X][det] * e1
    p1y = planes[F_P1Y][det] + planes[F_DP1Y][det] * e1
    p1z = planes[F_P1Z][det] + planes[F_DP1Z][det] * e1
    dpx, dpy, dpz = planes[F_DELTAPX][det], planes[F_DELTAPY][det], planes[F_DELTAPZ][det]
    x = p1x + (-p1y * dpx) / dpy + drift * sign
    z = p1z + (-p1y * dpz) / dpy
    w = np.where(mask, planes[F_RESOLUTION][det].astype(np.float64), 0.0)
    a00, a01, a11 = w.sum(1), (w * z).sum(1), (w * z * z).sum(1)
    b0, b1 = (w * x).sum(1), (w * x * z).sum(1)
    det2 = a00 * a11 - a01 * a01
    i00, i01, i11 = a11 / det2, -a01 / det2, a00 / det2
    return np.stack([i00 * b0 + i01 * b1, i01 * b0 + i11 * b1, np.sqrt(np.abs(i00)), np.sqrt(np.abs(i11))], axis=1)


def xz_error_report(ev, res, geo) -> dict:
    """Per-quantity error of the output tracks' XZ fit (x0, tx, err_x0, err_tx)
    against the float64 solve of the same problem."""
    planes, _ = geo.export()
    planes = planes.astype(np.float64)
    nt = res.n_tracks.astype(np.int64)
    valid = np.arange(res.tracks.shape[1])[None, :] < nt[:, None]
    ev_of = np.broadcast_to(np.arange(len(nt), dtype=np.int64)[:, None], valid.shape)[valid]
    tr = res.tracks[valid]
    k = np.arange(4)[None, :]
    gidx = ev.offsets[ev_of][:, None] + tr["hit_index"][:, :4].astype(np.int64)
    det = ev.det[gidx].astype(np.int64)
    mask = np.isin(det, X_DETS) & (k < tr["nhits"][:, None])
    det = np.where(mask, det, 16)
    elem = np.where(mask, ev.elem[gidx].astype(np.int64), 1)
    drift = np.where(mask, np.abs(ev.drift[gidx].astype(np.float64)), 0.0)
    sign = (((tr["sign_plus"][:, None] >> k) & 1).astype(np.float64) -
            ((tr["sign_minus"][:, None] >> k) & 1).astype(np.float64))
    ref = xz_fit_fp64(det, elem, drift, sign, mask, planes)
    dev = np.stack([tr["x0"], tr["tx"], tr["err_x0"], tr["err_tx"]], axis=1).astype(np.float64)
    out = {"tracks": int(len(dev)), "floors": XZ_FLOORS,
           "fit": "fit_2D_track_drift, FP32 in the chain vs float64 (trackanalyticalminimizer.cuh:78-119)"}
    for i, q in enumerate(XZ_QUANTITIES):
        err = np.abs(dev[:, i] - ref[:, i]) / np.maximum(np.abs(ref[:, i]), XZ_FLOORS[q])
        out[q] = {"max_rel": float(err.max()) if len(err) else 0.0,
                  "p99_rel": float(np.quantile(err, 0.99)) if len(err) else 0.0,
                  "median_rel": float(np.median(err)) if len(err) else 0.0}
    return out
