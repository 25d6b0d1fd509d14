"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (liboracle.so).

The oracle restates the reference chain on the host (or_oracle.cpp).  Only the
tests, the smoke check and bench.py's cpu_baseline leg load it, as the checker
or the timed CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, i32, f32 = C.c_void_p, C.c_int32, C.c_float
    lib.orc_geometry_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.orc_geometry_load.restype = i32
    lib.orc_geometry_free.argtypes = [vp]
    lib.orc_geometry_export.argtypes = [vp, vp, vp, vp]
    lib.orc_reconstruct.argtypes = [vp, vp, vp, vp, C.c_int]
    lib.orc_reconstruct.restype = i32
    lib.orc_fit2d.argtypes = [C.c_int, vp, vp, vp, vp]
    lib.orc_fit2d_drift.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp]
    lib.orc_matinv2.argtypes = [vp, vp]
    lib.orc_reduce_plane.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
    lib.orc_reduce_plane.restype = C.c_int
    lib.orc_dca.argtypes = [f32] * 10
    lib.orc_dca.restype = f32
    lib.orc_y_uvhit.argtypes = [vp, C.c_int, C.c_int, f32, C.c_int, f32, f32, C.POINTER(f32), C.POINTER(f32)]
    lib.orc_y_uvhit.restype = C.c_int
    lib.orc_xz_fy.argtypes = [vp, C.c_int, C.c_int, f32, C.c_int, f32, f32, C.POINTER(f32), C.POINTER(f32)]
    lib.orc_hodo_xz.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, f32, f32, f32, f32]
    lib.orc_hodo_xz.restype = C.c_int
    lib.orc_lr_xhits.argtypes = [vp, f32, f32, f32, f32, C.c_int, vp, vp, vp, vp, f32]
    lib.orc_lr_newhits.argtypes = [vp, f32, f32, f32, f32, f32, f32, f32, f32, C.c_int, vp, vp, vp, vp, f32]
    lib.orc_lr_single_xhits.argtypes = [vp, f32, f32, C.c_int, vp, vp, vp]
    lib.orc_lr_single_newhits.argtypes = [vp, f32, f32, f32, f32, C.c_int, vp, vp, vp]
    lib.orc_sagitta.argtypes = [vp, f32, f32, f32, f32, C.c_int, vp, vp]
    lib.orc_back_cleaning.argtypes = [C.c_int, vp, vp, vp]
    lib.orc_trace.argtypes = [C.c_int]
    lib.orc_trace_fetch.argtypes = [C.c_longlong, vp, vp, vp, vp, vp]
    lib.orc_trace_fetch.restype = C.c_longlong
    lib.orc_pairs_binned.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, vp, C.c_int, vp, vp]
    lib.orc_yz_state_check.argtypes = [C.c_int, vp]
    lib.orc_set_unit_filter.argtypes = [vp, C.c_int64]
    f32 = C.c_float
    lib.orc_prop_segment.argtypes = [vp, C.c_int, f32, vp, C.c_int, f32, vp, C.c_int, f32, vp, C.c_int, f32,
                                     f32, f32, f32, f32, f32, f32]
    lib.orc_prop_segment.restype = C.c_int
    lib.orc_global_cleaning.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp]
    lib.orc_dimuons.argtypes = [C.c_int64, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.orc_dimuon_swim.argtypes = [C.c_int, vp, vp, vp]
    lib.orc_rt_spline.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    lib.orc_rt_eval.argtypes = [C.c_int, vp, vp, C.c_int64, vp, vp]
    lib.orc_rt_calibrate.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, vp, vp, vp, vp]
    _lib = lib
    return lib


class OracleGeometry:
    def __init__(self, path: str):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.orc_geometry_load(path.encode(), C.byref(h))
        if st != 0:
            raise RuntimeError(f"oracle geometry load failed ({st})")
        self.handle = h

    def export(self):
        planes = np.zeros((27, 55), np.float32)
        wire = np.zeros((55, 404), np.float32)
        dw = np.zeros((55, 9), np.float32)
        self.lib.orc_geometry_export(self.handle, planes.ctypes.data, wire.ctypes.data, dw.ctypes.data)
        return planes, wire, dw

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.orc_geometry_free(self.handle)
            self.handle = None


def unit_keys(events, tracklet_ids, units) -> np.ndarray:
    """Sorted keys of searched YZ units for reconstruct(units=...): per unit its
    event index, tracklet_id and u3.v3 index over the window-list entries with
    an accepted hit (the index of the device's unit rows)."""
    k = (np.asarray(events, np.int64) << 36) | (np.asarray(tracklet_ids, np.int64) << 20) | np.asarray(units, np.int64)
    return np.unique(k)


def reconstruct(geo: OracleGeometry, params, ev, records: bool = False, threads: int = 0, vertex: bool = False,
                units=None):
    """Runs the oracle chain; returns the same Result layout as the product.

    units: optional sorted keys (unit_keys) restricting the YZ search of each
    listed tracklet to its listed units, as a pruned search leaves them;
    tracklets without a listed unit search all of theirs."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "e1039-gpu_amd"))
    import orchain as oc

    ne, nh = ev.n_events, ev.n_hits
    status = np.zeros(ne, np.uint32)
    keep = np.zeros(max(1, (nh + 31) // 32), np.uint32)
    ntr = np.zeros(ne, np.int32)
    tracks = np.zeros((ne, params.max_tracks_out), oc.TRACK_DTYPE)
    ntkl = np.zeros(ne, np.int32)
    rec = np.zeros((ne, params.max_tracklets_per_event, oc.TRACK_PARAMS), np.float32) if records else None
    vtx = np.zeros((ne, params.max_tracks_out, 6), np.float32) if vertex else None
    arrs = [np.ascontiguousarray(a) for a in (ev.det, ev.elem, ev.tdc, ev.drift, ev.offsets)]
    flag = None if ev.flag is None else np.ascontiguousarray(ev.flag)
    hits = oc.OrHits(ne, nh, arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data,
                     None if flag is None else flag.ctypes.data, arrs[4].ctypes.data, 0)
    res = oc.OrResult(status.ctypes.data, keep.ctypes.data, ntr.ctypes.data, tracks.ctypes.data, ntkl.ctypes.data,
                      None if rec is None else rec.ctypes.data, None if vtx is None else vtx.ctypes.data, None, 0)
    lib = load()
    keys = None if units is None else np.ascontiguousarray(units, np.int64)
    if keys is not None:
        lib.orc_set_unit_filter(keys.ctypes.data, int(keys.size))
    try:
        lib.orc_reconstruct(geo.handle, C.byref(params), C.byref(hits), C.byref(res), int(threads))
    finally:
        if keys is not None:
            lib.orc_set_unit_filter(None, 0)
    return oc.Result(status, keep, ntr, tracks, ntkl, rec, vtx)


# ---- R-T calibration (online_reconstruction.cu:220-242, 387-397) -------------

class OracleCalibration:
    """Per detector (tmin, tmax, T knots, R values) read from a calibration file
    as online_reconstruction.cu:226-242 reads it: a header line "detectorID nBin
    tmin tmax name", then nBin lines "iBin T R"; tmin/tmax kept as float32."""

    def __init__(self, tables: dict):
        self.tables = tables

    @classmethod
    def load(cls, path: str):
        tables = {}
        with open(path) as f:
            lines = f.read().split("\n")
        i = 0
        while i < len(lines):
            head = lines[i].split()
            i += 1
            if not head:
                continue
            d, n, t0, t1 = int(head[0]), int(head[1]), float(head[2]), float(head[3])
            rows = [lines[i + k].split() for k in range(n)]
            i += n
            tables[d] = (np.float32(t0), np.float32(t1), np.array([float(r[1]) for r in rows]),
                         np.array([float(r[2]) for r in rows]))
        return cls(tables)

    def flat(self):
        """nk[55], tmin[55], tmax[55], concatenated knots and values."""
        nk = np.zeros(55, np.int32)
        tmin = np.zeros(55, np.float32)
        tmax = np.zeros(55, np.float32)
        ts, rs = [], []
        for d in range(55):
            if d in self.tables:
                t0, t1, T, R = self.tables[d]
                nk[d], tmin[d], tmax[d] = len(T), t0, t1
                ts.append(T)
                rs.append(R)
        t = np.ascontiguousarray(np.concatenate(ts) if ts else np.zeros(1), np.float64)
        r = np.ascontiguousarray(np.concatenate(rs) if rs else np.zeros(1), np.float64)
        return nk, tmin, tmax, t, r


def rt_spline(T, R):
    """Coefficients (b, c, d) of the n-1 knot intervals of the "b1e1" spline."""
    T = np.ascontiguousarray(T, np.float64)
    R = np.ascontiguousarray(R, np.float64)
    n = len(T)
    b, c, d = (np.zeros(n - 1) for _ in range(3))
    load().orc_rt_spline(n, T.ctypes.data, R.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data)
    return b, c, d


def rt_eval(T, R, t):
    """The "b1e1" spline through (T, R) at the points t (TSpline3::Eval's interval choice)."""
    T = np.ascontiguousarray(T, np.float64)
    R = np.ascontiguousarray(R, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    out = np.zeros(len(t))
    load().orc_rt_eval(len(T), T.ctypes.data, R.ctypes.data, len(t), t.ctypes.data, out.ctypes.data)
    return out


def rt_calibrate(cal: OracleCalibration, cellwidth, det, tdc, drift):
    """Drift of every hit by the rule of online_reconstruction.cu:387-397."""
    nk, tmin, tmax, t, r = cal.flat()
    cw = np.ascontiguousarray(cellwidth, np.float32)
    det = np.ascontiguousarray(det, np.int16)
    tdc = np.ascontiguousarray(tdc, np.float32)
    drift = np.ascontiguousarray(drift, np.float32)
    out = np.zeros(len(det), np.float32)
    load().orc_rt_calibrate(nk.ctypes.data, tmin.ctypes.data, tmax.ctypes.data, t.ctypes.data, r.ctypes.data,
                            cw.ctypes.data, len(det), det.ctypes.data, tdc.ctypes.data, drift.ctypes.data,
                            out.ctypes.data)
    return out


# ---- dimuons and display histograms (reconstruction_kernels.cuh:2820-3195) ----

def dimuon_swim(a, b):
    """The dimuon closest-approach swim for pairs of tracks.  a, b: (n, 7) rows of
    (x0, tx, y0, ty, invp, charge, z of the most upstream hit), first and second
    track.  Returns (n, 10): vertex x, y, z; p_pos; p_neg; the z returned."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape and a.ndim == 2 and a.shape[1] == 7
    out = np.zeros((len(a), 10), np.float32)
    load().orc_dimuon_swim(len(a), a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def dimuons(chain, ev, z_array, max_out: int, trigger=None, histograms: bool = True):
    """Dimuons of a full-chain result with vertex (chain: the Result of reconstruct(..., vertex=True))."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "e1039-gpu_amd"))
    import orchain as oc
    ne = ev.n_events
    n_dim = np.zeros(ne, np.int32)
    dim = np.zeros((ne, oc.DIMUONS_MAX, oc.DIMUON_FIELDS), np.float32)
    hist = np.zeros(oc.HIST_TRIGGERS * oc.HIST_VARS * oc.HIST_BINS, np.float32) if histograms else None
    lost = np.zeros(1, np.int64)
    tracks = np.ascontiguousarray(chain.tracks)
    vtx = np.ascontiguousarray(chain.vertex, np.float32)
    det = np.ascontiguousarray(ev.det, np.int16)
    off = np.ascontiguousarray(ev.offsets, np.int64)
    z = np.ascontiguousarray(z_array, np.float32)
    status = np.ascontiguousarray(chain.status, np.uint32)
    ntr = np.ascontiguousarray(chain.n_tracks, np.int32)
    trig = None if trigger is None else np.ascontiguousarray(trigger, np.int32)
    load().orc_dimuons(ne, max_out, ntr.ctypes.data, status.ctypes.data, tracks.ctypes.data, vtx.ctypes.data,
                       det.ctypes.data, off.ctypes.data, z.ctypes.data, None if trig is None else trig.ctypes.data,
                       n_dim.ctypes.data, dim.ctypes.data, None if hist is None else hist.ctypes.data,
                       lost.ctypes.data)
    return oc.DimuonResult(n_dim, dim, hist, int(lost[0]))
