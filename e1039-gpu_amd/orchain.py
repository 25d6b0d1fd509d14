"""Python host side of the MI355X online-reconstruction chain.

Thin ctypes layer over libe1039or.so (C ABI in include/e1039_or.h and
include/e1039_or_gen.h).  It mirrors the reference's staging/launch sequence
(OnlineReconstruction/online_reconstruction.cu:776-894) as a library call:

    geo = Geometry.load(DEFAULT_GEOMETRY)
    sess = Session(geo)                      # device 0, reference caps
    ev = generate(geo, 10_000, seed=0xE1039)  # seeded synthetic events
    res = sess.reconstruct(ev)               # status, keep bits, tracks

Device buffers come from torch (plumbing only); all arithmetic runs in the
HIP kernels of the library.  There is no CPU fallback: a missing library or
device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libe1039or.so")
DEFAULT_GEOMETRY = os.path.join(HERE, "data", "basic_geometry_e906.txt")

N_DETECTORS = 55
MAX_ELEM = 404
TRACK_PARAMS = 148
MAX_HITS_PER_TRACK = 18

# per-event status bits (include/e1039_or.h)
EV_D0_MULT = 1 << 0
EV_D2_MULT = 1 << 1
EV_D3P_MULT = 1 << 2
EV_D3M_MULT = 1 << 3
EV_XZ_BIN_OVERFLOW = 1 << 4
EV_TRACKLET_OVERFLOW = 1 << 5
EV_PLANE_OVERFLOW = 1 << 6
EV_PAIR_OVERFLOW = 1 << 7
EV_CLUSTER_OVERFLOW = 1 << 8
EV_EVENT_TOO_LARGE = 1 << 9
EV_BAD_HIT = 1 << 10
EV_OUTPUT_OVERFLOW = 1 << 11
EV_REJECT_MASK = 0x7FF
TUNE_YZ_LIST_CAP = 1
TUNE_GRID_PERCENT = 2
TUNE_GRAPH = 3
TUNE_WAVES_PER_EVENT = 4
NOALLOC_MAX_EVENTS = 262144     # OR_NOALLOC_MAX_EVENTS
PRUNE_WEIGHTS = 5313       # OR_PRUNE_WEIGHTS: w1[64][16], b1[64], w2[64][64], b2[64], w3[64], b3
PRUNE_FEATURES = 16
UNIT_ROW = 20              # OR_UNIT_ROW: event, tracklet_id, unit, winner flag, 16 features
DEFAULT_PRUNER = os.path.join(HERE, "data", "pruner_u3v3.npz")
DEFAULT_CALIBRATION = os.path.join(HERE, "data", "calibration_e906.txt")
RT_MAX_KNOTS = 512                                                      # OR_RT_MAX_KNOTS
FIT_FIELDS = ["x0", "tx", "y0", "ty", "chi2", "ndf", "cond", "fp64"]   # OR_FIT_FIELDS per output track
FIT_COND_FP32 = 1.0e4                                                   # OR_FIT_COND_FP32

PARAM_FIELDS = [
    "max_hits_chamber_plane", "max_hits_hodo_plane", "max_hits_prop_plane", "max_reduced_per_plane",
    "max_pair_plane_hits", "max_pairs_per_bin", "max_pairs_window", "max_tracklets_per_slot",
    "max_tracklets_per_event", "max_d0_mult", "max_d2_mult", "max_d3_mult", "max_tracks_out",
    "max_hits_per_event", "full_chain",
]


class OrParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in PARAM_FIELDS]

    def as_dict(self):
        return {n: getattr(self, n) for n in PARAM_FIELDS}


class OrHits(C.Structure):
    _fields_ = [("n_events", C.c_int64), ("n_hits", C.c_int64), ("det", C.c_void_p), ("elem", C.c_void_p),
                ("tdc", C.c_void_p), ("drift", C.c_void_p), ("flag", C.c_void_p), ("event_offsets", C.c_void_p),
                ("on_device", C.c_int32), ("event_id", C.c_void_p)]


class OrResult(C.Structure):
    _fields_ = [("status", C.c_void_p), ("keep_mask", C.c_void_p), ("n_tracks", C.c_void_p), ("tracks", C.c_void_p),
                ("n_tracklets", C.c_void_p), ("records", C.c_void_p), ("vertex", C.c_void_p), ("fit", C.c_void_p),
                ("on_device", C.c_int32)]


class OrGenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_tracks_min", C.c_int32), ("n_tracks_max", C.c_int32),
                ("tx_max", C.c_float), ("ty_max", C.c_float), ("invp_min", C.c_float), ("invp_max", C.c_float),
                ("drift_sigma", C.c_float), ("noise_occupancy", C.c_float), ("afterpulse_frac", C.c_float),
                ("mult_mode", C.c_int32), ("mult_mean", C.c_float), ("mult_mean2", C.c_float),
                ("mult_frac2", C.c_float), ("mult_cap", C.c_int32)]


class OrGenEvents(C.Structure):
    _fields_ = [("n_events", C.c_int64), ("n_hits", C.c_int64), ("det", C.c_void_p), ("elem", C.c_void_p),
                ("tdc", C.c_void_p), ("drift", C.c_void_p), ("flag", C.c_void_p), ("event_offsets", C.c_void_p),
                ("n_true", C.c_void_p), ("truth", C.c_void_p)]


TRACK_DTYPE = np.dtype([
    ("x0", "<f4"), ("tx", "<f4"), ("y0", "<f4"), ("ty", "<f4"), ("invp", "<f4"),
    ("err_x0", "<f4"), ("err_tx", "<f4"), ("err_y0", "<f4"), ("err_ty", "<f4"), ("err_invp", "<f4"),
    ("chi2", "<f4"), ("hit_index", "<u2", (MAX_HITS_PER_TRACK,)), ("sign_plus", "<u4"), ("sign_minus", "<u4"),
    ("station_id", "i1"), ("nhits", "u1"), ("charge", "i1"), ("slot", "u1"), ("tracklet_id", "<u4"),
])
assert TRACK_DTYPE.itemsize == 96

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libe1039or.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (make -C e1039-gpu_amd/csrc)")
    # torch-ROCm ships its own HIP and HSA runtimes.  Importing torch first makes
    # the library's libamdhip64.so.7 dependency bind to torch's already-loaded
    # copy (same soname); loading the library first would bring in a second
    # runtime from /opt/rocm, and only one runtime per process can open the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    lib.or_params_default.argtypes = [C.POINTER(OrParams)]
    lib.or_geometry_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.or_geometry_load.restype = i32
    lib.or_geometry_free.argtypes = [vp]
    lib.or_geometry_export.argtypes = [vp, vp, vp]
    lib.or_geometry_export.restype = i32
    lib.or_session_create.argtypes = [vp, C.POINTER(OrParams), C.c_int, C.POINTER(vp)]
    lib.or_session_create.restype = i32
    lib.or_reconstruct.argtypes = [vp, C.POINTER(OrHits), C.POINTER(OrResult)]
    lib.or_reconstruct.restype = i32
    lib.or_reconstruct_noalloc.argtypes = [vp, C.POINTER(OrHits), C.POINTER(OrResult)]
    lib.or_reconstruct_noalloc.restype = i32
    lib.or_session_sync.argtypes = [vp]
    lib.or_session_sync.restype = i32
    lib.or_session_set_stream.argtypes = [vp, vp]
    lib.or_session_set_stream.restype = i32
    lib.or_session_last_kernel_ms.argtypes = [vp]
    lib.or_session_last_kernel_ms.restype = C.c_float
    lib.or_session_set_tuning.argtypes = [vp, C.c_int, i64]
    lib.or_session_set_tuning.restype = i32
    lib.or_session_get_tuning.argtypes = [vp, C.c_int, C.POINTER(i64)]
    lib.or_session_get_tuning.restype = i32
    lib.or_session_autotune.argtypes = [vp, C.POINTER(OrHits), C.POINTER(OrResult), C.c_int, C.POINTER(C.c_float)]
    lib.or_session_autotune.restype = i32
    lib.or_session_set_pruner.argtypes = [vp, C.POINTER(C.c_float), i64, C.c_float]
    lib.or_session_set_pruner.restype = i32
    lib.or_session_set_pruner_margin.argtypes = [vp, C.c_float]
    lib.or_session_set_pruner_margin.restype = i32
    lib.or_pruner_score.argtypes = [vp, vp, i64, vp]
    lib.or_pruner_score.restype = i32
    lib.or_pruner_layers.argtypes = [vp, vp, i64, vp, vp, vp]
    lib.or_pruner_layers.restype = i32
    lib.or_session_set_unit_dump.argtypes = [vp, vp, i64, vp]
    lib.or_session_set_unit_dump.restype = i32
    lib.or_calibration_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.or_calibration_load.restype = i32
    lib.or_calibration_create.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]
    lib.or_calibration_create.restype = i32
    lib.or_calibration_free.argtypes = [vp]
    lib.or_session_set_calibration.argtypes = [vp, vp]
    lib.or_session_set_calibration.restype = i32
    lib.or_rt_calibrate.argtypes = [vp, C.POINTER(OrHits), vp]
    lib.or_rt_calibrate.restype = i32
    lib.or_dimuons.argtypes = [vp, C.POINTER(OrHits), C.POINTER(OrResult), vp, C.POINTER(OrDimuonResult)]
    lib.or_dimuons.restype = i32
    lib.or_session_set_fit_cond_bound.argtypes = [vp, C.c_float]
    lib.or_session_set_fit_cond_bound.restype = i32
    lib.or_output_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.or_output_open.restype = i32
    lib.or_output_fill.argtypes = [vp, vp, C.POINTER(OrHits), C.POINTER(OrResult), i32, vp, i32]
    lib.or_output_fill.restype = i32
    lib.or_output_close.argtypes = [vp]
    lib.or_output_close.restype = i32
    lib.or_session_destroy.argtypes = [vp]
    lib.or_status_string.argtypes = [i32]
    lib.or_status_string.restype = C.c_char_p
    lib.or_gen_params_default.argtypes = [C.POINTER(OrGenParams)]
    lib.or_generate.argtypes = [vp, C.POINTER(OrGenParams), i64, C.c_int, C.POINTER(OrGenEvents)]
    lib.or_generate.restype = i32
    lib.or_gen_free.argtypes = [C.POINTER(OrGenEvents)]
    _lib = lib
    return lib


class OrError(RuntimeError):
    pass


def _check(st: int):
    if st != 0:
        raise OrError(f"libe1039or: {load_library().or_status_string(st).decode()} (status {st})")


def default_params(**overrides) -> OrParams:
    p = OrParams()
    load_library().or_params_default(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, int(v))
    return p


# Caps for high-occupancy workloads (hundreds to thousands of hits per event):
# the reference's staging, multiplicity, pair-list and tracklet-storage caps
# (reconstruction_constants.h:154,173-191; reconstruction_kernels.cuh:16,225,896)
# lifted to the bounds of the library's index encodings, so that the plain
# method runs on every event instead of rejecting it.
HIGH_OCCUPANCY_CAPS = dict(
    max_hits_chamber_plane=400, max_hits_hodo_plane=400, max_hits_prop_plane=400,
    max_reduced_per_plane=254, max_pair_plane_hits=254, max_pairs_per_bin=255, max_pairs_window=255,
    max_tracklets_per_slot=16384, max_tracklets_per_event=2048, max_d0_mult=1 << 30, max_d2_mult=1 << 30,
    max_d3_mult=1 << 30, max_tracks_out=64, max_hits_per_event=8192,
)


def high_occupancy_params(**overrides) -> OrParams:
    """or_params with HIGH_OCCUPANCY_CAPS, then overrides."""
    return default_params(**dict(HIGH_OCCUPANCY_CAPS, **overrides))


class Geometry:
    """Parsed geometry + wire table (or_geometry_load)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def load(cls, path: str = DEFAULT_GEOMETRY) -> "Geometry":
        h = C.c_void_p()
        _check(load_library().or_geometry_load(path.encode(), C.byref(h)))
        return cls(h)

    def export(self):
        planes = np.zeros((27, N_DETECTORS), np.float32)
        wire = np.zeros((N_DETECTORS, MAX_ELEM), np.float32)
        _check(load_library().or_geometry_export(self.handle, planes.ctypes.data, wire.ctypes.data))
        return planes, wire

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.or_geometry_free(self.handle)
            self.handle = None


class OrDimuonResult(C.Structure):
    _fields_ = [("n_dimuons", C.c_void_p), ("dimuons", C.c_void_p), ("histograms", C.c_void_p), ("lost", C.c_void_p)]


DIMUON_FIELDS = 36        # OR_DIMUON_FIELDS
DIMUONS_MAX = 16          # OR_DIMUONS_MAX
HIST_TRIGGERS, HIST_VARS, HIST_BINS = 11, 18, 256


@dataclass
class DimuonResult:
    n_dimuons: np.ndarray          # (n_events,) int32
    dimuons: np.ndarray            # (n_events, DIMUONS_MAX, DIMUON_FIELDS) float32
    histograms: np.ndarray | None  # (HIST_TRIGGERS * HIST_VARS * HIST_BINS,) float32
    lost: int


class Calibration:
    """R-T calibration (or_calibration_load / or_calibration_create)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def load(cls, path: str = DEFAULT_CALIBRATION) -> "Calibration":
        h = C.c_void_p()
        _check(load_library().or_calibration_load(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, tables: dict) -> "Calibration":
        """tables: detector id -> (tmin, tmax, knot times, values)."""
        ids = np.array(sorted(tables), np.int32)
        nk = np.array([len(tables[d][2]) for d in ids], np.int32)
        tmin = np.array([tables[d][0] for d in ids], np.float32)
        tmax = np.array([tables[d][1] for d in ids], np.float32)
        t = np.ascontiguousarray(np.concatenate([np.asarray(tables[d][2], np.float64) for d in ids])
                                 if len(ids) else np.zeros(1), np.float64)
        r = np.ascontiguousarray(np.concatenate([np.asarray(tables[d][3], np.float64) for d in ids])
                                 if len(ids) else np.zeros(1), np.float64)
        h = C.c_void_p()
        _check(load_library().or_calibration_create(len(ids), ids.ctypes.data, nk.ctypes.data, tmin.ctypes.data,
                                                    tmax.ctypes.data, t.ctypes.data, r.ctypes.data, C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.or_calibration_free(self.handle)
            self.handle = None


@dataclass
class Events:
    """Hits in the boundary layout: SoA, events contiguous (or_hits)."""
    det: np.ndarray        # int16
    elem: np.ndarray       # int16
    tdc: np.ndarray        # float32
    drift: np.ndarray      # float32
    flag: np.ndarray | None  # uint8 or None (all in time)
    offsets: np.ndarray    # int64, n_events + 1
    n_true: np.ndarray | None = None
    truth: np.ndarray | None = None

    @property
    def n_events(self) -> int:
        return len(self.offsets) - 1

    @property
    def n_hits(self) -> int:
        return int(self.offsets[-1])

    def subset(self, i0: int, i1: int) -> "Events":
        h0, h1 = int(self.offsets[i0]), int(self.offsets[i1])
        return Events(self.det[h0:h1].copy(), self.elem[h0:h1].copy(), self.tdc[h0:h1].copy(),
                      self.drift[h0:h1].copy(), None if self.flag is None else self.flag[h0:h1].copy(),
                      (self.offsets[i0:i1 + 1] - h0).astype(np.int64),
                      None if self.n_true is None else self.n_true[i0:i1].copy(),
                      None if self.truth is None else self.truth[i0:i1].copy())


def gen_params(**kw) -> OrGenParams:
    p = OrGenParams()
    load_library().or_gen_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def generate(geo: Geometry, n_events: int, threads: int = 0, **kw) -> Events:
    """Seeded synthetic events (include/e1039_or_gen.h)."""
    lib = load_library()
    p = gen_params(**kw)
    out = OrGenEvents()
    _check(lib.or_generate(geo.handle, C.byref(p), int(n_events), int(threads), C.byref(out)))
    try:
        nh, ne = out.n_hits, out.n_events

        def arr(ptr, dtype, n):
            if n == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), (n,)).copy()

        ev = Events(arr(out.det, np.int16, nh), arr(out.elem, np.int16, nh), arr(out.tdc, np.float32, nh),
                    arr(out.drift, np.float32, nh), arr(out.flag, np.uint8, nh),
                    arr(out.event_offsets, np.int64, ne + 1), arr(out.n_true, np.int8, ne),
                    arr(out.truth, np.float32, ne * 4 * 6).reshape(ne, 4, 6))
    finally:
        lib.or_gen_free(C.byref(out))
    return ev


@dataclass
class Result:
    status: np.ndarray       # uint32 per event
    keep_mask: np.ndarray    # uint32 words, bit h = hit h kept
    n_tracks: np.ndarray     # int32 per event
    tracks: np.ndarray       # TRACK_DTYPE, (n_events, max_tracks_out)
    n_tracklets: np.ndarray | None = None
    records: np.ndarray | None = None  # float32 (n_events, max_tracklets_per_event, 148)
    vertex: np.ndarray | None = None   # float32 (n_events, max_tracks_out, 6): vx, vy, vz, px, py, pz
    fit: np.ndarray | None = None      # float32 (n_events, max_tracks_out, 8): final fit (FIT_FIELDS)

    def valid_tracks(self) -> np.ndarray:
        """Tracks of all events, slots [0, n_tracks[e]) of each event in event
        order; the library leaves the other slots unwritten."""
        k = np.arange(self.tracks.shape[1])[None, :] < self.n_tracks[:, None]
        return np.ascontiguousarray(self.tracks[k])

    def keep_bits(self, n_hits: int) -> np.ndarray:
        b = np.unpackbits(self.keep_mask.view(np.uint8), bitorder="little")
        return b[:n_hits].astype(bool)


class Session:
    """A device session (or_session_create) bound to torch's current stream."""

    def __init__(self, geo: Geometry, params: OrParams | None = None, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise OrError("no GPU visible: the chain runs only on the device")
        self.lib = load_library()
        self.params = params if params is not None else default_params()
        self.device = device
        h = C.c_void_p()
        _check(self.lib.or_session_create(geo.handle, C.byref(self.params), device, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.or_session_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def last_kernel_ms(self) -> float:
        return float(self.lib.or_session_last_kernel_ms(self.handle))

    def set_tuning(self, key: int, value: int):
        """or_session_set_tuning; key e.g. TUNE_YZ_LIST_CAP."""
        _check(self.lib.or_session_set_tuning(self.handle, key, value))

    def get_tuning(self, key: int) -> int:
        """or_session_get_tuning."""
        v = C.c_int64(0)
        _check(self.lib.or_session_get_tuning(self.handle, key, C.byref(v)))
        return int(v.value)

    def set_calibration(self, cal: "Calibration | None"):
        """or_session_set_calibration: the chain derives drift from tdc for calibrated detectors (None: off)."""
        _check(self.lib.or_session_set_calibration(self.handle, None if cal is None else cal.handle))

    def rt_calibrate(self, ev: "Events") -> np.ndarray:
        """or_rt_calibrate on host buffers: the drift of every hit under the session's calibration."""
        cont = [np.ascontiguousarray(a) for a in (ev.det, ev.elem, ev.tdc, ev.drift, ev.offsets)]
        hits = OrHits(ev.n_events, ev.n_hits, cont[0].ctypes.data, cont[1].ctypes.data, cont[2].ctypes.data,
                      cont[3].ctypes.data, None, cont[4].ctypes.data, 0)
        out = np.zeros(ev.n_hits, np.float32)
        _check(self.lib.or_rt_calibrate(self.handle, C.byref(hits), out.ctypes.data))
        return out

    def rt_calibrate_device(self, t):
        """or_rt_calibrate on a device-resident batch (upload()); returns a device tensor (asynchronous)."""
        import torch
        stream = torch.cuda.current_stream(self.device)
        _check(self.lib.or_session_set_stream(self.handle, C.c_void_p(stream.cuda_stream)))
        hits = OrHits(t["n_events"], t["n_hits"], t["det"].data_ptr(), t["elem"].data_ptr(), t["tdc"].data_ptr(),
                      t["drift"].data_ptr(), None, t["offsets"].data_ptr(), 1)
        out = torch.empty(max(1, t["n_hits"]), dtype=torch.float32, device=t["det"].device)
        _check(self.lib.or_rt_calibrate(self.handle, C.byref(hits), C.c_void_p(out.data_ptr())))
        return out[: t["n_hits"]]

    # --- dimuons and display histograms ------------------------------------------
    def alloc_dimuons(self, n_events: int, histograms: bool = True):
        import torch
        dev = torch.device("cuda", self.device)
        return {
            "n_dimuons": torch.empty(max(1, n_events), dtype=torch.int32, device=dev),
            "dimuons": torch.empty(max(1, n_events * DIMUONS_MAX * DIMUON_FIELDS), dtype=torch.float32, device=dev),
            "histograms": (torch.zeros(HIST_TRIGGERS * HIST_VARS * HIST_BINS, dtype=torch.float32, device=dev)
                           if histograms else None),
            "lost": torch.zeros(1, dtype=torch.int64, device=dev),
        }

    def run_dimuons(self, t, o, d, trigger=None):
        """or_dimuons on device buffers: t from upload(), o the outputs of a full-chain run with vertex,
        d from alloc_dimuons(), trigger an (n_events,) int32 device tensor or None (asynchronous)."""
        hits, res = self._device_args(t, o)
        out = OrDimuonResult(d["n_dimuons"].data_ptr(), d["dimuons"].data_ptr(),
                             None if d["histograms"] is None else d["histograms"].data_ptr(), d["lost"].data_ptr())
        trig = None if trigger is None else C.c_void_p(trigger.contiguous().data_ptr())
        _check(self.lib.or_dimuons(self.handle, C.byref(hits), C.byref(res), trig, C.byref(out)))

    @staticmethod
    def fetch_dimuons(d, n_events: int) -> DimuonResult:
        return DimuonResult(
            n_dimuons=d["n_dimuons"].cpu().numpy()[:n_events],
            dimuons=d["dimuons"].cpu().numpy()[: n_events * DIMUONS_MAX * DIMUON_FIELDS].reshape(
                n_events, DIMUONS_MAX, DIMUON_FIELDS),
            histograms=None if d["histograms"] is None else d["histograms"].cpu().numpy(),
            lost=int(d["lost"].cpu().numpy()[0]),
        )

    def set_fit_cond_bound(self, bound: float):
        """or_session_set_fit_cond_bound: condition bound of the final fit's FP32 solve
        (0 sends every track to the FP64 path; per device, shared by its sessions)."""
        _check(self.lib.or_session_set_fit_cond_bound(self.handle, float(bound)))

    # --- learned YZ unit pruning -----------------------------------------------
    def set_pruner(self, weights: np.ndarray | None, threshold: float = 0.0):
        """or_session_set_pruner: flat fp32 weights (PRUNE_WEIGHTS), or None for the plain chain."""
        if weights is None:
            _check(self.lib.or_session_set_pruner(self.handle, None, 0, 0.0))
            return
        w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
        _check(self.lib.or_session_set_pruner(self.handle, w.ctypes.data_as(C.POINTER(C.c_float)), w.size,
                                              float(threshold)))

    def set_pruner_margin(self, margin: float = float("inf")):
        """or_session_set_pruner_margin: keep units within margin of their tracklet's best score."""
        _check(self.lib.or_session_set_pruner_margin(self.handle, float(margin)))

    def score_units(self, features):
        """or_pruner_score on a (n, 16) float32 device tensor; returns the (n,) scores."""
        import torch
        f = features.contiguous()
        out = torch.empty(f.shape[0], dtype=torch.float32, device=f.device)
        stream = torch.cuda.current_stream(self.device)
        _check(self.lib.or_session_set_stream(self.handle, C.c_void_p(stream.cuda_stream)))
        _check(self.lib.or_pruner_score(self.handle, C.c_void_p(f.data_ptr()), f.shape[0], C.c_void_p(out.data_ptr())))
        return out

    def score_layers(self, features):
        """or_pruner_layers: (scores, layer-1 pre-activations, layer-2 pre-activations), (n,), (n, 64), (n, 64)."""
        import torch
        f = features.contiguous()
        n = f.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=f.device)
        p1 = torch.empty((n, 64), dtype=torch.float32, device=f.device)
        p2 = torch.empty((n, 64), dtype=torch.float32, device=f.device)
        stream = torch.cuda.current_stream(self.device)
        _check(self.lib.or_session_set_stream(self.handle, C.c_void_p(stream.cuda_stream)))
        _check(self.lib.or_pruner_layers(self.handle, C.c_void_p(f.data_ptr()), n, C.c_void_p(out.data_ptr()),
                                         C.c_void_p(p1.data_ptr()), C.c_void_p(p2.data_ptr())))
        return out, p1, p2

    def set_unit_dump(self, rows=None, count=None):
        """or_session_set_unit_dump with a (cap, UNIT_ROW) float32 and a (1,) int64 device tensor (zero it first)."""
        if rows is None:
            _check(self.lib.or_session_set_unit_dump(self.handle, None, 0, None))
            return
        _check(self.lib.or_session_set_unit_dump(self.handle, C.c_void_p(rows.data_ptr()), rows.shape[0],
                                                 C.c_void_p(count.data_ptr())))

    # --- device-resident path -------------------------------------------------
    def upload(self, ev: Events):
        import torch
        dev = torch.device("cuda", self.device)
        t = {
            "det": torch.from_numpy(ev.det).to(dev),
            "elem": torch.from_numpy(ev.elem).to(dev),
            "tdc": torch.from_numpy(ev.tdc).to(dev),
            "drift": torch.from_numpy(ev.drift).to(dev),
            "flag": None if ev.flag is None else torch.from_numpy(ev.flag).to(dev),
            "offsets": torch.from_numpy(ev.offsets).to(dev),
            "n_events": ev.n_events, "n_hits": ev.n_hits,
        }
        return t

    def alloc_outputs(self, n_events: int, n_hits: int, tracklets: bool = False, records: bool = False,
                      vertex: bool = False, fit: bool = False):
        import torch
        dev = torch.device("cuda", self.device)
        p = self.params
        o = {
            "status": torch.empty(n_events, dtype=torch.int32, device=dev),
            "keep_mask": torch.empty(max(1, (n_hits + 31) // 32), dtype=torch.int32, device=dev),
            "n_tracks": torch.empty(n_events, dtype=torch.int32, device=dev),
            "tracks": torch.empty(max(1, n_events * p.max_tracks_out * 96), dtype=torch.uint8, device=dev),
            "n_tracklets": torch.empty(n_events, dtype=torch.int32, device=dev) if tracklets else None,
            "records": (torch.empty(n_events * p.max_tracklets_per_event * TRACK_PARAMS, dtype=torch.float32,
                                    device=dev) if records else None),
            "vertex": (torch.empty(max(1, n_events * p.max_tracks_out * 6), dtype=torch.float32, device=dev)
                       if vertex else None),
            "fit": (torch.empty(max(1, n_events * p.max_tracks_out * len(FIT_FIELDS)), dtype=torch.float32,
                                device=dev) if fit else None),
        }
        return o

    def _device_args(self, t, o):
        import torch
        stream = torch.cuda.current_stream(self.device)
        _check(self.lib.or_session_set_stream(self.handle, C.c_void_p(stream.cuda_stream)))
        hits = OrHits(t["n_events"], t["n_hits"], t["det"].data_ptr(), t["elem"].data_ptr(), t["tdc"].data_ptr(),
                      t["drift"].data_ptr(), None if t["flag"] is None else t["flag"].data_ptr(),
                      t["offsets"].data_ptr(), 1)
        res = OrResult(o["status"].data_ptr(), o["keep_mask"].data_ptr(), o["n_tracks"].data_ptr(),
                       o["tracks"].data_ptr(), None if o["n_tracklets"] is None else o["n_tracklets"].data_ptr(),
                       None if o["records"] is None else o["records"].data_ptr(),
                       None if o.get("vertex") is None else o["vertex"].data_ptr(),
                       None if o.get("fit") is None else o["fit"].data_ptr(), 1)
        return hits, res

    def run_device(self, t, o):
        """Device call on torch's current stream: inputs and outputs already in HBM (asynchronous)."""
        hits, res = self._device_args(t, o)
        _check(self.lib.or_reconstruct(self.handle, C.byref(hits), C.byref(res)))

    def run_device_noalloc(self, t, o):
        """or_reconstruct_noalloc: the device call without any device allocation (fixed per-device scratch)."""
        hits, res = self._device_args(t, o)
        _check(self.lib.or_reconstruct_noalloc(self.handle, C.byref(hits), C.byref(res)))

    def autotune(self, t, o, reps: int = 2) -> float:
        """or_session_autotune on a device-resident batch; returns the best mean chain time in ms."""
        hits, res = self._device_args(t, o)
        best = C.c_float(0.0)
        _check(self.lib.or_session_autotune(self.handle, C.byref(hits), C.byref(res), reps, C.byref(best)))
        return float(best.value)

    @staticmethod
    def fetch(o, n_events: int, max_tracks_out: int, max_tkl: int) -> Result:
        tr = o["tracks"].cpu().numpy().view(TRACK_DTYPE)[: n_events * max_tracks_out].reshape(n_events, max_tracks_out)
        return Result(
            status=o["status"].cpu().numpy().view(np.uint32),
            keep_mask=o["keep_mask"].cpu().numpy().view(np.uint32),
            n_tracks=o["n_tracks"].cpu().numpy(),
            tracks=tr,
            n_tracklets=None if o["n_tracklets"] is None else o["n_tracklets"].cpu().numpy(),
            records=(None if o["records"] is None else
                     o["records"].cpu().numpy().reshape(n_events, max_tkl, TRACK_PARAMS)),
            vertex=(None if o.get("vertex") is None else
                    o["vertex"].cpu().numpy()[: n_events * max_tracks_out * 6].reshape(n_events, max_tracks_out, 6)),
            fit=(None if o.get("fit") is None else
                 o["fit"].cpu().numpy()[: n_events * max_tracks_out * len(FIT_FIELDS)].reshape(
                     n_events, max_tracks_out, len(FIT_FIELDS))),
        )

    def reconstruct(self, ev: Events, tracklets: bool = True, records: bool = False, vertex: bool = False,
                    fit: bool = False) -> Result:
        """Uploads, runs on the device, downloads."""
        import torch
        t = self.upload(ev)
        o = self.alloc_outputs(ev.n_events, ev.n_hits, tracklets, records, vertex, fit)
        self.run_device(t, o)
        torch.cuda.synchronize(self.device)
        return self.fetch(o, ev.n_events, self.params.max_tracks_out, self.params.max_tracklets_per_event)

    def reconstruct_host(self, ev: Events, records: bool = False, vertex: bool = False, fit: bool = False) -> Result:
        """Host buffers in and out; the library stages through device memory."""
        p = self.params
        ne, nh = ev.n_events, ev.n_hits
        status = np.zeros(ne, np.uint32)
        keep = np.zeros(max(1, (nh + 31) // 32), np.uint32)
        ntr = np.zeros(ne, np.int32)
        tracks = np.zeros((ne, p.max_tracks_out), TRACK_DTYPE)
        ntkl = np.zeros(ne, np.int32)
        rec = np.zeros((ne, p.max_tracklets_per_event, TRACK_PARAMS), np.float32) if records else None
        vtx = np.zeros((ne, p.max_tracks_out, 6), np.float32) if vertex else None
        fitv = np.zeros((ne, p.max_tracks_out, len(FIT_FIELDS)), np.float32) if fit else None
        cont = [np.ascontiguousarray(a) for a in (ev.det, ev.elem, ev.tdc, ev.drift, ev.offsets)]
        flag = None if ev.flag is None else np.ascontiguousarray(ev.flag)
        hits = OrHits(ne, nh, cont[0].ctypes.data, cont[1].ctypes.data, cont[2].ctypes.data, cont[3].ctypes.data,
                      None if flag is None else flag.ctypes.data, cont[4].ctypes.data, 0)
        res = OrResult(status.ctypes.data, keep.ctypes.data, ntr.ctypes.data, tracks.ctypes.data, ntkl.ctypes.data,
                       None if rec is None else rec.ctypes.data, None if vtx is None else vtx.ctypes.data,
                       None if fitv is None else fitv.ctypes.data, 0)
        _check(self.lib.or_reconstruct(self.handle, C.byref(hits), C.byref(res)))
        return Result(status, keep, ntr, tracks, ntkl, rec, vtx, fitv)


# ---- columnar event output (or_output_*, the reference's OROutput tree) -------

OUTPUT_MIN_STATION_ID = 7   # the reference writes tracks of stationID >= 7 (online_reconstruction.cu:1132)


class OutputWriter:
    """or_output_open / or_output_fill / or_output_close: one .npz archive with
    one array per OROutput branch (OROutput.cxx:24-102), filled from host-side
    results batch by batch.  Use as a context manager or call close()."""

    def __init__(self, path: str):
        self.lib = load_library()
        self.handle = C.c_void_p()
        _check(self.lib.or_output_open(str(path).encode(), C.byref(self.handle)))

    def fill(self, geo: Geometry, ev: Events, res: Result, max_tracklets_per_event: int,
             dim: "DimuonResult | None" = None, event_id=None, min_station_id: int = OUTPUT_MIN_STATION_ID):
        if res.records is None or res.n_tracklets is None:
            raise ValueError("the writer needs the tracklet records and counts (records=True)")
        keep = []   # arrays referenced by pointer until the call returns
        def ptr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data
        hits = OrHits(ev.n_events, ev.n_hits, ptr(ev.det, np.int16), ptr(ev.elem, np.int16), ptr(ev.tdc, np.float32),
                      ptr(ev.drift, np.float32), None, ptr(ev.offsets, np.int64), 0,
                      None if event_id is None else ptr(event_id, np.int32))
        out = OrResult(ptr(res.status, np.uint32), ptr(res.keep_mask, np.uint32), None, None,
                       ptr(res.n_tracklets, np.int32), ptr(res.records, np.float32), None, None, 0)
        dptr = None
        if dim is not None:
            d = OrDimuonResult(ptr(dim.n_dimuons, np.int32), ptr(dim.dimuons, np.float32), None, None)
            keep.append(d)
            dptr = C.addressof(d)
        _check(self.lib.or_output_fill(self.handle, geo.handle, C.byref(hits), C.byref(out),
                                       int(max_tracklets_per_event), dptr, int(min_station_id)))

    def close(self):
        if self.handle:
            h, self.handle = self.handle, C.c_void_p()
            _check(self.lib.or_output_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
