"""The C ABI library: loads on a host without a GPU, exports every function the
headers declare, reports a missing device as an error, and the generator is
deterministic."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("e1039_or.h", "e1039_or_gen.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(or_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function(oc):
    lib = oc.load_library()
    names = declared_functions()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", oc.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, f"declared but not exported: {missing}"
    for n in names:
        getattr(lib, n)


def test_status_strings(oc):
    lib = oc.load_library()
    for st in range(9):
        assert lib.or_status_string(st)
    assert b"unknown" in lib.or_status_string(99)


def test_default_params_are_reference_caps(oc):
    p = oc.default_params().as_dict()
    assert p["max_hits_chamber_plane"] == 200 and p["max_hits_hodo_plane"] == 60 and p["max_hits_prop_plane"] == 120
    assert p["max_pairs_per_bin"] == 20 and p["max_tracklets_per_event"] == 512
    assert (p["max_d0_mult"], p["max_d2_mult"], p["max_d3_mult"]) == (210, 120, 90)


def test_session_without_device_reports_error(oc, geo):
    import torch
    if torch.cuda.is_available():
        return
    lib = oc.load_library()
    h = C.c_void_p()
    p = oc.default_params()
    st = lib.or_session_create(geo.handle, C.byref(p), 0, C.byref(h))
    assert st == 6  # OR_ERR_NO_DEVICE


def test_generator_deterministic_and_thread_independent(oc, geo):
    a = oc.generate(geo, 3000, threads=1, seed=99, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    b = oc.generate(geo, 3000, threads=8, seed=99, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    for k in ("det", "elem", "tdc", "drift", "flag", "offsets", "n_true", "truth"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    c = oc.generate(geo, 3000, seed=100, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    assert not np.array_equal(a.elem[:1000], c.elem[:1000])
    counts = np.diff(a.offsets)
    assert 280 < counts.mean() < 320 and counts.max() <= 1000


def test_generator_layout(oc, geo):
    ev = oc.generate(geo, 500, seed=5)
    for e in range(ev.n_events):
        h0, h1 = ev.offsets[e], ev.offsets[e + 1]
        d, el = ev.det[h0:h1].astype(int), ev.elem[h0:h1].astype(int)
        key = d * 1000 + el
        assert (np.diff(key) >= 0).all()
        assert ((d >= 1) & (d <= 54) & ~((d >= 7) & (d <= 12))).all()
    assert (ev.flag == 2).all()
    assert 1.0 <= ev.n_true.mean() <= 2.0
