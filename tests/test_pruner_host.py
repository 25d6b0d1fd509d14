"""Host side of the learned unit pruner: weight layout, bf16 rounding, the
reference scoring against explicit loops, the shipped weights file, and the C
ABI's argument checks (no GPU needed)."""
import ctypes as C
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def pruner():
    import pruner as p
    return p


def test_pack_unpack_round_trip(pruner):
    rng = np.random.default_rng(0)
    parts = (rng.normal(size=(64, 16)), rng.normal(size=64), rng.normal(size=(64, 64)), rng.normal(size=64),
             rng.normal(size=64), 0.25)
    flat = pruner.pack(*parts)
    assert flat.size == pruner.NW == 5313
    back = pruner.unpack(flat)
    for a, b in zip(parts[:5], back[:5]):
        assert np.array_equal(np.asarray(a, np.float32), b)
    assert back[5] == np.float32(0.25)
    with pytest.raises(ValueError):
        pruner.unpack(flat[:-1])


def test_bf16_rounding(pruner):
    import torch
    x = torch.tensor([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.0e-3, 1.0 + 2**-9], dtype=torch.float32)
    got = pruner.bf16_round(x).tolist()
    # ties go to even; 1 + 2^-9 is below the halfway point
    assert got == [1.0, 1.0, 1.015625, -2.5, pytest.approx(3.0e-3, rel=4e-3), 1.0]


def test_reference_scoring_against_loops(pruner):
    import torch
    rng = np.random.default_rng(1)
    flat = pruner.pack(rng.normal(0, 0.5, (64, 16)), rng.normal(0, 0.2, 64), rng.normal(0, 0.2, (64, 64)),
                       rng.normal(0, 0.2, 64), rng.normal(0, 0.3, 64), -0.3)
    w1, b1, w2, b2, w3, b3 = pruner.unpack(flat)
    x = rng.normal(0, 1, (5, 16)).astype(np.float32)
    got = pruner.score_reference(flat, torch.as_tensor(x)).numpy()

    def r(v):   # bf16 round trip of a python float
        return float(pruner.bf16_round(torch.tensor([v], dtype=torch.float32))[0])

    for n in range(5):
        h1 = [max(0.0, sum(r(float(w1[i, k])) * r(float(x[n, k])) for k in range(16)) + float(b1[i])) for i in range(64)]
        h2 = [max(0.0, sum(r(float(w2[i, k])) * r(h1[k]) for k in range(64)) + float(b2[i])) for i in range(64)]
        s = sum(h2[i] * float(w3[i]) for i in range(64)) + b3
        assert got[n] == pytest.approx(s, rel=1e-4, abs=1e-4)


def test_shipped_weights_file(oc, pruner):
    if not os.path.exists(oc.DEFAULT_PRUNER):
        pytest.skip("no shipped pruner weights")
    flat, thr, meta = pruner.load()
    assert flat.size == pruner.NW and np.isfinite(flat).all() and np.isfinite(thr)
    assert meta["chosen"]["track_recall"] >= 0.995 and meta["chosen"]["tracks"] > 50000
    assert 0.0 < meta["margin"] < 1.0e6
    assert meta["features"] == pruner.FEATURE_NAMES and len(meta["features"]) == pruner.NF


def test_abi_argument_checks(oc):
    lib = oc.load_library()
    w = np.zeros(5313, np.float32)
    assert lib.or_session_set_pruner(None, w.ctypes.data_as(C.POINTER(C.c_float)), 5313, 0.0) == 1
    assert lib.or_pruner_score(None, None, 0, None) == 1
    assert lib.or_session_set_unit_dump(None, None, 0, None) == 1
    v = C.c_int64(0)
    assert lib.or_session_get_tuning(None, 1, C.byref(v)) == 1
