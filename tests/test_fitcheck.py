"""The float64 reference of the final track fit (e1039-gpu_amd/fitcheck.py):
exact lines are recovered, noisy problems agree with numpy.linalg.lstsq on
the whitened design matrix (a different method), and chi2 is the weighted
residual sum."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def fc():
    import fitcheck
    return fitcheck


def planes_like(rng, n_rows, nh=18):
    z = np.sort(rng.uniform(100.0, 1900.0, (n_rows, nh)), axis=1)
    ang = rng.choice([0.0, 0.2443, -0.2443], (n_rows, nh))
    c, s = np.cos(ang), np.sin(ang)
    r = rng.choice([0.04, 0.05, 0.1], (n_rows, nh))
    return z, c, s, 1.0 / (r * r)


def test_exact_lines(fc):
    rng = np.random.default_rng(3)
    z, c, s, w = planes_like(rng, 200)
    p = np.stack([rng.uniform(-50, 50, 200), rng.uniform(-0.15, 0.15, 200), rng.uniform(-30, 30, 200),
                  rng.uniform(-0.1, 0.1, 200)], 1)
    u = c * (p[:, :1] + p[:, 1:2] * z) + s * (p[:, 2:3] + p[:, 3:4] * z)
    mask = np.ones_like(z, bool)
    out = fc.fp64_fit(u, z, c, s, w, mask)
    assert np.allclose(out[:, :4], p, rtol=0, atol=1e-8)
    assert np.all(out[:, 4] < 1e-12)


def test_against_lstsq_with_masked_hits(fc):
    rng = np.random.default_rng(4)
    z, c, s, w = planes_like(rng, 50)
    mask = np.ones_like(z, bool)
    mask[:, 14:] = rng.uniform(size=(50, 4)) < 0.5          # 14 to 18 hits
    u = rng.normal(0, 30, (50, 1)) + rng.normal(0, 0.1, z.shape) * 5 + c * 0.05 * z
    out = fc.fp64_fit(u, z, c, s, w, mask)
    for t in range(50):
        m = mask[t]
        sw = np.sqrt(w[t, m])
        H = np.stack([c[t, m], c[t, m] * z[t, m], s[t, m], s[t, m] * z[t, m]], 1) * sw[:, None]
        p, *_ = np.linalg.lstsq(H, u[t, m] * sw, rcond=None)
        assert np.allclose(out[t, :4], p, rtol=1e-9, atol=1e-9 * (1 + np.abs(p).max()))
        chi2 = float(((H @ p - u[t, m] * sw) ** 2).sum())
        assert out[t, 4] == pytest.approx(chi2, rel=1e-9, abs=1e-9)


def test_fit_cond_bound_argument_checks(oc):
    lib = oc.load_library()
    assert lib.or_session_set_fit_cond_bound(None, 1.0) == 1
    assert oc.FIT_COND_FP32 == 1.0e4
