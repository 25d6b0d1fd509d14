"""Final straight-line fit of the output tracks (or_result.fit) on the device
against the float64 least-squares solution (fitcheck.fp64_fit), per fitted
quantity (relative, floors 1 cm for x0/y0, 1e-3 for tx/ty, 1 for chi2): the
FP32 compensated path within 1e-6 (the target is 1e-5; float32 rounding of the
result alone is 6e-8), and the FP64 path (forced with a zero condition bound)
within 1e-6 as well.  The fit output leaves the chain's results unchanged, and
the host-buffer path returns the same fits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_final_fit_against_fp64(oc, geo, session):
    import fitcheck
    ev = oc.generate(geo, 3000, seed=31, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    plain = session.reconstruct(ev)
    got = session.reconstruct(ev, fit=True)
    assert np.array_equal(plain.n_tracks, got.n_tracks)
    assert np.array_equal(plain.valid_tracks().view(np.uint8), got.valid_tracks().view(np.uint8))
    rep = fitcheck.error_report(ev, got, geo)
    assert rep["tracks"] > 1000
    assert rep["fp64_fraction"] == 0.0, rep      # well-conditioned after equilibration
    for q in ("x0", "tx", "y0", "ty", "chi2"):
        assert rep[q]["max_rel"] <= 1e-6, (q, rep[q])
    valid = np.arange(got.tracks.shape[1])[None, :] < got.n_tracks[:, None]
    f = got.fit[valid]
    assert np.array_equal(f[:, 5], got.valid_tracks()["nhits"].astype(np.float32) - 4.0)
    assert set(np.unique(f[:, 7]).tolist()) <= {0.0, 1.0}
    assert ((f[:, 7] == 1.0) | (f[:, 6] <= oc.FIT_COND_FP32)).all()
    hb = session.reconstruct_host(ev, fit=True)
    assert np.array_equal(hb.fit[valid].view(np.uint32), f.view(np.uint32))


def test_final_fit_fp64_path(oc, geo, session):
    import fitcheck
    ev = oc.generate(geo, 1000, seed=32, mult_mode=1, mult_mean=300.0, mult_cap=1000)
    fp32 = session.reconstruct(ev, fit=True)
    session.set_fit_cond_bound(0.0)
    try:
        got = session.reconstruct(ev, fit=True)
    finally:
        session.set_fit_cond_bound(oc.FIT_COND_FP32)
    valid = np.arange(got.tracks.shape[1])[None, :] < got.n_tracks[:, None]
    f = got.fit[valid]
    assert len(f) > 300 and (f[:, 7] == 1.0).all()
    assert np.array_equal(f[:, 6], fp32.fit[valid][:, 6])           # same condition estimate
    rep = fitcheck.error_report(ev, got, geo, cond_bound=0.0)
    assert rep["fp64_fraction"] == 1.0
    for q in ("x0", "tx", "y0", "ty", "chi2"):
        assert rep[q]["max_rel"] <= 1e-6, (q, rep[q])
    again = session.reconstruct(ev, fit=True)
    assert (again.fit[valid][:, 7] == 0.0).all()
