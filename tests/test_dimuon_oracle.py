"""The oracle's dimuon building and display histograms (oracle/or_oracle_dimuon.cpp,
restating reconstruction_kernels.cuh:2820-3195 and reconstruction_helper.cuh:
1211-1476), pinned without retyping it:
  * kinematics: each record's mass, pT, x1, x2, xF, costheta and phi against
    float64 closed forms of the 4-vectors it stores (Minkowski products, the
    boost to the centre-of-mass frame written as gamma (pz - beta E));
  * the reference's cuts hold for every record, 4-vector energies are
    E = sqrt(m_mu^2 + p^2), pair counts never exceed the ordered opposite-charge
    pairs the reference enumerates, the closest-approach vertex lies on the
    swim's z grid;
  * histograms against a vectorised numpy fill of the same records and tracks
    (every bin's predicate evaluated in float32);
  * the not-a-bit-test trigger mapping and the "all" row.
The closest-approach swim on its own is pinned against a continuous-model
integration in tests/test_dimuon_swim.py."""
import numpy as np
import pytest

MMU, MP, EBEAM, PBEAM = 0.105658, 0.938272, 120.0, 119.996331801


@pytest.fixture(scope="module")
def case(oc, orc, geo, ogeo):
    p = oc.default_params(max_hits_per_event=2048)
    p.full_chain = 1
    ev = oc.generate(geo, 3000, seed=77, mult_mode=1, mult_mean=150.0, mult_cap=600, n_tracks_min=2, n_tracks_max=3)
    chain = orc.reconstruct(ogeo, p, ev, vertex=True, threads=8)
    trig = np.random.default_rng(3).integers(0, 1024, ev.n_events).astype(np.int32)
    z = geo.export()[0][0]
    d = orc.dimuons(chain, ev, z, p.max_tracks_out, trig)
    return p, ev, chain, trig, d


def records(d):
    rows = np.arange(d.dimuons.shape[1])[None, :] < d.n_dimuons[:, None]
    return d.dimuons[rows].astype(np.float64), np.nonzero(rows)[0]


def test_counts(case):
    p, ev, chain, trig, d = case
    assert d.n_dimuons.sum() > 200 and d.lost == 0
    for e in np.nonzero(d.n_dimuons)[0]:
        t = chain.tracks[e, : chain.n_tracks[e]]
        t = t[t["station_id"] >= 7]
        pairs = 0
        for a in range(len(t)):
            for b in range(len(t)):
                if t["charge"][a] * t["charge"][b] > 0:
                    continue
                if t["slot"][a] < t["slot"][b] or (t["slot"][a] == t["slot"][b]):
                    pairs += 1          # same slot: both orders
        assert d.n_dimuons[e] <= pairs


def test_kinematics_closed_forms(case):
    _, _, _, _, d = case
    R, _ = records(d)
    P, N, Ps, Ns = R[:, 18:22], R[:, 22:26], R[:, 26:30], R[:, 30:34]
    for v in (P, N, Ps, Ns):
        assert np.allclose(v[:, 3], np.sqrt(MMU ** 2 + (v[:, :3] ** 2).sum(1)), rtol=1e-6)
    v1 = 0.5 * (R[:, 12:15] + R[:, 15:18])
    single = np.all(np.abs(R[:, 9:12] - v1) <= 1e-6 * (1 + np.abs(v1)), axis=1)
    q = np.where(single[:, None], Ps + Ns, P + N)
    m = np.sqrt(q[:, 3] ** 2 - (q[:, :3] ** 2).sum(1))
    pT = np.hypot(q[:, 0], q[:, 1])
    s = (EBEAM + MP) ** 2 - PBEAM ** 2
    x1 = q[:, 3] / (EBEAM + MP)
    x2 = (EBEAM * q[:, 3] - PBEAM * q[:, 2]) / (EBEAM * (EBEAM + MP) - PBEAM * PBEAM)
    beta = PBEAM / (EBEAM + MP)
    gamma = 1 / np.sqrt(1 - beta * beta)
    xF = 2 * gamma * (q[:, 2] - beta * q[:, 3]) / (np.sqrt(s) * (1 - m * m / s))
    cth = 2 * (N[:, 3] * P[:, 2] - N[:, 2] * P[:, 3]) / (m * np.sqrt(m * m + pT * pT))
    phi = np.arctan2(2 * np.sqrt(m * m + pT * pT) * (N[:, 0] * P[:, 1] - N[:, 1] * P[:, 0]),
                     m * (P[:, 0] ** 2 - N[:, 0] ** 2 + P[:, 1] ** 2 - N[:, 1] ** 2))
    assert np.allclose(R[:, 0], m, rtol=5e-4)          # float32 E^2 - p^2 cancels ~60 GeV against ~4
    assert np.allclose(R[:, 1], pT, rtol=1e-4, atol=1e-5)
    assert np.allclose(R[:, 3], x1, rtol=1e-5)
    assert np.allclose(R[:, 4], x2, rtol=5e-4, atol=1e-6)   # float32 Minkowski products cancel
    assert np.allclose(R[:, 2], xF, rtol=1e-3, atol=1e-4)
    assert np.allclose(R[:, 5], cth, rtol=2e-3, atol=1e-3)
    assert np.allclose(R[:, 6], phi, rtol=2e-3, atol=1e-3)
    assert (~single).sum() > 0.5 * len(R)


def test_cuts_and_vertex_grid(case):
    _, _, _, _, d = case
    R, _ = records(d)
    P, N = R[:, 18:22], R[:, 22:26]
    assert np.all(-1.054 * (P[:, 0] - N[:, 0]) >= 0)
    assert np.all((R[:, 0] >= 0) & (R[:, 0] <= 10) & (np.abs(R[:, 2]) <= 1))
    assert np.all((R[:, 3] >= 0) & (R[:, 3] <= 1) & (R[:, 4] >= 0) & (R[:, 4] <= 1) & (np.abs(R[:, 5]) <= 1))
    pz = P[:, 2] + N[:, 2]
    assert np.all((pz >= 30) & (pz <= 120) & (np.abs(P[:, 0] + N[:, 0]) <= 3))
    assert np.all(R[:, 9] ** 2 + R[:, 10] ** 2 <= 225)
    assert np.all(R[:, [7, 8, 34, 35]] == 0)
    # closest-approach vertices sit on the swim's z grid: FMAG steps of 2 x 2.5146 cm from
    # z = 502.92, then target steps of 5 cm
    v1 = 0.5 * (R[:, 12:15] + R[:, 15:18])
    dca = ~np.all(np.abs(R[:, 9:12] - v1) <= 1e-6 * (1 + np.abs(v1)), axis=1)
    grid = np.concatenate([502.92 - 2 * 2.5146 * np.arange(101), 502.92 - 502.92 - 5.0 * np.arange(1, 101)])
    dist = np.abs(R[dca, 11][:, None] - grid[None, :]).min(1)
    assert dist.max() < 1e-2


def test_histograms_against_numpy_fill(case):
    p, ev, chain, trig, d = case
    nb = np.array([100, 100, 100, 100, 100, 200, 50, 50, 50, 50, 50, 50, 100, 100, 100, 100, 100, 200])
    lo = np.array([-20, -20, 0, -20, -20, -500, 0, 0, 0, 0, 0, -3.141592654, -20, -20, 0, -20, -20, -500], np.float32)
    hi = np.array([20, 20, 120, 20, 20, 500, 10, 1, 1, 0.5, 5, 3.141592654, 20, 20, 120, 20, 20, 500], np.float32)
    want = np.zeros((11, 18, 256), np.float64)
    trg_on = np.array([[(t & n) != 0 or n == 10 for n in range(11)] for t in trig])      # (events, 11)

    def fill(k, vals, ev_of):
        hw = np.float32((hi[k] - lo[k]) * np.float32(0.5) / np.float32(nb[k]))
        j = np.arange(nb[k], dtype=np.float32)
        x = (lo[k] + hw * (np.float32(2) * j + np.float32(1))).astype(np.float32)
        v = vals.astype(np.float32)[:, None]
        inbin = ((x - hw).astype(np.float32) <= v) & (v < (x + hw).astype(np.float32))   # (n, nbins)
        want[:, k, : nb[k]] += trg_on[ev_of].T.astype(np.float64) @ inbin.astype(np.float64)

    valid = np.arange(chain.tracks.shape[1])[None, :] < chain.n_tracks[:, None]
    ev_of = np.broadcast_to(np.arange(ev.n_events)[:, None], valid.shape)[valid]
    tr = chain.tracks[valid]
    vx = chain.vertex[valid]
    sel = (tr["station_id"] >= 7) & ~(tr["chi2"] / (tr["nhits"].astype(np.float32) - np.float32(5)) > 250)
    for k, col in enumerate([3, 4, 5, 0, 1, 2]):
        fill(k, vx[sel, col], ev_of[sel])
    R, dev = records(d)
    for k, col in enumerate([0, 2, 3, 4, 1, 6]):
        fill(6 + k, R[:, col], dev)
    for k, (a, b) in enumerate([(18, 22), (19, 23), (20, 24)]):
        fill(12 + k, (R[:, a].astype(np.float32) + R[:, b].astype(np.float32)), dev)
    for k, col in enumerate([9, 10, 11]):
        fill(15 + k, R[:, col], dev)
    got = d.histograms.reshape(11, 18, 256)
    assert np.array_equal(got, want.astype(np.float32))
    assert got[0].sum() == 0            # trigger 0: bits & 0 is never set
    assert got[10].sum() > 0


def test_no_trigger_bits_fills_only_all(oc, orc, case, geo):
    p, ev, chain, trig, d = case
    d0 = orc.dimuons(chain, ev, geo.export()[0][0], p.max_tracks_out, None)
    h = d0.histograms.reshape(11, 18, 256)
    assert h[:10].sum() == 0
    assert np.array_equal(h[10], d.histograms.reshape(11, 18, 256)[10])
    assert np.array_equal(d0.n_dimuons, d.n_dimuons)


def test_abi_argument_checks(oc):
    import ctypes as C
    lib = oc.load_library()
    hits = oc.OrHits(0, 0, None, None, None, None, None, None, 1)
    res = oc.OrResult(None, None, None, None, None, None, None, None, 1)
    out = oc.OrDimuonResult(None, None, None, None)
    assert lib.or_dimuons(None, C.byref(hits), C.byref(res), None, C.byref(out)) == 1
