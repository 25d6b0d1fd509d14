"""Geometry reading and wire table (online_reconstruction.cu:115-218).

The product reader and the oracle reader are written independently and must
agree bit for bit; both are pinned against values read straight from the
file text and against the wire end-point model of the same file.
"""
import os

import numpy as np
import pytest

FIELDS = ["z", "nelem", "cellwidth", "spacing", "xoffset", "scalex", "x0", "x1", "x2", "costheta", "scaley",
          "y0", "y1", "y2", "sintheta", "resolution", "p1x", "p1y", "p1z", "deltapx", "deltapy", "deltapz",
          "dp1x", "dp1y", "dp1z", "slope_max", "inter_max"]


def test_product_and_oracle_readers_agree(geo, ogeo):
    pl, wire = geo.export()
    opl, owire, _ = ogeo.export()
    assert np.array_equal(pl.view(np.uint32), opl.view(np.uint32))
    assert np.array_equal(wire.view(np.uint32), owire.view(np.uint32))


def test_columns_match_file_text(oc, geo):
    pl, _ = geo.export()
    merged = {45, 47, 48}  # lines whose resolution and p1x tokens run together
    with open(oc.DEFAULT_GEOMETRY) as f:
        for line in f:
            if line.startswith("#"):
                continue
            tok = line.split()
            d = int(tok[0])
            if d in merged:
                continue
            vals = [float(t) for t in tok[1:26]]
            for k in range(25):
                assert pl[k, d] == np.float32(vals[k]), (d, FIELDS[k])


def test_merged_tokens_follow_stream_extraction(geo):
    # "2.749630657183.4666022" (basic_geometry_e906.txt:48): resolution 2.749630657183, p1x .4666022
    pl, _ = geo.export()
    for d, p1x in ((47, 0.4666022), (48, 0.468201)):
        assert pl[15, d] == np.float32(2.749630657183)
        assert pl[16, d] == np.float32(p1x)
    # "11.34348941-145.1692364": the minus sign ends the first number
    assert pl[15, 45] == np.float32(11.34348941)
    assert pl[16, 45] == np.float32(-145.1692364)


def test_slope_and_intercept_limits_pairing(geo):
    # even planes above 1 add (res_i + res_{i-1}) / (z_i - z_{i-1}) to themselves and to plane i-1
    pl, _ = geo.export()
    ct, st, z, res, smax, imax = pl[9], pl[14], pl[0], pl[15], pl[25], pl[26]
    for d in range(2, 55, 2):
        base_i = np.float32(ct[d] * np.float32(0.15) + st[d] * np.float32(0.1))
        base_j = np.float32(ct[d - 1] * np.float32(0.15) + st[d - 1] * np.float32(0.1))
        with np.errstate(divide="ignore"):  # hodoscope pairs share a z: the limit becomes inf
            ds = float(np.float32(res[d] + res[d - 1]) / np.float32(z[d] - z[d - 1]))
        assert smax[d] == np.float32(float(base_i) + ds)
        assert smax[d - 1] == np.float32(float(base_j) + ds)
        di = ds * float(z[d])
        bi = np.float32(ct[d] * np.float32(150) + st[d] * np.float32(50))
        assert imax[d] == np.float32(float(bi) + di)
    assert smax[1] != 0


def test_wire_table_spacing_and_centre(geo):
    pl, wire = geo.export()
    for d in list(range(1, 31)) + list(range(31, 47)):
        n = int(pl[1, d])
        if n == 0:
            continue
        steps = np.diff(wire[d, 1:n + 1].astype(np.float64))
        assert np.allclose(steps, pl[3, d], atol=2e-5 * max(1.0, abs(wire[d, 1:n + 1]).max()))
    # prop tubes: the alignment offset changes every 8 tubes
    for d in range(47, 55):
        n = int(pl[1, d])
        steps = np.diff(wire[d, 1:n + 1].astype(np.float64))
        inner = [k for k in range(n - 1) if (k + 1) % 8 != 0]
        assert np.allclose(steps[inner], pl[3, d], atol=1e-4)


def test_wire_table_matches_end_point_model(geo):
    # the wire coordinate x*cos + y*sin along the end-point wire line agrees with the table
    pl, wire = geo.export()
    for d in list(range(1, 7)) + list(range(13, 47)):
        n = int(pl[1, d])
        e = np.arange(1, n + 1)
        p1x = pl[16, d] + pl[22, d] * (e - 1)
        p1y = pl[17, d] + pl[23, d] * (e - 1)
        u = p1x * pl[9, d] + p1y * pl[14, d]
        assert np.abs(u - wire[d, 1:n + 1]).max() < 0.02, d


def test_geometry_errors(oc, tmp_path):
    with pytest.raises(oc.OrError, match="cannot open"):
        oc.Geometry.load(str(tmp_path / "missing.txt"))
    bad = tmp_path / "bad.txt"
    bad.write_text("99 1.0 2 0.5 0.5 0 1 0 0 0 1 1 0 0 0 0 0.1 0 0 0 0 0 0 0 0 0 0\n")
    with pytest.raises(oc.OrError, match="detector id"):
        oc.Geometry.load(str(bad))
