"""EnSight Gold export (src/ExportGold.cc:31-307) against the reference's own
output set test/sedov/sedov.{case,geo,zr,ze,zp} (45x45 Sedov after 10
cycles, written by the reference with basename ./test/sedov/sedov).

The writer runs on host arrays here: the oracle's state after 10 cycles and
the host mesh.  Comparison rules for values follow test_oracle_golden.py:
the golden's far field (zone rows >= 15, point rows >= 18) carries spurious
bands and is excluded from the value comparison; every header line, every
connectivity line and every z coordinate must match exactly.
"""
import os

import numpy as np
import pytest

import pennant_amd as pa
from oracle_lib import Oracle
from golden_decks import HERE, ref_deck

GOLD = HERE / "golden" / "sedov"
NZX = 45


def _write_oracle_gold(tmp_path, text, ncyc, base):
    o = Oracle(text)
    assert o.run(1000) == ncyc
    hm = pa.HostMesh(text)
    sc = o.scalars()
    px = o.get("px")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        os.makedirs(os.path.dirname(base), exist_ok=True)
        pa.write_gold_host(base, int(sc["cycle"]), sc["time"], o.get("zr"), o.get("ze"),
                           o.get("zp"), hm.get("znump"), px[:, 0], px[:, 1], hm.get("mapsp1"))
    finally:
        os.chdir(cwd)
    return o, hm


@pytest.fixture(scope="module")
def sedov_gold(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("gold")
    _write_oracle_gold(tmp, ref_deck("sedov"), 10, "./test/sedov/sedov")
    return tmp / "test" / "sedov"


def test_case_file_identical(sedov_gold):
    assert (sedov_gold / "sedov.case").read_bytes() == (GOLD / "sedov.case").read_bytes()


def test_geo_file_matches(sedov_gold):
    got = (sedov_gold / "sedov.geo").read_text().split("\n")
    want = (GOLD / "sedov.geo").read_text().split("\n")
    assert len(got) == len(want)
    npts = int(want[8])
    assert npts == (NZX + 1) ** 2
    assert got[:9] == want[:9]                       # cycle, t, headers, point count
    rows = np.arange(npts) // (NZX + 1)
    for c in range(2):                               # x, then y coordinates
        off = 9 + c * npts
        for p in np.nonzero(rows < 18)[0]:
            assert got[off + p] == want[off + p], (c, p, got[off + p], want[off + p])
    zoff = 9 + 2 * npts
    assert got[zoff:zoff + npts] == want[zoff:zoff + npts]   # z = 0
    # element block: quad4, count, zone ids, corner points
    assert got[zoff + npts:] == want[zoff + npts:]
    assert want[zoff + npts] == "quad4"


@pytest.mark.parametrize("var", ["zr", "ze", "zp"])
def test_var_files_match(sedov_gold, var):
    got = (sedov_gold / f"sedov.{var}").read_text().split("\n")
    want = (GOLD / f"sedov.{var}").read_text().split("\n")
    assert len(got) == len(want)
    assert got[:4] == want[:4] == [var, "part", "         1", "quad4"]
    nz = NZX * NZX
    g, w = got[4:4 + nz], want[4:4 + nz]
    rows = np.arange(nz) // NZX
    matched = 0
    for z in np.nonzero(rows < 15)[0]:
        if abs(float(w[z])) >= 1e-30:
            assert g[z] == w[z], (z, g[z], w[z])
            matched += 1
        else:
            assert abs(float(g[z]) - float(w[z])) <= 1e-30
    assert matched >= (600 if var == "zr" else 30)


def test_pie_mesh_triangles_then_quads(tmp_path):
    """Pie mesh (no reference output): triangles first, then quads, each with
    ascending 1-based zone ids; the connectivity read back equals mapsp1+1."""
    text = ref_deck("nohsmall")
    o, hm = _write_oracle_gold(tmp_path, text, Oracle(text).run(1000), "out/noh")
    lines = (tmp_path / "out" / "noh.geo").read_text().split("\n")
    znump, mapsp1 = hm.get("znump"), hm.get("mapsp1")
    npts = int(lines[8])
    i = 9 + 3 * npts
    zsbeg = np.concatenate([[0], np.cumsum(znump)])
    seen = []
    for kind, nv in (("tria3", 3), ("quad4", 4)):
        zl = np.nonzero(znump == nv)[0]
        if zl.size == 0:
            continue
        assert lines[i] == kind and int(lines[i + 1]) == zl.size
        ids = [int(x) for x in lines[i + 2:i + 2 + zl.size]]
        assert ids == list(zl + 1)
        conn = lines[i + 2 + zl.size:i + 2 + 2 * zl.size]
        for z, row in zip(zl, conn):
            assert [int(x) for x in row.split()] == list(mapsp1[zsbeg[z]:zsbeg[z + 1]] + 1)
            assert row == "".join("%10d" % (p + 1) for p in mapsp1[zsbeg[z]:zsbeg[z + 1]])
        seen += list(zl)
        i += 2 + 2 * zl.size
    assert sorted(seen) == list(range(znump.size))
    assert lines[i] == ""
    zr = (tmp_path / "out" / "noh.zr").read_text().split("\n")
    assert zr[3] == "tria3"
    assert float(zr[4]) == pytest.approx(o.get("zr")[0], rel=1e-5)


def test_nsided_block(tmp_path):
    """A zone with more than 4 sides goes to the nsided block: ids, then the
    side counts, then the corner points (src/ExportGold.cc:185-207)."""
    znump = np.array([5, 3, 4, 6], dtype=np.int32)
    mapsp1 = np.arange(18, dtype=np.int32) % 9
    px = np.linspace(0.0, 1.0, 9)
    py = -px
    zr = np.array([1.0, 2.0, 3.0, 4.0])
    pa.write_gold_host(tmp_path / "m", 7, 0.5, zr, zr * 10, zr * 100, znump, px, py, mapsp1)
    geo = (tmp_path / "m.geo").read_text().split("\n")
    assert geo[0] == "cycle =        7" and geo[1] == "t =  5.00000000e-01"
    i = 9 + 27
    # zone 1 (the triangle) owns sides 5-7, zone 2 (the quad) sides 8-11
    assert geo[i:i + 4] == ["tria3", "         1", "         2", "         6         7         8"]
    assert geo[i + 4:i + 8] == ["quad4", "         1", "         3",
                                "         9         1         2         3"]
    assert geo[i + 8:i + 10] == ["nsided", "         2"]
    assert geo[i + 10:i + 12] == ["         1", "         4"]          # zone ids
    assert geo[i + 12:i + 14] == ["         5", "         6"]          # side counts
    assert geo[i + 14] == "".join("%10d" % (p + 1) for p in mapsp1[0:5])
    assert geo[i + 15] == "".join("%10d" % (p + 1) for p in mapsp1[12:18])
    zrf = (tmp_path / "m.zr").read_text().split("\n")
    assert zrf == ["zr", "part", "         1", "tria3", " 2.00000e+00", "quad4", " 3.00000e+00",
                   "nsided", " 1.00000e+00", " 4.00000e+00", ""]
