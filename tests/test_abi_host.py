"""CPU checks of the product library: it loads, exports every entry point the
header declares, and its host logic (deck parsing, error behaviour, one-piece
mesh numbering, point-corner CSR) agrees with the oracle.  No device needed.
"""
import ctypes
import re

import numpy as np
import pytest

import pennant_amd as pa
from oracle_lib import Oracle, ROOT
from golden_decks import ref_deck
from pennant_amd import decks as wd

HEADER = ROOT / "include" / "pennant_hip.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(pennant_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = pa.lib()
    names = header_functions()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding, and back
    assert set(names) == set(pa.exported_symbols())


def test_task_entry_points_cover_reference_task_ids():
    names = set(header_functions())
    tasks = ["calc_ctrs", "calc_vols", "calc_side_fracs", "calc_surf_vecs", "calc_edge_len",
             "calc_char_len", "adv_pos_half", "calc_rho", "calc_crnr_mass", "sum_crnr_force",
             "calc_accel", "adv_pos_full", "calc_work", "calc_work_rate", "calc_energy",
             "calc_dt_new", "calc_dvol", "calc_dt", "calc_state_half", "calc_force_pgas",
             "calc_force_tts", "set_corner_div", "set_qcn_force", "set_force_qcs", "set_vel_diff",
             "apply_fixed_bc", "calc_global_dt"]
    assert {f"pennant_hip_{t}" for t in tasks} <= names


@pytest.mark.parametrize("deck", ["sedov", "sedovsmall", "noh", "nohsmall", "leblanc"])
def test_host_mesh_matches_oracle(deck):
    text = ref_deck(deck)
    o = Oracle(text)
    h = pa.HostMesh(text)
    for name in ("mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4", "znump"):
        np.testing.assert_array_equal(h.get(name), o.map(name).astype(np.int32), err_msg=name)
    px = o.get("px")
    # coordinates bit for bit (pie points use the host libm's cos/sin, as the
    # reference does)
    assert np.array_equal(h.get("px"), px[:, 0]) and np.array_equal(h.get("py"), px[:, 1])


@pytest.mark.parametrize("text", [wd.noh_pie(7, 5, cstop=1), wd.sedov_rect(6, cstop=1)])
def test_point_corner_csr(text):
    h = pa.HostMesh(text)
    p1 = h.get("mapsp1")
    pcbeg, pcside = h.get("pcbeg"), h.get("pcside")
    nump = pcbeg.size - 1
    assert pcbeg[0] == 0 and pcbeg[-1] == p1.size
    assert np.array_equal(np.sort(pcside), np.arange(p1.size))
    for p in range(nump):
        s = pcside[pcbeg[p]:pcbeg[p + 1]]
        assert np.all(p1[s] == p) and np.all(np.diff(s) > 0)
    zsbeg, znump = h.get("zsbeg"), h.get("znump")
    assert np.array_equal(np.diff(zsbeg), znump)


def test_pie_origin_valence():
    h = pa.HostMesh(wd.noh_pie(9, 12, cstop=1))
    pcbeg = h.get("pcbeg")
    assert pcbeg[1] - pcbeg[0] == 12          # one corner per centre triangle


@pytest.mark.parametrize("text,msg", [
    ("cstop 5\nmeshparams 4 4\n", "must specify meshtype"),
    ("cstop 5\nmeshtype tri\nmeshparams 4 4\n", "invalid meshtype tri"),
    ("cstop 5\nmeshtype rect\n", "must specify meshparams"),
    ("cstop 5\nmeshtype rect\nmeshparams 4 4 1 1 1\n", "meshparams must have <= 4 values"),
    ("cstop 5\nmeshtype rect\nmeshparams 0 4\n", "meshparams values must be positive"),
    ("cstop 5\nmeshtype pie\nmeshparams 4 4 360\n", "meshparams theta must be < 360"),
    ("meshtype rect\nmeshparams 4 4\n", "Must specify either cstop or tstop"),
    ("cstop 5\ncstop 6\nmeshtype rect\nmeshparams 4 4\n", "Duplicate key cstop in input file"),
    ("cstop 5\nmeshtype rect\nmeshparams 4 4\nsubregion 0 1 0\n", "subregion must have 4 entries"),
    ("cstop 5\nmeshtype hex\nmeshparams 4\n", "hex is not supported"),
    # like the reference's getline/eof loop, a last line without a newline is not read
    ("meshtype rect\nmeshparams 4 4\ncstop 5", "Must specify either cstop or tstop"),
])
def test_deck_errors_match_reference_messages(text, msg):
    with pytest.raises(pa.PennantError, match=re.escape(msg)):
        pa.HostMesh(text)
    if "hex" not in msg:
        from oracle_lib import OracleError
        with pytest.raises(OracleError, match=re.escape(msg)):
            Oracle(text)


def test_deck_comments_and_defaults():
    h = pa.HostMesh("# a comment\ncstop 5\n\nmeshtype rect\nmeshparams 3\n")
    assert h.get("znump").size == 9          # nzy defaults to nzx
    assert np.max(h.get("px")) == 1.0        # lenx defaults to 1.0
