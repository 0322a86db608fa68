"""The driver loop above the cycle (Driver::run, src/Driver.cc:102-187) and the
total-energy report.

  * pennant_drive prints the reference's report line after loop cycle 0 and
    every dtreport cycles, and its end-of-run summary; the printed cycle,
    time and dt equal the oracle's at those cycles, runs stopping at cstop
    and at tstop (inside a report interval) included, in task and fused mode;
  * pennant_energy's sums equal the same sums over the oracle's state
    (numpy, ascending order) at several cycles, and the total-energy drift of
    a Sedov run is stated.
"""
import math
import re
import sys

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from oracle_lib import Oracle
from golden_decks import override, ref_deck

pytestmark = pytest.mark.gpu

WALL = re.compile(r"wall = +\d+ us")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def expected_report(text):
    """The report lines and summary Driver::run prints for this deck,
    computed from the oracle cycle by cycle (wall times left out)."""
    o = Oracle(text)
    kv = {ln.split()[0]: ln.split()[1:] for ln in text.splitlines() if ln.split()}
    cstop = int(kv["cstop"][0]) if "cstop" in kv else 999999
    tstop = float(kv["tstop"][0]) if "tstop" in kv else 1e99
    dtreport = int(kv["dtreport"][0]) if "dtreport" in kv else 10
    lines = []
    for cyc in range(cstop):
        if o.done():
            break
        assert o.run(1) == 1
        if cyc == 0 or (cyc + 1) % dtreport == 0:
            s = o.scalars()
            lines.append("End cycle %6d, time = %11.5g, dt = %11.5g, wall = X us\n"
                         % (int(s["cycle"]), s["time"], s["dt"]))
    s = o.scalars()
    summary = ("\nRun complete\n" + "cycle = %6d,        cstop = %8d\n" % (int(s["cycle"]), cstop) +
               "time = %14.6g, tstop = %8.6g\n\n" % (s["time"], tstop) +
               "************************************\n")
    return lines, summary, int(s["cycle"])


DRIVE_DECKS = {
    "sedov45_cstop10": lambda: ref_deck("sedov"),
    "noh_pie_dtreport3": lambda: override(ref_deck("noh"), dtreport=3, cstop=17),
    # stops at tstop inside a report interval
    "sedov20_tstop_dtreport7": lambda: override(wd.sedov_rect(20, tstop=2.0e-3), dtreport=7),
    "leblanc_dtreport1": lambda: override(ref_deck("leblanc"), dtreport=1),
}


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_TASKS], ids=["fused", "tasks"])
@pytest.mark.parametrize("deck", list(DRIVE_DECKS))
def test_drive_report_lines_match_oracle(deck, mode):
    text = DRIVE_DECKS[deck]()
    lines, summary, ncyc = expected_report(text)
    say(f"[drive] {deck} mode {mode}: {ncyc} cycles, {len(lines)} report lines")
    with pa.Problem(text) as pb:
        n, out, us = pb.drive(mode)
    assert n == ncyc
    got = WALL.sub("wall = X us", out)
    want = "".join(lines) + summary
    assert got.startswith(want), f"\n--- got\n{got}\n--- want\n{want}"
    tail = got[len(want):]
    m = re.fullmatch(r"hydro cycle run time= +([0-9.e+-]+) us\n\*{36}\n", tail)
    assert m, tail
    assert float(m.group(1)) == pytest.approx(us, rel=1e-6) and us > 0
    # the per-report wall times are whole microseconds
    walls = [int(w) for w in re.findall(r"wall = +(\d+) us", out)]
    assert len(walls) == len(lines) and all(w >= 0 for w in walls)


def test_drive_mixed_and_learned_complete():
    from pennant_amd.closure import Mlp
    from pennant_amd.train_closure import DEFAULT_OUT
    text = override(wd.sedov_rect(48, tstop=5e-3), dtreport=25)
    m = Mlp.load(DEFAULT_OUT)
    for mode in (pa.MODE_MIXED, pa.MODE_LEARNED):
        with pa.Problem(text) as pb, pa.Problem(text) as ref:
            if mode == pa.MODE_LEARNED:
                pb.set_closure(m)
                ref.set_closure(m)
            n, out, _ = pb.drive(mode)
            n2 = ref.run(10 ** 6, mode)
            assert n == n2 and out.count("End cycle") == 1 + n // 25
            for f in ("zr", "ze", "px"):
                assert np.array_equal(pb.get(f).view(np.uint64), ref.get(f).view(np.uint64))
    m.close()


def oracle_energy(o, hm_maps):
    """PENNANT's energy sums over the oracle's state: internal = sum zetot;
    kinetic = sum over corners of zm * cvol / zvol * |u|^2 / 2 with
    cvol = zarea(x) * x_p * (smf_s + smf_s3) / 2, zarea of the current
    positions, accumulated in ascending zone and side order."""
    zsbeg, znump, p1, p2, s3 = hm_maps
    zetot, zm, zvol, smf = (o.get(f) for f in ("zetot", "zm", "zvol", "smf"))
    px, pu = o.get("px"), o.get("pu")
    ei = 0.0
    for v in zetot:
        ei += v
    ek = 0.0
    for z in range(zm.size):
        sb, n = int(zsbeg[z]), int(znump[z])
        cx = cy = 0.0
        for k in range(n):
            cx += px[p1[sb + k], 0] / n
            cy += px[p1[sb + k], 1] / n
        za = 0.0
        for k in range(n):
            a, b = p1[sb + k], p2[sb + k]
            za += 0.5 * ((px[b, 0] - px[a, 0]) * (cy - px[a, 1]) - (px[b, 1] - px[a, 1]) * (cx - px[a, 0]))
        for k in range(n):
            s = sb + k
            p = p1[s]
            cvol = za * px[p, 0] * 0.5 * (smf[s] + smf[s3[s]])
            ek += zm[z] * cvol / zvol[z] * 0.5 * (pu[p, 0] * pu[p, 0] + pu[p, 1] * pu[p, 1])
    return ei, ek


@pytest.mark.parametrize("deck", ["sedov", "noh"])
def test_energy_sums_equal_oracle_state_sums(deck):
    text = override(ref_deck(deck), cstop=30)
    hm = pa.HostMesh(text)
    maps = tuple(hm.get(n) for n in ("zsbeg", "znump", "mapsp1", "mapsp2", "mapss3"))
    o = Oracle(text)
    with pa.Problem(text) as pb:
        for c in (0, 1, 10, 30):
            while o.scalars()["cycle"] < c:
                o.run(1)
                pb.run(1, pa.MODE_FUSED)
            ei, ek = pb.energy()
            wi, wk = oracle_energy(o, maps)
            say(f"  {deck} cycle {c}: internal {ei!r} kinetic {ek!r}")
            assert ei == pytest.approx(wi, rel=1e-12, abs=0)
            assert ek == pytest.approx(wk, rel=1e-12, abs=1e-300)


def test_sedov_energy_drift_is_reported():
    """Sedov 45^2 to t = 1 in the FP64 form: E(0) is the deposited energy
    per radian, and the drift E(t)/E(0) - 1 is stated (and bounded, so that
    a change in it is seen)."""
    text = override(ref_deck("sedov"), cstop=10 ** 6)
    with pa.Problem(text) as pb:
        e0 = sum(pb.energy())
        while not pb.scalars()["done"]:
            pb.run(4096, pa.MODE_FUSED)
        e1 = sum(pb.energy())
    deposited = 5027.7 * (0.025 ** 3 / 2.0)         # einitsub x RZ volume per radian
    say(f"  E(0) {e0!r} (deposited {deposited!r}), E(t=1) {e1!r}, drift {e1 / e0 - 1:+.4e}")
    assert e0 == pytest.approx(deposited, rel=1e-12)
    assert 0.0 < e1 / e0 - 1 < 0.2
    assert math.isfinite(e1)
