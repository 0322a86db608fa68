"""A mesh run as several pieces on one device, against the one-piece run.

The reference splits a run into numpcs pieces (-n <numpcs>, src/main.cc:61-80)
with private, master and ghost points (src/Mesh.cc:432-617) and reduces
shared points' corner mass and force through reduction instances
(src/Hydro.cc:571-574,766-769).  Here each piece has its own arrays; the
owner of a master point sums its corners from every piece in ascending
one-piece side order, so the run in pieces must give the one-piece run's
bits (the one-piece FP64 run is itself bit-identical to the CPU oracle,
tests/test_gpu_parity.py), in the FP64 and in the mixed-precision form.

Cases: Sedov rect with far BCs, Noh pie (triangles at the origin, which all
pieces of the first row reach), a 96-sector pie whose origin (96 corners, 72
of them received) takes the owner's block-wide sum, an elongated LeBlanc
mesh, piece counts whose block rounding leaves pieces empty, and Sedov
1440 x 1440 in 8 pieces.
"""
import sys

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

STATE = ("px", "pu", "zr", "ze", "zp", "zetot", "zwrate", "zvol", "zm")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bits_equal(name, got, want):
    g, w = bits(got), bits(want)
    if g.shape != w.shape or not np.array_equal(g, w):
        bad = np.nonzero(g.reshape(-1) != w.reshape(-1))[0]
        raise AssertionError(f"{name}: {bad.size} of {g.size} values differ; first at {bad[:4]}")


def check_layout(pp, one):
    lay = [pp.layout(i) for i in range(pp.count)]
    assert sum(d["zones"] for d in lay) == one.numz
    assert sum(d["sides"] for d in lay) == one.nums
    assert sum(d["private"] + d["master"] for d in lay) == one.nump
    if pp.count > 1:
        assert sum(d["ghost"] for d in lay) > 0
        assert sum(d["master"] for d in lay) > 0
        assert sum(d["received"] for d in lay) > 0
    return lay


def run_both(text, numpcs, mode, cycles):
    with pa.Problem(text) as one, pa.Pieces(text, numpcs) as pp:
        lay = check_layout(pp, one)
        say(f"  {numpcs} pieces -> {pp.count} ({lay[0]['numpcx']} x {lay[0]['numpcy']}), "
            f"ghosts {[d['ghost'] for d in lay]}, received corners {[d['received'] for d in lay]}")
        n1 = one.run(cycles, mode)
        n2 = pp.run(cycles, mode)
        assert n1 == n2 == cycles, (n1, n2)
        s1, s2 = one.scalars(), pp.scalars()
        for k in ("time", "dt", "dtrec", "cycle", "dtnext", "dvovmax", "dtnew"):
            assert bits(np.array([s2[k]]))[0] == bits(np.array([s1[k]]))[0], (k, s2[k], s1[k])
        for f in STATE:
            assert_bits_equal(f"{numpcs} pieces {f}", pp.get(f), one.get(f))


CASES = {
    "sedov30_far": (wd.sedov_rect(30, cstop=400, bc_far=True), (2, 3, 4, 6, 9)),
    "noh_pie_24x20": (wd.noh_pie(20, 24, cstop=400), (2, 4, 5, 8)),
    "pie_96_sectors": (wd.noh_pie(8, 96, cstop=400), (4, 6)),
    "leblanc_64x8": (wd.leblanc_rect(64, 8, cstop=400), (4, 8)),
    "sedov10_empty_pieces": (wd.sedov_rect(10, cstop=400), (7,)),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED], ids=["fp64", "mixed"])
def test_pieces_bit_identical_to_one_piece(case, mode):
    text, counts = CASES[case]
    for k in counts:
        say(f"[pieces] {case} mode {mode} numpcs {k}")
        run_both(text, k, mode, 60)


def test_pieces_against_oracle():
    """Pieces against the CPU oracle directly: Noh pie in 4 pieces, 25 cycles."""
    text = wd.noh_pie(16, 12, cstop=25)
    o = Oracle(text)
    assert o.run(100) == 25
    with pa.Pieces(text, 4) as pp:
        assert pp.run(100, pa.MODE_FUSED) == 25
        for f in ("px", "pu", "zr", "ze", "zp", "zetot"):
            assert_bits_equal(f, pp.get(f), o.get(f))


def test_pieces_mixed_then_fp64_cycles():
    """Mixed cycles leave zr and ze derived; FP64 cycles after them store and
    read them in every piece as the one-piece run does."""
    text = wd.sedov_rect(24, cstop=400)
    with pa.Problem(text) as one, pa.Pieces(text, 4) as pp:
        for mode in (pa.MODE_MIXED, pa.MODE_FUSED, pa.MODE_MIXED):
            assert one.run(15, mode) == pp.run(15, mode) == 15
        for f in STATE:
            assert_bits_equal(f, pp.get(f), one.get(f))


def test_pieces_stop_and_bad_input():
    text = wd.sedov_rect(12, cstop=7)
    with pa.Pieces(text, 4) as pp:
        assert pp.run(100, pa.MODE_FUSED) == 7
        assert pp.run(10, pa.MODE_FUSED) == 0
        assert pp.scalars()["done"] == 1
    with pytest.raises(pa.PennantError):
        pa.Pieces(text, 0)
    with pa.Pieces(text, 2) as pp:
        with pytest.raises(pa.PennantError):
            pp.run(1, pa.MODE_LEARNED)


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED], ids=["fp64", "mixed"])
def test_pieces_graph_replays_between_eager_cycles(mode):
    """Runs of 16 cycles replay a captured graph (each phase's piece kernels
    on parallel branches); eager cycles in between: the state equals the
    one-piece run's after every run."""
    text = wd.noh_pie(20, 24, cstop=400)
    with pa.Problem(text) as one, pa.Pieces(text, 6) as pp:
        for n in (16, 5, 32, 3, 16):
            assert one.run(n, mode) == pp.run(n, mode) == n
            for f in ("px", "pu", "zr", "ze", "zetot"):
                assert_bits_equal(f"{f} after {n}", pp.get(f), one.get(f))


def test_pieces_time_cycles_advances_like_run():
    """pennant_pieces_time_cycles queues cycles between events (a graph for
    the first 16): the state it leaves equals pennant_pieces_run's."""
    text = wd.noh_pie(16, 12, cstop=400)
    with pa.Pieces(text, 3) as a, pa.Pieces(text, 3) as b:
        ms = a.time_cycles(21, pa.MODE_MIXED)
        assert ms > 0
        assert b.run(21, pa.MODE_MIXED) == 21
        assert a.scalars()["cycle"] == b.scalars()["cycle"] == 21
        for f in STATE:
            assert_bits_equal(f, a.get(f), b.get(f))


def test_pieces_sedov_1440_eight_pieces():
    """Full size: Sedov 1440^2 (2.07M zones) as 8 pieces, 5 cycles, bit-identical."""
    text = wd.sedov_1440(cstop=10 ** 6)
    say("[pieces] sedov 1440^2, 8 pieces")
    run_both(text, 8, pa.MODE_FUSED, 5)
