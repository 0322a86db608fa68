"""Batched problems on the GPU (pennant_batch_*).

  * every problem of a batch in the plain FP64 form is bit-identical to its
    own single-problem run (and so to the CPU oracle), with problems of
    different mesh types, sizes, parameters and stop conditions;
  * the mixed form of a batch matches each problem's own mixed run;
  * graph replay, grid shapes, computed rect maps and the autotuner do not
    change results, and the autotuner does not advance the problems;
  * a batch records training tuples and runs the learned closure.
"""
import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from pennant_amd.accuracy import rel_l1
from pennant_amd.device import DeviceBuffer
from golden_decks import PARAMS, override, ref_deck

pytestmark = pytest.mark.gpu

STATE = ("px", "pu", "zr", "ze", "zp", "zetot", "zwrate", "zvol")


def say(msg):
    print(msg, flush=True)


def mixed_decks():
    return [wd.sedov_rect(37, cstop=40), wd.noh_pie(40, 24, cstop=35, dtinit=1e-4),
            ref_deck("sedov"), ref_deck("noh"), wd.leblanc_rect(64, 8, cstop=50),
            wd.sedov_rect(300, energy=0.7, cstop=25), wd.noh_pie(300, 257, speed=1.7, cstop=30, dtinit=1e-5),
            ref_deck("nohsmall"),
            # per-problem parameters away from the other problems'
            override(wd.sedov_rect(30, cstop=35), **PARAMS),
            override(wd.noh_pie(20, 16, cstop=25, dtinit=1e-4), **dict(PARAMS, q1=0.5, q2=1.2, qgamma=1.2))]


def single(text, mode, ncyc=10 ** 6):
    with pa.Problem(text) as pb:
        n = pb.run(ncyc, mode)
        return n, pb.scalars(), {f: pb.get(f) for f in STATE}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_fp64_bit_identical_to_single_runs(graph):
    decks = mixed_decks()
    say(f"[batch fp64] {len(decks)} problems, graph={graph}")
    with pa.Batch(decks) as b:
        b.set_tuning("graph", graph)
        b.run(100, pa.MODE_FUSED)          # every problem stops before 100 cycles
        assert b.active() == 0
        for i, text in enumerate(decks):
            n, sc, want = single(text, pa.MODE_FUSED)
            got = b.scalars(i)
            assert got["cycle"] == sc["cycle"] and got["done"] == 1, i
            for k in ("time", "dt", "dtrec", "dtnew", "dvovmax"):
                assert bits(got[k]) == bits(sc[k]), (i, k)
            for f in STATE:
                assert np.array_equal(bits(b.get(i, f)), bits(want[f])), (i, f)


def test_batch_mixed_bit_identical_to_single_mixed_runs():
    """The mixed form is compiled without contraction, so a batched problem
    gives the same bits as its own mixed run."""
    decks = mixed_decks()
    with pa.Batch(decks) as b:
        b.run(100, pa.MODE_MIXED)
        for i, text in enumerate(decks):
            _, sc, want = single(text, pa.MODE_MIXED)
            assert b.scalars(i)["cycle"] == sc["cycle"]
            assert bits(b.scalars(i)["time"]) == bits(sc["time"]), i
            for f in STATE:
                assert np.array_equal(bits(b.get(i, f)), bits(want[f])), (i, f)
            assert rel_l1(b.get(i, "zr"), want["zr"]) == 0.0


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED])
def test_batch_rect_maps_do_not_change_results(mode):
    """Rect problems placed after pie problems in the arena (non-zero first
    side and point) give the same bits with their points and corner slots
    computed from the zone index as with the maps read."""
    decks = mixed_decks()
    out = []
    for rect in (1, 0):
        with pa.Batch(decks) as b:
            b.set_tuning("rect_maps", rect)
            b.run(100, mode)
            out.append([(b.scalars(i)["cycle"], [bits(b.get(i, f)) for f in STATE])
                        for i in range(len(decks))])
    for i, (on, off) in enumerate(zip(*out)):
        assert on[0] == off[0], i
        for f, x, y in zip(STATE, on[1], off[1]):
            assert np.array_equal(x, y), (i, f)


@pytest.mark.parametrize("maxb,xcd", [(1, 0), (7, 1), (64, 1), (65535, 0)])
def test_batch_grid_shapes(maxb, xcd):
    decks = mixed_decks()[:5]
    with pa.Batch(decks) as ref, pa.Batch(decks) as b:
        b.set_tuning("max_blocks", maxb)
        b.set_tuning("xcd", xcd)
        ref.run(20, pa.MODE_MIXED)
        b.run(20, pa.MODE_MIXED)
        for i in range(len(decks)):
            for f in STATE:
                assert np.array_equal(bits(b.get(i, f)), bits(ref.get(i, f))), (i, f)


def test_autotune_does_not_advance_or_change_results():
    decks = mixed_decks()[:5]
    with pa.Batch(decks) as ref, pa.Batch(decks) as b:
        ref.run(7, pa.MODE_MIXED)
        b.run(7, pa.MODE_MIXED)
        before = [b.scalars(i)["cycle"] for i in range(len(decks))]
        ms = b.autotune(pa.MODE_MIXED, cycles=3)
        say(f"[autotune] best {ms:.3f} ms / 3 cycles: max_blocks {b.get_tuning('max_blocks')} "
            f"xcd {b.get_tuning('xcd')}")
        assert [b.scalars(i)["cycle"] for i in range(len(decks))] == before
        ref.run(20, pa.MODE_MIXED)
        b.run(20, pa.MODE_MIXED)
        for i in range(len(decks)):
            for f in STATE:
                assert np.array_equal(bits(b.get(i, f)), bits(ref.get(i, f))), (i, f)


def test_ensemble_runs_and_captures_tuples():
    decks, n, par = wd.ensemble(24, cstop=10 ** 6)
    say(f"[ensemble] 24 problems, {int((n.astype(np.int64) ** 2).sum())} zones")
    with pa.Batch(decks) as b:
        b.autotune(pa.MODE_MIXED, cycles=2)
        b.run(32, pa.MODE_MIXED)
        buf = DeviceBuffer((1 << 20, 12), np.float32)
        b.set_capture(buf, 1 << 20)
        b.run(1, pa.MODE_MIXED)
        cnt = b.capture_count()
        b.set_capture(None, 0)
        say(f"  captured {cnt} tuples")
        assert cnt > 1000
        tup = buf.numpy(rows=min(cnt, 1 << 20))
        assert np.all(np.isfinite(tup))
        assert np.all(tup[:, 10] > 0) and np.all(tup[:, 10] <= 1.0 + 1e-6)
        assert np.all(tup[:, 11] >= 1.0 - 1e-6) and np.all(tup[:, 11] < 2.0 + 1e-6)
        # each problem's state stays finite and its cycle count advanced
        for i in range(24):
            sc = b.scalars(i)
            assert sc["cycle"] == 33
            assert np.all(np.isfinite(b.get(i, "zr")))


def test_batch_learned_bit_identical_to_single_learned_runs():
    """The batch evaluates the closure in one pass, a single problem in two
    (flag, then recompute the flagged zones); both apply the network to the
    same gated corners and give the same bits."""
    from pennant_amd.closure import Mlp
    from pennant_amd.train_closure import DEFAULT_OUT
    m = Mlp.load(DEFAULT_OUT)
    decks = mixed_decks()[:6]
    with pa.Batch(decks) as b:
        b.set_closure(m)
        b.run(60, pa.MODE_LEARNED)
        nb = b.closure_stats()
        tot = [0, 0]
        for i, text in enumerate(decks):
            with pa.Problem(text) as pb:
                pb.set_closure(m)
                pb.run(60, pa.MODE_LEARNED)
                st = pb.closure_stats()
                tot[0] += st[0]
                tot[1] += st[1]
                for f in STATE:
                    assert np.array_equal(bits(b.get(i, f)), bits(pb.get(f))), (i, f)
        say(f"[batch learned] zone evaluations batch {nb}, single runs {tuple(tot)}")
        assert nb == tuple(tot) and nb[0] > 0
    m.close()
