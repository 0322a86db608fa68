"""The pipelined mixed cycle (tuning key "pipeline"; off by default, tried by
the autotuner).

One zone pass holds a cycle's corrector and the next cycle's predictor; the
predictor assumes the next dt computed without the volume-change limit
(src/Hydro.cc:1647-1667, src/Driver.cc:213-245) and is redone when the
finalize finds that the limit set it.  Checked here against the unpipelined
mixed cycle (separate predictor and corrector passes), bit for bit:

  * meshes with triangles, quads and pie origins, both mixed forms, with and
    without graph replay, stopping at cstop and tstop;
  * decks whose small cflv makes the volume-change limit bind (redone
    look-ahead predictors);
  * interleaved cycles of other modes, state writes and odd launch counts;
  * the learned form's closure counts.
"""
import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from golden_decks import ref_deck

pytestmark = pytest.mark.gpu

STATE = ("px", "pu", "zr", "ze", "zp", "zetot", "zwrate", "zvol")
SCAL = ("time", "dt", "dtrec", "cycle", "dtnew", "dvovmax", "dtnext", "done")


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def tight_cflv(text, cflv):
    """The deck with its volume-change limit (cflv, 0.1 when absent) set."""
    if "cflv " in text:
        assert "cflv 0.1\n" in text
        return text.replace("cflv 0.1\n", f"cflv {cflv}\n")
    return text + f"cflv {cflv}\n"


def decks():
    return [wd.sedov_rect(37, cstop=40), wd.noh_pie(40, 24, cstop=35, dtinit=1e-4),
            ref_deck("sedov"), ref_deck("noh"), wd.leblanc_rect(64, 8, cstop=50),
            wd.sedov_rect(120, energy=0.7, cstop=45), ref_deck("nohsmall")]


def closure():
    from pennant_amd.closure import Mlp
    from pennant_amd.train_closure import DEFAULT_OUT
    return Mlp.load(DEFAULT_OUT)


def run(text, mode, pipeline, steps, mlp=None, graph=1):
    """Runs the launch sequence `steps` (cycle counts) and returns the state,
    scalars, closure counts and look-ahead redos."""
    with pa.Problem(text) as pb:
        pb.set_tuning("pipeline", pipeline)
        pb.set_tuning("graph", graph)
        if mlp is not None:
            pb.set_closure(mlp)
        for n in steps:
            pb.run(n, mode)
        st = {f: pb.get(f) for f in STATE}
        sc = pb.scalars()
        cs = pb.closure_stats() if mlp is not None else None
        return st, sc, cs, pb.lookahead_redos()


def same(a, b, tag):
    sa, ca, csa, _ = a
    sb, cb, csb, _ = b
    for k in SCAL:
        assert bits(ca[k]) == bits(cb[k]), (tag, k, ca[k], cb[k])
    for f in STATE:
        assert np.array_equal(bits(sa[f]), bits(sb[f])), (tag, f)
    assert csa == csb, (tag, csa, csb)


@pytest.mark.parametrize("graph", [1, 0])
def test_pipelined_mixed_bit_identical(graph):
    for i, text in enumerate(decks()):
        a = run(text, pa.MODE_MIXED, 1, [100], graph=graph)
        b = run(text, pa.MODE_MIXED, 0, [100], graph=graph)
        say(f"[pipeline mixed] deck {i}: cycle {a[1]['cycle']:.0f}, redos {a[3]}")
        same(a, b, i)


def test_pipelined_learned_bit_identical():
    m = closure()
    for i, text in enumerate(decks()[:5]):
        a = run(text, pa.MODE_LEARNED, 1, [60], mlp=m)
        b = run(text, pa.MODE_LEARNED, 0, [60], mlp=m)
        say(f"[pipeline learned] deck {i}: closure zones {a[2]}, redos {a[3]}")
        same(a, b, i)
    m.close()


@pytest.mark.parametrize("cflv", [0.02, 0.002])
def test_redone_lookahead_bit_identical(cflv):
    """A small cflv makes the volume-change limit set dt on some cycles (on
    nearly all at 0.002): the look-ahead predictor is redone there."""
    m = closure()
    for text in (tight_cflv(wd.sedov_rect(48, cstop=80), cflv),
                 tight_cflv(wd.noh_pie(40, 24, cstop=60, dtinit=1e-4), cflv)):
        for mode, mlp in ((pa.MODE_MIXED, None), (pa.MODE_LEARNED, m)):
            a = run(text, mode, 1, [100], mlp=mlp)
            b = run(text, mode, 0, [100], mlp=mlp)
            say(f"[redo cflv {cflv}] mode {mode}: cycle {a[1]['cycle']:.0f}, redos {a[3]}")
            assert a[3] > 0
            assert b[3] == 0
            same(a, b, (cflv, mode))
    m.close()


def test_interleaved_modes_and_state_writes():
    text = wd.sedov_rect(64, cstop=500)

    def seq(pipeline):
        with pa.Problem(text) as pb:
            pb.set_tuning("pipeline", pipeline)
            pb.run(7, pa.MODE_MIXED)
            pb.run(5, pa.MODE_FUSED)
            pb.run(33, pa.MODE_MIXED)
            ze = pb.get("ze")
            ze[::7] *= 1.001
            pb.set("ze", ze)
            pb.run(16, pa.MODE_MIXED)
            pb.device_ptr("zr")          # a caller may write through it
            pb.run(2, pa.MODE_MIXED)
            sc = pb.scalars()
            vals = [sc["time"], sc["dt"], sc["dtrec"], sc["cycle"], sc["dtnew"], sc["dvovmax"],
                    sc["numsbad_pred"], sc["numsbad_corr"], sc["dtnext"], sc["done"]]
            pb.set_scalars(vals)
            pb.run(3, pa.MODE_MIXED)
            pb.run(1, pa.MODE_TASKS)
            pb.run(19, pa.MODE_MIXED)
            return {f: pb.get(f) for f in STATE}, pb.scalars(), None, pb.lookahead_redos()

    a, b = seq(1), seq(0)
    say(f"[interleaved] cycle {a[1]['cycle']:.0f}")
    assert a[1]["cycle"] == 7 + 5 + 33 + 16 + 2 + 3 + 1 + 19
    same(a, b, "interleaved")


def test_stops_at_tstop():
    """The run's last cycle (here ended by tstop) runs no look-ahead; zp is
    that cycle's predictor pressure."""
    text = wd.sedov_rect(32, cstop=10 ** 6, tstop=2e-3)
    a = run(text, pa.MODE_MIXED, 1, [10 ** 5])
    b = run(text, pa.MODE_MIXED, 0, [10 ** 5])
    say(f"[tstop] cycle {a[1]['cycle']:.0f} time {a[1]['time']:.6e}")
    assert a[1]["done"] == 1
    same(a, b, "tstop")


def test_autotune_learned_keeps_state_and_counts():
    """The autotuner (which also tries the pipelined cycle for the mixed
    forms) restores the state and the closure counters: a tuned problem goes
    on exactly like an untuned one."""
    m = closure()
    text = wd.sedov_rect(64, cstop=500)
    with pa.Problem(text) as a, pa.Problem(text) as b:
        for pb in (a, b):
            pb.set_closure(m)
            pb.run(20, pa.MODE_LEARNED)
        before = b.closure_stats()
        b.autotune(pa.MODE_LEARNED, cycles=3)
        say(f"[autotune learned] max_blocks {b.get_tuning('max_blocks')} xcd {b.get_tuning('xcd')} "
            f"pipeline {b.get_tuning('pipeline')}")
        assert b.closure_stats() == before
        assert b.scalars()["cycle"] == 20
        a.run(37, pa.MODE_LEARNED)
        b.run(37, pa.MODE_LEARNED)
        for f in STATE:
            assert np.array_equal(bits(a.get(f)), bits(b.get(f))), f
        assert a.closure_stats() == b.closure_stats()
    m.close()


@pytest.mark.parametrize("pipeline", [0, 1])
def test_learned_list_slices_keep_results(pipeline):
    """The learned pass's zone list holds one slice per XCD band of the unit
    mapping, and a single slice when the mapping is off (xcd 0, or a grid
    that is not a multiple of 8).  Every launch shape gives the same state
    and closure counts, bit for bit, on a mesh with a triangle run and a quad
    run of 50 units."""
    m = closure()
    text = wd.noh_pie(128, 100, cstop=500, dtinit=1e-4)
    ref = None
    for cap, xcd in ((2048, 1), (48, 1), (16, 1), (13, 1), (48, 0)):
        with pa.Problem(text) as pb:
            pb.set_closure(m)
            pb.set_tuning("max_blocks", cap)
            pb.set_tuning("xcd", xcd)
            pb.set_tuning("pipeline", pipeline)
            pb.run(40, pa.MODE_LEARNED)
            out = ([bits(pb.get(f)) for f in STATE], pb.closure_stats())
        say(f"[list slices] pipeline {pipeline} max_blocks {cap} xcd {xcd}: closure {out[1]}")
        assert out[1][0] > 0
        if ref is None:
            ref = out
            continue
        for f, x, y in zip(STATE, ref[0], out[0]):
            assert np.array_equal(x, y), f
        assert out[1] == ref[1]
    m.close()
