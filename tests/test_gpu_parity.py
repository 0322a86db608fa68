"""GPU parity: the gfx950 path against the CPU oracle, bit for bit.

The FP64 plain form must reproduce the reference's serial arithmetic
exactly (tolerance 0: every double compared by its bit pattern), so these
tests compare raw bits, not values within a tolerance.

  * step parity: the oracle's state after k warm-up cycles is loaded into the
    device problem; then each of the 30 doCycle steps runs through the task
    entry points on the device and through the oracle on the host, and every
    field (49 arrays) plus the reduction scalars must match after every step;
  * cycle parity: whole cycles in task mode and in fused mode against the
    oracle (rect with and without far BCs, pie with triangles, a pie origin
    past the per-thread corner sum, two-state LeBlanc, decks setting the
    parameters no reference deck sets, and the first cycles of every
    reference deck but the hex one: sedovsmall, sedovbig, leblancbig and
    nohsquare's radial inflow on a rect mesh);
  * full size: Sedov 1440^2 (2.07M zones) -- a few cycles against the oracle,
    and fused mode against task mode;
  * the C ABI called directly on caller-owned device buffers.
"""
import ctypes
import sys

import numpy as np
import pytest

import pennant_amd as pa
from oracle_lib import FIELDS, Oracle
from golden_decks import PARAMS, override, ref_deck
from pennant_amd import decks as wd

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
        gv, wv = np.asarray(got).reshape(-1), np.asarray(want).reshape(-1)
        bad = np.nonzero(bits(gv) != bits(wv))[0]
        raise AssertionError(f"{name}: {bad.size} of {gv.size} values differ; first at {bad[:4]}: "
                             f"got {gv[bad[:4]]}, want {wv[bad[:4]]}")


def load_oracle_state(pb, o, dt_current):
    for f in FIELDS:
        pb.set(f, o.get(f))
    s = o.scalars()
    pb.set_scalars([s["time"], s["dt"], s["dtrec"], s["cycle"], s["dtnew"], s["dvovmax"],
                    0, 0, dt_current, 0])


STEP_DECKS = {
    "sedov32_params": lambda: override(wd.sedov_rect(32, cstop=10 ** 6), **PARAMS),
    "sedov45": lambda: ref_deck("sedov"),
    "noh_pie30x100": lambda: ref_deck("noh"),
    "nohsmall_pie2x20": lambda: ref_deck("nohsmall"),
    "leblanc10x90": lambda: ref_deck("leblanc"),
}


@pytest.mark.parametrize("deck", list(STEP_DECKS))
def test_step_parity_every_task(deck):
    text = STEP_DECKS[deck]()
    say(f"[step parity] {deck}")
    o = Oracle(text)
    o.run(3)
    pb = pa.Problem(text)
    dt = o.begin_cycle()
    load_oracle_state(pb, o, dt)
    for f in FIELDS:
        assert_bits_equal(f"load {f}", pb.get(f), o.get(f))
    for k in range(pa.NSTEPS):
        assert o.step(k) == 0
        assert pb.step(k) == 0, f"step {k}"
        for f in FIELDS:
            assert_bits_equal(f"step {k} field {f}", pb.get(f), o.get(f))
        if k in (26, 27, 28):
            so, sp = o.scalars(), pb.scalars()
            for key in ("dtnew", "dvovmax", "dtrec"):
                assert_bits_equal(f"step {k} {key}", sp[key], so[key])
    pb.close()


CYCLE_DECKS = {
    "sedov45_farbc": (lambda: ref_deck("sedov"), 10),
    "sedov90": (lambda: wd.sedov_90(cstop=40), 40),
    "noh_pie30x100": (lambda: ref_deck("noh"), 10),
    "noh_pie64": (lambda: wd.noh_pie(64, 48, cstop=25, dtinit=1e-4), 25),
    # 96 sectors: the origin has 96 corners, more than a K2 thread takes, so
    # its block-wide sum reads corner slots past the first four positions
    "noh_pie24x96": (lambda: wd.noh_pie(24, 96, cstop=12, dtinit=1e-4), 12),
    "sedov40_params": (lambda: override(wd.sedov_rect(40, cstop=40), **PARAMS), 40),
    "noh_pie20x16_params": (lambda: override(wd.noh_pie(20, 16, cstop=25, dtinit=1e-4),
                                             **dict(PARAMS, q1=0.5, q2=1.2, qgamma=1.2)), 25),
    "leblanc10x90": (lambda: ref_deck("leblanc"), 10),
    "leblanc256x16": (lambda: wd.leblanc_rect(256, 16, cstop=30), 30),
    # the reference's other decks (hex nohpoly aside): the first 10 cycles
    "sedovsmall9x9": (lambda: ref_deck("sedovsmall"), 10),
    "sedovbig540": (lambda: override(ref_deck("sedovbig"), cstop=10), 10),
    "leblancbig160x1440": (lambda: override(ref_deck("leblancbig"), cstop=10), 10),
    # rect mesh with a radial inflow (uinitradial) and alfa 0.1; meshparams 360
    # (one value: nzy = nzx, unit lengths)
    "nohsquare360": (lambda: override(ref_deck("nohsquare"), cstop=10), 10),
}


@pytest.mark.parametrize("mode", [pa.MODE_TASKS, pa.MODE_FUSED], ids=["tasks", "fused"])
@pytest.mark.parametrize("deck", list(CYCLE_DECKS))
def test_cycle_parity(deck, mode):
    make, ncyc = CYCLE_DECKS[deck]
    text = make()
    say(f"[cycle parity] {deck} mode={mode} cycles={ncyc}")
    o = Oracle(text)
    assert o.run(1000) == ncyc
    pb = pa.Problem(text)
    assert pb.run(1000, mode) == ncyc
    for f in STATE:
        assert_bits_equal(f, pb.get(f), o.get(f))
    if mode == pa.MODE_TASKS:
        for f in FIELDS:
            assert_bits_equal(f, pb.get(f), o.get(f))
    so, sp = o.scalars(), pb.scalars()
    assert sp["cycle"] == so["cycle"]
    for key in ("time", "dt", "dtrec", "dtnew", "dvovmax"):
        assert_bits_equal(key, sp[key], so[key])
    assert sp["done"] == 1
    pb.close()


@pytest.mark.parametrize("k1,maxb,xcd", [(0, 1, 1), (0, 7, 1), (0, 8, 1), (0, 16, 0), (0, 512, 1),
                                         (1, 1, 1), (1, 5, 1), (1, 24, 1), (1, 4096, 0)])
def test_fused_launch_shapes_do_not_change_results(k1, maxb, xcd):
    """Grid-stride zone kernels with tiny and large grids, with and without
    the XCD band mapping, both K1 layouts, on a pie mesh (a triangle run and a
    quad run) and a rect mesh."""
    for text, ncyc in ((wd.noh_pie(40, 24, cstop=12, dtinit=1e-4), 12),
                       (wd.sedov_rect(37, cstop=12), 12)):
        o = Oracle(text)
        o.run(100)
        with pa.Problem(text) as pb:
            pb.set_tuning("k1_layout", k1)
            pb.set_tuning("max_blocks", maxb)
            pb.set_tuning("xcd", xcd)
            assert pb.run(100, pa.MODE_FUSED) == ncyc
            for f in STATE:
                assert_bits_equal(f"{f} k1={k1} maxb={maxb}", pb.get(f), o.get(f))
            assert_bits_equal("dtrec", pb.scalars()["dtrec"], o.scalars()["dtrec"])


@pytest.mark.parametrize("mode,pipeline", [(pa.MODE_FUSED, 0), (pa.MODE_MIXED, 0), (pa.MODE_MIXED, 1)],
                         ids=["fp64", "mixed", "mixed-pipelined"])
def test_rect_maps_do_not_change_results(mode, pipeline):
    """On a rect mesh the zone kernels compute each zone's points and corner
    slots from its index (tuning "rect_maps", on by default) instead of
    reading mapsp1 and the slot map: the state is the same bit for bit as
    with the maps read, on rect meshes with one, a few and many zones per
    row and with fixed boundaries (a pie mesh does not qualify); in FP64 it
    is the oracle's."""
    decks = [wd.sedov_rect(37, cstop=40), wd.leblanc_rect(64, 8, cstop=40), ref_deck("sedov"),
             "\n".join(ln for ln in wd.sedov_rect(1, cstop=40).splitlines() if not ln.startswith("meshparams"))
             + "\nmeshparams 1 5 1.0 1.0\n"]
    with pa.Problem(wd.noh_pie(8, 6, cstop=4)) as pie:
        assert pie.get_tuning("rect_maps") == 0
    for text in decks:
        out = []
        for rm in (1, 0):
            with pa.Problem(text) as pb:
                pb.set_tuning("rect_maps", rm)
                pb.set_tuning("pipeline", pipeline)
                assert pb.get_tuning("rect_maps") == rm
                n = pb.run(30, mode)
                out.append((n, {f: pb.get(f) for f in STATE}, pb.scalars()["dtrec"]))
        assert out[0][0] == out[1][0] > 0
        for f in STATE:
            assert_bits_equal(f"{f} rect_maps", out[0][1][f], out[1][1][f])
        if mode == pa.MODE_FUSED:
            o = Oracle(text)
            o.run(out[0][0])
            for f in STATE:
                assert_bits_equal(f"{f} vs oracle", out[0][1][f], o.get(f))


def test_sedov_xy_identical_to_oracle(tmp_path):
    text = ref_deck("sedov")
    o = Oracle(text)
    o.run(1000)
    o.write_xy(tmp_path / "ora.xy")
    with pa.Problem(text) as pb:
        pb.run(1000, pa.MODE_FUSED)
        pb.write_xy(tmp_path / "gpu.xy")
    assert (tmp_path / "gpu.xy").read_text() == (tmp_path / "ora.xy").read_text()


@pytest.mark.parametrize("deck", ["sedov", "nohsmall"])
def test_gold_export_identical_to_oracle(tmp_path, deck):
    """EnSight Gold files of the device state equal those written from the
    oracle's state byte for byte (the host writer itself is checked against
    the reference's sedov.{case,geo,zr,ze,zp} by the CPU tests)."""
    text = ref_deck(deck)
    o = Oracle(text)
    o.run(1000)
    hm = pa.HostMesh(text)
    sc = o.scalars()
    px = o.get("px")
    pa.write_gold_host(tmp_path / "ora", int(sc["cycle"]), sc["time"], o.get("zr"), o.get("ze"),
                       o.get("zp"), hm.get("znump"), px[:, 0], px[:, 1], hm.get("mapsp1"))
    with pa.Problem(text) as pb:
        pb.run(1000, pa.MODE_FUSED)
        pb.write_gold(tmp_path / "gpu")
    for ext in ("geo", "zr", "ze", "zp"):
        assert (tmp_path / f"gpu.{ext}").read_bytes() == (tmp_path / f"ora.{ext}").read_bytes(), ext


def test_run_to_tstop_in_chunks():
    # cycles queued past the stop condition do nothing (device predicate)
    text = wd.sedov_rect(40, tstop=3e-4)
    o = Oracle(text)
    n = o.run(100000)
    with pa.Problem(text) as pb:
        done = 0
        while not pb.scalars()["done"]:
            pb.launch(16, pa.MODE_FUSED)
            pb.sync()
            done += 16
        assert pb.scalars()["cycle"] == n
        assert_bits_equal("time", pb.scalars()["time"], o.scalars()["time"])
        for f in STATE:
            assert_bits_equal(f, pb.get(f), o.get(f))


def test_bad_sides_reported_like_reference():
    # a first dt far above the Courant limit inverts zones on cycle 0
    text = wd.noh_pie(20, 16, cstop=5, dtinit=0.5)
    o = Oracle(text)
    with pytest.raises(Exception, match="negative side volumes on cycle 0"):
        o.run(5)
    for mode in (pa.MODE_TASKS, pa.MODE_FUSED):
        with pa.Problem(text) as pb:
            with pytest.raises(pa.PennantError, match="negative side volumes on cycle 0"):
                pb.run(5, mode)


def test_full_size_sedov1440_against_oracle():
    text = wd.sedov_1440(cstop=3)
    say("[full size] sedov 1440^2: oracle 3 cycles")
    o = Oracle(text)
    assert o.run(3) == 3
    say("[full size] sedov 1440^2: fused 3 cycles")
    with pa.Problem(text) as pb:
        assert (pb.numz, pb.nums, pb.nump) == (2073600, 8294400, 2076481)
        assert pb.run(3, pa.MODE_FUSED) == 3
        for f in STATE:
            assert_bits_equal(f, pb.get(f), o.get(f))
        for key in ("time", "dtrec"):
            assert_bits_equal(key, pb.scalars()[key], o.scalars()[key])


def test_full_size_fused_equals_tasks_and_conserves_mass():
    text = wd.sedov_1440(cstop=12)
    say("[full size] sedov 1440^2: tasks vs fused, 12 cycles")
    with pa.Problem(text) as a, pa.Problem(text) as b:
        zm0 = a.get("zm")
        a.run(12, pa.MODE_TASKS)
        b.run(12, pa.MODE_FUSED)
        for f in STATE:
            assert_bits_equal(f, b.get(f), a.get(f))
        assert np.array_equal(b.get("zm"), zm0)
        np.testing.assert_allclose(b.get("zr") * b.get("zvol"), zm0, rtol=1e-13)


def test_noh_pie_1024_fused_equals_tasks():
    text = wd.noh_pie_1024(cstop=6)
    say("[full size] noh pie 1024^2: tasks vs fused, 6 cycles")
    with pa.Problem(text) as a, pa.Problem(text) as b:
        assert (a.numz, a.nums, a.nump) == (1048576, 4193280, 1049601)
        a.run(6, pa.MODE_TASKS)
        b.run(6, pa.MODE_FUSED)
        for f in STATE:
            assert_bits_equal(f, b.get(f), a.get(f))


def test_task_abi_on_caller_owned_buffers():
    """Calls TID_CALCCTRS / TID_CALCEDGELEN / TID_CALCDTNEW entry points on
    device buffers the caller allocated with hipMalloc, as a Legion task
    wrapper would with acc.ptr(rect)."""
    from hip_rt import DeviceArray, synchronize
    text = wd.noh_pie(12, 9, cstop=1)
    o = Oracle(text)
    L = pa.lib()
    rng = np.random.default_rng(7)
    with pa.Problem(text) as pb:
        m = pb.mesh_desc()
        px = o.get("px")
        pxx, pxy = DeviceArray(px[:, 0].copy()), DeviceArray(px[:, 1].copy())
        ex = [DeviceArray.zeros(pb.nums) for _ in range(2)]
        zx = [DeviceArray.zeros(pb.numz) for _ in range(2)]
        el = DeviceArray.zeros(pb.nums)
        synchronize()
        assert L.pennant_hip_calc_ctrs(m, pa.vec(pxx.ptr, pxy.ptr), pa.vec(ex[0].ptr, ex[1].ptr),
                                       pa.vec(zx[0].ptr, zx[1].ptr), None) == 0
        assert L.pennant_hip_calc_edge_len(m, pa.vec(pxx.ptr, pxy.ptr), el.ptr, None) == 0
        synchronize()
        assert_bits_equal("zx", np.stack([zx[0].numpy(), zx[1].numpy()], axis=1), o.get("zx"))
        assert_bits_equal("ex", np.stack([ex[0].numpy(), ex[1].numpy()], axis=1), o.get("ex"))
        p1, p2 = o.map("mapsp1"), o.map("mapsp2")
        d = px[p2] - px[p1]
        np.testing.assert_allclose(el.numpy(), np.hypot(d[:, 0], d[:, 1]), rtol=1e-15)
        # reduction entry point with a caller workspace
        a = rng.random(pb.numz) + 0.5
        u = rng.random(pb.numz)
        s = rng.random(pb.numz)
        da, du, ds = DeviceArray(a), DeviceArray(u), DeviceArray(s)
        ws = DeviceArray.zeros(L.pennant_hip_reduce_workspace_bytes() // 8)
        out = DeviceArray.zeros(1)
        assert L.pennant_hip_calc_dt_new(m, da.ptr, du.ptr, ds.ptr, 0.6, out.ptr, ws.ptr, None) == 0
        synchronize()
        want = np.min(a * 0.6 / np.maximum(u, np.maximum(s, 1e-99)))
        assert out.numpy()[0] == want


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED])
def test_autotune_restores_state_and_keeps_results(mode):
    """The autotuner times launch shapes from a snapshot and restores it: the
    problem does not advance, and the run afterwards matches an untuned run
    bit for bit (the FP64 form also matches the oracle)."""
    text = wd.noh_pie(60, 40, cstop=30, dtinit=1e-4)
    with pa.Problem(text) as ref, pa.Problem(text) as pb:
        ref.run(5, mode)
        pb.run(5, mode)
        pb.autotune(mode, cycles=2)
        say(f"[autotune] mode {mode}: k1 {pb.get_tuning('k1_layout')} max_blocks "
            f"{pb.get_tuning('max_blocks')} xcd {pb.get_tuning('xcd')}")
        assert pb.scalars()["cycle"] == 5
        assert ref.run(100, mode) == pb.run(100, mode) == 25
        for f in STATE:
            assert_bits_equal(f, pb.get(f), ref.get(f))
    if mode == pa.MODE_FUSED:
        o = Oracle(text)
        o.run(100)
        with pa.Problem(text) as pb:
            pb.autotune(mode, cycles=2)
            pb.run(100, mode)
            for f in STATE:
                assert_bits_equal(f, pb.get(f), o.get(f))


@pytest.mark.parametrize("mode", [pa.MODE_FUSED, pa.MODE_MIXED])
def test_graph_replay_matches_direct_launches(mode):
    """Runs of 16 cycles replay one captured HIP graph; the result equals
    direct launches bit for bit (and the oracle in the FP64 form), also when
    the problem stops inside a replayed run."""
    text = wd.sedov_rect(50, cstop=45)
    with pa.Problem(text) as g, pa.Problem(text) as d:
        d.set_tuning("graph", 0)
        g.launch(64, mode)
        d.launch(64, mode)
        g.sync()
        d.sync()
        assert g.scalars()["cycle"] == d.scalars()["cycle"] == 45
        for f in STATE:
            assert_bits_equal(f, g.get(f), d.get(f))
    if mode == pa.MODE_FUSED:
        o = Oracle(text)
        o.run(100)
        with pa.Problem(text) as g:
            g.launch(48, mode)
            g.sync()
            for f in STATE:
                assert_bits_equal(f, g.get(f), o.get(f))
