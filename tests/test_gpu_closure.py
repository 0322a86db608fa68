"""The learned QCS closure on the GPU, against the CPU oracle
(oracle/mlp_oracle.py, pinned by tests/test_mlp_oracle.py).

  * inference: the training GEMM path and the wave-level MFMA routine of the
    fused predictor against the oracle's forward pass (BF16 rounding points
    emulated);
  * gradients: loss and every parameter gradient of the GPU backward pass
    against the oracle's;
  * training: deterministic, and it lowers the loss on exact QCS targets;
  * capture: the tuples a mixed cycle records against the oracle's features
    and targets of the same state;
  * in the loop: the learned form against the exact QCS in the same mixed
    cycle and against FP64 (error and time per cycle), the envelope
    fallback, and, as a sanity check of the network, the same closure with
    constant coefficients or an untrained network.

Tolerances: BF16 keeps 8 significant bits, so single activations can round
differently when the FP32 accumulation order differs; outputs are compared
to 1e-2 absolute (typical differences are 1e-3).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from pennant_amd.accuracy import rel_l1
from pennant_amd.closure import Mlp, collect_tuples
from pennant_amd.device import DeviceBuffer
from oracle_lib import FIELDS, Oracle
from golden_decks import ref_deck

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def random_params(seed, scale=0.25):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, mo.NPARAM).astype(np.float32)
    W1 = p[mo.OFF_W1:mo.OFF_W2].reshape(mo.NH, mo.K1)
    W1[:, mo.NF + 1:] = 0.0
    return p


def random_features(seed, n):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1.5, (n, mo.NF)).astype(np.float32)


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 70001])
def test_inference_paths_against_oracle(n):
    say(f"[closure inference] n={n}")
    m = Mlp(0)
    p = random_params(11)
    m.set_params(p)
    x = random_features(5, n)
    want, _ = mo.forward(p, x, rounding=True)
    for path in (0, 1):
        got = m.forward(x, path)
        err = np.max(np.abs(got - want))
        say(f"  path {path}: max abs err {err:.2e}")
        assert err <= 1e-2, (path, err)
    # the two GPU paths round at the same points
    assert np.max(np.abs(m.forward(x, 0) - m.forward(x, 1))) <= 1e-2
    m.close()


def test_gradients_against_oracle():
    say("[closure gradients]")
    m = Mlp(0)
    p = random_params(21)
    m.set_params(p)
    rng = np.random.default_rng(4)
    t = np.concatenate([random_features(6, 3000), rng.uniform(0, 2, (3000, 2)).astype(np.float32)], axis=1)
    loss, g = m.loss_grad(t)
    lref, gref = mo.loss_grad(p, t, rounding=True)
    say(f"  loss {loss:.6e} vs {lref:.6e}; grad rel err "
        f"{np.linalg.norm(g - gref) / np.linalg.norm(gref):.2e}")
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-2 * np.linalg.norm(gref)
    for name, a, b in (("W1", mo.OFF_W1, mo.OFF_W2), ("W2", mo.OFF_W2, mo.OFF_B2), ("b2", mo.OFF_B2, mo.OFF_W3),
                       ("W3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.NPARAM)):
        e = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-30)
        assert e <= 2e-2, (name, e)
    W1g = g[mo.OFF_W1:mo.OFF_W2].reshape(mo.NH, mo.K1)
    assert np.all(W1g[:, mo.NF + 1:] == 0.0)
    m.close()


def synthetic_tuples(n, seed):
    """Features with f9 spread over [0, 1] and the exact targets of f9."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, mo.NF)).astype(np.float32)
    x[:, 9] = rng.uniform(0, 1, n)
    r = x[:, 9].astype(np.float64)
    nrm = np.sqrt((1 - r) ** 2 + r ** 2)
    t = np.stack([(1 - r) / nrm, 1 + r / nrm], axis=1)
    return np.concatenate([x, t.astype(np.float32)], axis=1)


def test_training_is_deterministic_and_learns():
    say("[closure training]")
    tup = synthetic_tuples(20000, 1)
    a, b = Mlp(3), Mlp(3)
    la = a.train(tup, steps=600, batch=1024, lr=2e-3, seed=9)
    lb = b.train(tup, steps=600, batch=1024, lr=2e-3, seed=9)
    assert np.array_equal(la, lb)
    assert np.array_equal(a.get_params(), b.get_params())
    say(f"  loss {la[0]:.3e} -> {la[-1]:.3e}")
    assert la[-20:].mean() <= 0.05 * la[:20].mean()
    # one Adam step from the oracle's gradient reproduces the GPU's update
    c = Mlp(3)
    p0 = c.get_params()
    small = tup[:256]
    loss, g = c.loss_grad(small)
    c.train(small, steps=1, batch=256, lr=1e-3, seed=0)
    p1 = c.get_params()
    moved = np.abs(p1 - p0)
    assert moved.max() > 0 and moved.max() <= 1.01e-3 * 1.0001 + 1e-9   # Adam's first step is lr*sign
    for m in (a, b, c):
        m.close()


def _capture_vs_oracle(text, ncyc, q1, q2, qgamma):
    o = Oracle(text)
    o.run(ncyc)
    with pa.Problem(text) as pb:
        for f in FIELDS:
            pb.set(f, o.get(f))
        s = o.scalars()
        dt = o.begin_cycle()
        pb.set_scalars([s["time"], s["dt"], s["dtrec"], s["cycle"], s["dtnew"], s["dvovmax"], 0, 0, dt, 0])
        for k in range(14):
            assert o.step(k) == 0
        fields = {f: o.get(f) for f in ("pu0", "elen", "cdiv", "cevol", "cdu", "zss", "zrp", "zr", "zuc")}
        maps = {m: o.map(m) for m in ("mapsp1", "mapsp2", "mapss3", "mapsz")}
        _, want = mo.qcs_tuples(fields, maps, q1, q2, qgamma)
        cap = 4 * fields["cdiv"].size + 16
        buf = DeviceBuffer((cap, 12), np.float32)
        pb.set_capture(buf, cap)
        pb.run(1, pa.MODE_MIXED)
        n = pb.capture_count()
        got = buf.numpy(rows=min(n, cap))
        pb.set_capture(None, 0)
    return got, want


def test_capture_matches_oracle_features_uniform_compression():
    """Noh's first cycle compresses every corner: the mixed cycle records
    exactly the oracle's corners, with the same features and targets (FP32
    vs FP64), compared as multisets (each column sorted)."""
    got, want = _capture_vs_oracle(ref_deck("noh"), 0, 0.1, 1.0, 5.0 / 3.0)
    say(f"[capture noh] {got.shape[0]} tuples (oracle {want.shape[0]})")
    assert got.shape[0] == want.shape[0] > 1000
    for j in range(12):
        a, b = np.sort(got[:, j].astype(np.float64)), np.sort(want[:, j])
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.max(np.abs(a - b)) <= 1e-4 * scale, (j, np.max(np.abs(a - b)), scale)


def test_capture_matches_oracle_features_sedov():
    """At a Sedov front a few corners have a divergence so close to zero that
    FP32 and FP64 disagree on its sign: the counts agree to 2% and the
    feature and target distributions (quantiles) agree closely."""
    got, want = _capture_vs_oracle(ref_deck("sedov"), 6, 0.1, 1.0, 5.0 / 3.0)
    say(f"[capture sedov] {got.shape[0]} tuples (oracle {want.shape[0]})")
    assert abs(got.shape[0] - want.shape[0]) <= 0.02 * want.shape[0]
    q = np.linspace(0.05, 0.95, 19)
    for j in range(12):
        a, b = np.quantile(got[:, j], q), np.quantile(want[:, j], q)
        scale = max(np.max(np.abs(want[:, j])), 1e-30)
        assert np.max(np.abs(a - b)) <= 2e-2 * scale, (j, a, b)


def train_closure(seed=5):
    """Trains the closure on tuples of a few short mixed runs (Sedov and Noh
    at two sizes and energies)."""
    decks = [wd.sedov_rect(48, energy=0.5, cstop=400), wd.sedov_rect(64, cstop=300),
             wd.noh_pie(40, 30, cstop=300, dtinit=1e-4), wd.noh_pie(64, 48, speed=1.5, cstop=200, dtinit=1e-4)]
    tup = collect_tuples(decks, cycles=400, every=10, cap_per_run=1 << 18, max_total=400000)
    say(f"  {tup.shape[0]} training tuples")
    m = Mlp(seed)
    losses = m.train(tup, steps=1500, batch=4096, lr=2e-3, seed=seed)
    m.fit_envelope(tup)
    say(f"  training loss {losses[0]:.3e} -> {losses[-20:].mean():.3e}")
    return m, tup, losses


@pytest.fixture(scope="module")
def trained():
    say("[closure] training")
    m, tup, losses = train_closure()
    yield m, tup, losses
    m.close()


def run_form(text, mode, mlp=None, ncyc=10 ** 7):
    with pa.Problem(text) as pb:
        if mlp is not None:
            pb.set_closure(mlp)
        n = pb.run(ncyc, mode)
        stats = pb.closure_stats() if mlp is not None else None
        return n, pb.scalars(), {f: pb.get(f) for f in ("zr", "ze", "zp", "px")}, stats


def test_trained_network_fits_exact_targets(trained):
    m, tup, losses = trained
    assert losses[-20:].mean() <= 1e-3
    y = m.forward(tup[:20000, :10], 1)
    err = np.abs(y - tup[:20000, 10:])
    say(f"  held-in max err {err.max():.3e} mean {err.mean():.3e}")
    assert err.mean() <= 2e-2


def test_learned_cycle_in_loop(trained):
    m, _, _ = trained
    text = wd.sedov_rect(96, cstop=500)
    n64, s64, ref, _ = run_form(text, pa.MODE_FUSED)
    nmx, smx, mix, _ = run_form(text, pa.MODE_MIXED)
    nle, sle, lea, stats = run_form(text, pa.MODE_LEARNED, m)
    assert n64 == nmx == nle == 500
    e_mix = rel_l1(mix["zr"], ref["zr"])
    e_lea = rel_l1(lea["zr"], ref["zr"])
    say(f"  L1 zr: mixed {e_mix:.2e} learned {e_lea:.2e}; zones learned/fallback {stats}")
    assert stats[0] > 0
    assert e_lea <= 1e-2


def test_envelope_fallback_is_exact_qcs(trained):
    """With an empty envelope every gated zone falls back to the exact
    coefficient: the learned form then gives the mixed form's bits."""
    m, _, _ = trained
    lo, hi = m.get_envelope()
    m.set_envelope(np.full(10, 1e30, np.float32), np.full(10, -1e30, np.float32))
    try:
        text = wd.sedov_rect(40, cstop=100)
        _, _, mix, _ = run_form(text, pa.MODE_MIXED)
        _, _, lea, stats = run_form(text, pa.MODE_LEARNED, m)
    finally:
        m.set_envelope(lo, hi)
    say(f"  fallback zones {stats}")
    assert stats[0] == 0 and stats[1] > 0
    for f in ("zr", "ze", "zp", "px"):
        assert np.array_equal(lea[f].view(np.uint64), mix[f].view(np.uint64)), f


def test_learned_against_exact_qcs(trained):
    """What the trained network gives, measured against the method it stands
    in for -- the exact QCS coefficient (src/QCS.cc:285-296) in the same
    mixed-precision cycle -- and against plain FP64, on Sedov 96^2 to t = 1.

    The learned coefficient is an approximation of the exact one, so it
    cannot beat it on accuracy; what it costs is reported here: its density
    and energy errors against FP64 beside the exact-QCS mixed form's (the
    density error must stay within 1%, the headline bound), and its time per
    cycle beside the mixed form's (the network pass adds work; the learned
    form is the slower of the two on every measured workload).

    Sanity check of the network itself: the same closure with the network
    replaced by constant coefficients y0 = y1 = 1 (linear plus quadratic
    viscosity without the Kurapatenko blend) is at least 5x further from
    FP64, and an untrained network is worse still or inverts a zone."""
    m, _, _ = trained
    text = wd.sedov_rect(96)
    _, s64, ref, _ = run_form(text, pa.MODE_FUSED)
    _, smx, mix, _ = run_form(text, pa.MODE_MIXED)
    _, sle, lea, stats = run_form(text, pa.MODE_LEARNED, m)
    assert s64["done"] == smx["done"] == sle["done"] == 1
    e_mix, e_lea = rel_l1(mix["zr"], ref["zr"]), rel_l1(lea["zr"], ref["zr"])
    e_mix_e, e_lea_e = rel_l1(mix["ze"], ref["ze"]), rel_l1(lea["ze"], ref["ze"])
    # time per cycle of the two mixed forms on the same state (Sedov 360^2, mid-run)
    big = wd.sedov_rect(360)
    us = {}
    for name, mode in (("mixed", pa.MODE_MIXED), ("learned", pa.MODE_LEARNED)):
        with pa.Problem(big) as pb:
            if mode == pa.MODE_LEARNED:
                pb.set_closure(m)
            pb.run(2000, mode)
            ms, _ = pb.time_cycles(200, mode, phases=False)
            us[name] = ms * 1e3 / 200
    say(f"  Sedov 96^2 to t=1, density L1 vs FP64: exact QCS (mixed) {e_mix:.2e}, learned {e_lea:.2e}; "
        f"energy L1 {e_mix_e:.2e} / {e_lea_e:.2e}; zones learned/fallback {stats}")
    say(f"  Sedov 360^2 cycle: exact QCS (mixed) {us['mixed']:.1f} us, learned {us['learned']:.1f} us")
    assert stats[0] > 0
    assert e_lea <= 1e-2
    # sanity check: the network against no network and an untrained one
    lo, hi = m.get_envelope()
    const = Mlp(0)
    p = np.zeros(mo.NPARAM, np.float32)
    p[mo.OFF_B3:mo.NPARAM] = 1.0
    const.set_params(p)
    const.set_envelope(lo, hi)
    short = wd.sedov_rect(96, cstop=500)
    _, _, ref5, _ = run_form(short, pa.MODE_FUSED)
    _, _, lea5, _ = run_form(short, pa.MODE_LEARNED, m)
    _, _, cst5, _ = run_form(short, pa.MODE_LEARNED, const)
    untrained = Mlp(5)
    untrained.set_envelope(lo, hi)
    try:
        _, _, unt5, _ = run_form(short, pa.MODE_LEARNED, untrained)
        e_u = rel_l1(unt5["zr"], ref5["zr"])
    except pa.PennantError as ex:           # random coefficients can invert zones
        say(f"  untrained network: run stopped ({ex})")
        e_u = float("inf")
    e_l, e_c = (rel_l1(x["zr"], ref5["zr"]) for x in (lea5, cst5))
    say(f"  500 cycles, L1 zr vs FP64: trained {e_l:.2e}, constant coefficients {e_c:.2e}, untrained {e_u:.2e}")
    assert e_l * 5 <= e_c
    assert e_l * 5 <= e_u
    const.close()
    untrained.close()
