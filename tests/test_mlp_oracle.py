"""Pins the learned-closure oracle (oracle/mlp_oracle.py) on the CPU.

  * BF16 rounding: round-to-nearest-even cases fixed by the format;
  * backward pass: the float64 analytic gradient against central finite
    differences of the float64 loss (a dropped term, a wrong mask or a
    transposed product fails it);
  * targets: y0, y1 reproduce the reference QCS coefficient, checked against
    the hydro oracle's own crmu (src/QCS.cc:285-296) on Sedov and Noh
    states: crmu = (q1 y0 zss + G y1 du) zrp cevol for every compressing
    corner;
  * features: f8 = log(1 + du/zss) and f9 are consistent with the targets
    (y0, y1 are functions of f9 alone), and compressing corners are exactly
    those with cdiv < 0.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402

from oracle_lib import Oracle  # noqa: E402
from golden_decks import ref_deck  # noqa: E402
from pennant_amd import decks as wd  # noqa: E402


def deck_value(text, key, default):
    for line in text.splitlines():
        parts = line.split()
        if parts and parts[0] == key:
            return float(parts[1])
    return default


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)               # BF16 spacing at 1
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0],
                 dtype=np.float32)
    got = mo.bf16(x)
    assert got[0] == one
    assert got[1] == one                        # tie -> even (1.0)
    assert got[2] == one + 2 * ulp              # tie -> even (1 + 2^-6)
    assert got[3] == one + ulp                  # above the tie -> up
    assert got[4] == np.float32(-2.5) and got[5] == 0.0


def _random_params(rng, scale=0.3):
    p = rng.normal(0, scale, mo.NPARAM)
    W1 = p[mo.OFF_W1:mo.OFF_W2].reshape(mo.NH, mo.K1)
    W1[:, mo.NF + 1:] = 0.0
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = _random_params(rng)
    t = np.concatenate([rng.normal(0, 1, (40, mo.NF)), rng.uniform(0, 2, (40, mo.NO))], axis=1)
    loss, g = mo.loss_grad(p, t, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_W1, mo.OFF_W2).reshape(mo.NH, mo.K1)[:, :mo.NF + 1]
                                     .reshape(-1), 20, replace=False),
                          rng.choice(np.arange(mo.OFF_W2, mo.NPARAM), 60, replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, t, rounding=False)[0] - mo.loss_grad(pm, t, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the padded input columns carry no gradient
    gW1 = g[mo.OFF_W1:mo.OFF_W2].reshape(mo.NH, mo.K1)
    assert np.all(gW1[:, mo.NF + 1:] == 0.0)


def test_rounded_gradient_close_to_exact():
    """BF16 rounding changes the gradient by about the BF16 precision only."""
    rng = np.random.default_rng(3)
    p = _random_params(rng)
    t = np.concatenate([rng.normal(0, 1, (500, mo.NF)), rng.uniform(0, 2, (500, mo.NO))], axis=1)
    _, g64 = mo.loss_grad(p, t, rounding=False)
    _, g16 = mo.loss_grad(p, t, rounding=True)
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)


@pytest.mark.parametrize("deck,ncyc", [("sedov", 6), ("noh", 4), ("leblanc", 5)])
def test_targets_reproduce_reference_qcs(deck, ncyc):
    text = ref_deck(deck)
    q1 = deck_value(text, "q1", 0.0)
    q2 = deck_value(text, "q2", 2.0)
    qgamma = deck_value(text, "qgamma", 5.0 / 3.0)
    G = q2 * 0.25 * (qgamma + 1.0)
    o = Oracle(text)
    o.run(ncyc)
    o.begin_cycle()
    for k in range(14):                       # through setQCnForce
        assert o.step(k) == 0
    fields = {f: o.get(f) for f in ("pu0", "elen", "cdiv", "cevol", "cdu", "zss", "zrp", "zr", "zuc", "crmu")}
    maps = {m: o.map(m) for m in ("mapsp1", "mapsp2", "mapss3", "mapsz")}
    c, tup = mo.qcs_tuples(fields, maps, q1, q2, qgamma)
    assert c.size > 0
    assert np.array_equal(np.sort(c), np.nonzero(fields["cdiv"] < 0)[0])
    z = maps["mapsz"][c]
    zss = fields["zss"][z]
    du = fields["cdu"][c]
    zkur = q1 * tup[:, 10] * zss + G * tup[:, 11] * du
    crmu = zkur * fields["zrp"][z] * fields["cevol"][c]
    np.testing.assert_allclose(crmu, fields["crmu"][c], rtol=1e-12, atol=1e-300)
    # the targets depend on f9 alone: y0 = (1-r)/sqrt((1-r)^2 + r^2)
    r = tup[:, 9]
    nrm = np.sqrt((1 - r) ** 2 + r ** 2)
    np.testing.assert_allclose(tup[:, 10], (1 - r) / nrm, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(tup[:, 11], 1 + r / nrm, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(tup[:, 8], np.log1p(du / zss), rtol=1e-14)
    assert np.all(tup[:, 6] > 0) and np.all(tup[:, 6] <= 1)


def test_uniform_compression_features():
    """Noh's first cycle: every corner compresses (uniform radial inflow)."""
    text = wd.noh_pie(8, 6, cstop=3, dtinit=1e-4)
    o = Oracle(text)
    o.begin_cycle()
    for k in range(14):
        o.step(k)
    fields = {f: o.get(f) for f in ("pu0", "elen", "cdiv", "cevol", "cdu", "zss", "zrp", "zr", "zuc")}
    maps = {m: o.map(m) for m in ("mapsp1", "mapsp2", "mapss3", "mapsz")}
    c, tup = mo.qcs_tuples(fields, maps, 1.0, 1.0, 5.0 / 3.0)
    assert c.size == fields["cdiv"].size
    assert np.all(tup[:, 4] < 0)              # cdiv evol / zss
    assert np.all(tup[:, 7] > 1)              # density rises over the half step
