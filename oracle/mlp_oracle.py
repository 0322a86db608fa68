"""Learned QCS closure oracle -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the closure network (forward, loss, backward) and of
its per-corner features and training targets.  Only tests/ may import it, as
the reference the GPU kernels are compared with; the product never does.

Network (csrc/closure.h): 10 features -> 64 -> 64 -> 2, relu, BF16 weights
and activations, FP32 accumulation.  With `rounding=True` the oracle rounds
to BF16 at the points where the GPU does (inputs, weights, relu outputs, the
gradients fed to the next product) and accumulates in float64; with
`rounding=False` everything is float64, which is what the finite-difference
test pins.

Features and targets are computed from the CPU hydro oracle's QCS fields at
the predictor half step (after setQCnForce, src/QCS.cc:246-327); the
targets are pinned against the oracle's own crmu (src/QCS.cc:285-296) by
the identity zkur = q1 y0 zss + G y1 du.

Parity: forward/backward pinned by central finite differences (float64);
features and targets pinned by the crmu identity above.
"""
from __future__ import annotations

import numpy as np

NF, NH, NO, K1 = 10, 64, 2, 16
OFF_W1 = 0
OFF_W2 = OFF_W1 + NH * K1
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NO * NH
NPARAM = OFF_B3 + NO


def bf16(x) -> np.ndarray:
    """Round float32 values to BF16 (round to nearest even), as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def unpack(p):
    p = np.asarray(p)
    W1 = p[OFF_W1:OFF_W2].reshape(NH, K1)
    W2 = p[OFF_W2:OFF_B2].reshape(NH, NH)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(NO, NH)
    b3 = p[OFF_B3:NPARAM]
    return W1, W2, b2, W3, b3


def _inputs(x):
    x = np.asarray(x, dtype=np.float64).reshape(-1, NF)
    X = np.zeros((x.shape[0], K1))
    X[:, :NF] = x
    X[:, NF] = 1.0
    return X


def forward(p, x, rounding=True):
    """Outputs (n x 2) and the activations the backward pass needs."""
    W1, W2, b2, W3, b3 = unpack(p)
    r = (lambda a: bf16(a).astype(np.float64)) if rounding else (lambda a: np.asarray(a, np.float64))
    X = r(_inputs(x))
    W1r, W2r, W3r = r(W1), r(W2), r(W3)
    Z1 = X @ W1r.T
    H1 = r(np.maximum(Z1, 0.0).astype(np.float32) if rounding else np.maximum(Z1, 0.0))
    Z2 = H1 @ W2r.T + np.asarray(b2, np.float64)
    H2 = r(np.maximum(Z2, 0.0).astype(np.float32) if rounding else np.maximum(Z2, 0.0))
    Y = H2 @ W3r.T + np.asarray(b3, np.float64)
    return Y, (X, H1, H2, W2r, W3r)


def loss_grad(p, tuples, rounding=True):
    """Loss (mean over tuples of the summed squared output error) and its
    gradient in the flat parameter layout."""
    t = np.asarray(tuples, dtype=np.float64).reshape(-1, NF + NO)
    B = t.shape[0]
    Y, (X, H1, H2, W2r, W3r) = forward(p, t[:, :NF], rounding)
    T = t[:, NF:]
    d = Y - T
    loss = float(np.sum(d * d) / B)
    r = (lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)) if rounding else (lambda a: a)
    dY = r(2.0 * d / B)
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (dY.T @ H2).reshape(-1)
    g[OFF_B3:NPARAM] = dY.sum(axis=0)
    dZ2 = r((dY @ W3r) * (H2 > 0))
    g[OFF_W2:OFF_B2] = (dZ2.T @ H1).reshape(-1)
    g[OFF_B2:OFF_W3] = dZ2.sum(axis=0)
    dZ1 = r((dZ2 @ W2r) * (H1 > 0))
    gW1 = dZ1.T @ X
    gW1[:, NF + 1:] = 0.0
    g[OFF_W1:OFF_W2] = gW1.reshape(-1)
    return loss, g


def slog(x):
    """sign(x) log(1 + |x|)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.log1p(np.abs(x))


def exact_targets(q1, G, zss, du):
    """Linear and quadratic coefficients whose combination q1 y0 zss + G y1 du
    is the Kurapatenko coefficient zkur = t2 + sqrt(t1^2 + t2^2)."""
    t1 = q1 * np.asarray(zss, np.float64)
    t2 = G * np.asarray(du, np.float64)
    s = np.sqrt(t1 * t1 + t2 * t2)
    safe = np.where(s > 0, s, 1.0)
    y0 = np.where(s > 0, t1 / safe, 1.0)
    y1 = np.where(s > 0, 1.0 + t2 / safe, 1.0)
    return y0, y1


def qcs_tuples(fields: dict, maps: dict, q1: float, q2: float, qgamma: float):
    """Tuples (features, targets) of the compressing corners, from the CPU
    hydro oracle's fields after setQCnForce: pu0, elen, cdiv, cevol, cdu,
    zss, zrp, zr, zuc, and the maps mapsp1, mapsp2, mapss3, mapsz.  Returns
    (corner indices, n x 12 array)."""
    G = q2 * 0.25 * (qgamma + 1.0)
    cdiv = fields["cdiv"]
    c = np.nonzero(cdiv < 0)[0]
    z = maps["mapsz"][c]
    p = maps["mapsp1"][c]
    s3 = maps["mapss3"][c]
    p1 = maps["mapsp1"][s3]
    p2 = maps["mapsp2"][c]
    pu = fields["pu0"]
    zss = fields["zss"][z]
    evol = fields["cevol"][c]
    du = fields["cdu"][c]
    el = fields["elen"]
    f = np.zeros((c.size, NF + NO))
    f[:, 0:2] = slog((pu[p] - pu[p1]) / zss[:, None])
    f[:, 2:4] = slog((pu[p2] - pu[p]) / zss[:, None])
    f[:, 4] = slog(cdiv[c] * evol / zss)
    f[:, 5] = evol / np.minimum(el[s3], el[c])
    up1 = 0.5 * (pu[p] + pu[p2])
    up3 = 0.5 * (pu[p1] + pu[p])
    duav = 0.25 * (pu[p] + up1 + fields["zuc"][z] + up3)
    f[:, 6] = zss / (zss + np.sqrt(np.sum(duav * duav, axis=1)))
    f[:, 7] = fields["zrp"][z] / fields["zr"][z]
    f[:, 8] = slog(du / zss)
    tl, tq = q1 * zss, G * du
    den = tl + tq
    f[:, 9] = np.where(den > 0, tq / np.where(den > 0, den, 1.0), 0.0)
    y0, y1 = exact_targets(q1, G, zss, du)
    f[:, 10] = y0
    f[:, 11] = y1
    return c, f
