"""Trains the learned QCS closure on tuples that the exact QCS of the mixed
form generates while a batched ensemble of Sedov and Noh problems runs, and
saves the network (parameters and feature envelope) to an .npz file.

    python -m pennant_amd.train_closure --out closure.npz

Everything runs on the GPU: the ensemble (one launch per phase over all
problems), the tuple capture, and the training kernels.  Deterministic for a
given seed.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import MODE_MIXED, Batch
from . import decks as wd
from .closure import Mlp, TUPLE
from .device import DeviceBuffer

DEFAULT_OUT = Path(__file__).resolve().parent / "data" / "closure_ensemble.npz"


def capture_ensemble(count: int, seed: int, cycles: int, every: int, cap: int, max_tuples: int,
                     sample_seed: int = 0, log=print) -> np.ndarray:
    """Runs `count` ensemble problems for `cycles` mixed cycles, capturing one
    cycle in every `every`; keeps a uniform random sample of at most
    max_tuples of the captured tuples."""
    decks, n, _ = wd.ensemble(count, seed=seed, cstop=10 ** 7)
    log(f"ensemble: {count} problems, {int((n.astype(np.int64) ** 2).sum())} zones")
    rng = np.random.default_rng(sample_seed)
    out = []
    per = max(1, max_tuples // max(1, cycles // every))
    buf = DeviceBuffer((cap, TUPLE), np.float32)
    t0 = time.perf_counter()
    with Batch(decks) as b:
        done = 0
        while done < cycles:
            b.run(every - 1, MODE_MIXED, raise_on_error=False)
            b.set_capture(buf, cap)
            b.run(1, MODE_MIXED, raise_on_error=False)
            k = min(b.capture_count(), cap)
            b.set_capture(None, 0)
            done += every
            if k:
                t = buf.numpy(rows=k)
                if k > per:
                    t = t[np.sort(rng.choice(k, per, replace=False))]
                out.append(t)
        st = b.stopped()
        if st:
            log(f"problems stopped on negative side volumes (their earlier tuples are kept): {st}")
    tup = np.concatenate(out) if out else np.zeros((0, TUPLE), np.float32)
    log(f"captured {tup.shape[0]} tuples in {time.perf_counter() - t0:.1f}s")
    return tup


def train(tuples: np.ndarray, steps: int, batch: int, lr: float, seed: int, log=print) -> Mlp:
    m = Mlp(seed)
    t0 = time.perf_counter()
    losses = m.train(tuples, steps=steps, batch=batch, lr=lr, seed=seed)
    m.fit_envelope(tuples)
    log(f"trained {steps} steps in {time.perf_counter() - t0:.1f}s: loss {losses[0]:.3e} -> "
        f"{losses[-50:].mean():.3e}")
    return m


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_OUT))
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cycles", type=int, default=640)
    ap.add_argument("--every", type=int, default=16)
    ap.add_argument("--max-tuples", type=int, default=2_000_000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    tup = capture_ensemble(args.count, args.seed, args.cycles, args.every, cap=1 << 22,
                           max_tuples=args.max_tuples, log=log)
    m = train(tup, args.steps, args.batch, args.lr, seed=1, log=log)
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    m.save(args.out)
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
