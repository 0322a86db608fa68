"""Learned QCS closure: the network (training and inference run on the GPU
through libpennant_hip.so) and the workflow around it -- collecting training
tuples from mixed-precision runs, training, fitting the feature envelope,
saving and loading.

Features, outputs and training targets are defined in csrc/closure.h; a
tuple is 12 float32 values: 10 features, then the exact linear and quadratic
coefficients.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import MODE_MIXED, PennantError, Problem, lib
from .device import DeviceBuffer

NFEAT, NOUT, TUPLE = 10, 2, 12


def _chk(rc, what):
    if rc != 0:
        raise PennantError(f"{what} failed ({rc})")


class Mlp:
    """The 10 -> 64 -> 64 -> 2 closure network (BF16 weights and activations,
    FP32 accumulation), with FP32 master parameters on the device."""

    def __init__(self, seed: int = 0):
        h = lib().pennant_mlp_create(int(seed))
        if not h:
            raise PennantError("pennant_mlp_create failed")
        self._h = ctypes.c_void_p(h)
        self.nparam = lib().pennant_mlp_num_params()

    # -- parameters -------------------------------------------------------
    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        _chk(lib().pennant_mlp_get_params(self._h, out.ctypes.data, self.nparam), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters")
        _chk(lib().pennant_mlp_set_params(self._h, p.ctypes.data, self.nparam), "set_params")

    def get_envelope(self):
        lo = np.empty(NFEAT, dtype=np.float32)
        hi = np.empty(NFEAT, dtype=np.float32)
        _chk(lib().pennant_mlp_get_envelope(self._h, lo.ctypes.data, hi.ctypes.data), "get_envelope")
        return lo, hi

    def set_envelope(self, lo, hi):
        lo = np.ascontiguousarray(lo, dtype=np.float32)
        hi = np.ascontiguousarray(hi, dtype=np.float32)
        _chk(lib().pennant_mlp_set_envelope(self._h, lo.ctypes.data, hi.ctypes.data), "set_envelope")

    def fit_envelope(self, tuples: np.ndarray):
        """The envelope is the bounding box of the training features."""
        f = np.asarray(tuples, dtype=np.float32)[:, :NFEAT]
        self.set_envelope(f.min(axis=0), f.max(axis=0))

    # -- compute ----------------------------------------------------------
    def forward(self, x: np.ndarray, path: int = 0) -> np.ndarray:
        """Outputs (n x 2) for features x (n x 10); path 0 runs the training
        GEMM kernels, path 1 the wave-level routine of the fused predictor."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEAT)
        n = x.shape[0]
        dx = DeviceBuffer.from_host(x)
        dy = DeviceBuffer((n, NOUT), np.float32)
        _chk(lib().pennant_mlp_forward(self._h, dx.ptr, n, dy.ptr, int(path)), "forward")
        return dy.numpy()

    def loss_grad(self, tuples: np.ndarray):
        """(loss, gradient) over all the tuples (n x 12)."""
        t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, TUPLE)
        dt = DeviceBuffer.from_host(t)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = ctypes.c_float()
        _chk(lib().pennant_mlp_loss_grad(self._h, dt.ptr, t.shape[0], g.ctypes.data, ctypes.byref(loss)),
             "loss_grad")
        return float(loss.value), g

    def train(self, tuples, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        """Adam steps on minibatches drawn from the tuples (a host array or a
        DeviceBuffer); returns the per-step minibatch losses."""
        if isinstance(tuples, DeviceBuffer):
            dt, n = tuples, tuples.shape[0]
        else:
            t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, TUPLE)
            dt, n = DeviceBuffer.from_host(t), t.shape[0]
        losses = np.empty(max(steps, 1), dtype=np.float32)
        _chk(lib().pennant_mlp_train(self._h, dt.ptr, int(n), int(steps), int(batch), float(lr), int(seed),
                                     losses.ctypes.data), "train")
        return losses[:steps]

    # -- persistence ------------------------------------------------------
    def save(self, path):
        lo, hi = self.get_envelope()
        np.savez(path, params=self.get_params(), lo=lo, hi=hi)

    @classmethod
    def load(cls, path) -> "Mlp":
        d = np.load(path, allow_pickle=False)
        m = cls(0)
        m.set_params(d["params"])
        m.set_envelope(d["lo"], d["hi"])
        return m

    def close(self):
        if getattr(self, "_h", None):
            lib().pennant_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def collect_tuples(deck_texts, cycles: int, every: int = 1, cap_per_run: int = 1 << 20,
                   max_total: int | None = None) -> np.ndarray:
    """Runs each deck in mixed precision for `cycles` cycles, capturing the
    tuples of one cycle in every `every` (the exact QCS of the mixed form
    generates them); returns them concatenated (n x 12 float32)."""
    out, total = [], 0
    buf = DeviceBuffer((cap_per_run, TUPLE), np.float32)
    for text in deck_texts:
        with Problem(text) as pb:
            done = 0
            while done < cycles:
                k = min(every - 1, cycles - done)
                if k > 0:
                    pb.run(k, MODE_MIXED)
                    done += k
                if done >= cycles or pb.scalars()["done"]:
                    break
                pb.set_capture(buf, cap_per_run)
                pb.run(1, MODE_MIXED)
                n = min(pb.capture_count(), cap_per_run)
                pb.set_capture(None, 0)
                done += 1
                if n:
                    out.append(buf.numpy(rows=n))
                    total += n
                if max_total is not None and total >= max_total:
                    break
                if pb.scalars()["done"]:
                    break
        if max_total is not None and total >= max_total:
            break
    if not out:
        return np.zeros((0, TUPLE), dtype=np.float32)
    t = np.concatenate(out)
    return t[:max_total] if max_total is not None else t
