"""ctypes wrapper of the CPU oracle (oracle/liboracle_pennant.so).

Test infrastructure only: the oracle is what the HIP path is compared
against, never part of the product.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle_pennant.so"

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(str(ORACLE_SO))
        vp, dp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.ora_destroy.argtypes = [vp]
        L.ora_error.restype = ctypes.c_char_p
        L.ora_error.argtypes = [vp]
        L.ora_set_omp.argtypes = [vp, ctypes.c_int]
        L.ora_omp_threads.restype = ctypes.c_int
        L.ora_run.argtypes = [vp, ctypes.c_int]
        L.ora_run.restype = ctypes.c_int
        L.ora_done.argtypes = [vp]
        L.ora_begin_cycle.argtypes = [vp]
        L.ora_begin_cycle.restype = ctypes.c_double
        L.ora_step.argtypes = [vp, ctypes.c_int]
        L.ora_step.restype = ctypes.c_int
        L.ora_end_cycle.argtypes = [vp]
        L.ora_nsteps.restype = ctypes.c_int
        L.ora_sizes.argtypes = [vp, i64p]
        L.ora_scalars.argtypes = [vp, dp]
        L.ora_set_scalars.argtypes = [vp, dp]
        L.ora_field.restype = dp
        L.ora_field.argtypes = [vp, ctypes.c_char_p, i64p, ctypes.POINTER(ctypes.c_int)]
        L.ora_map.restype = vp
        L.ora_map.argtypes = [vp, ctypes.c_char_p, i64p, ctypes.POINTER(ctypes.c_int)]
        L.ora_num_bcs.argtypes = [vp]
        L.ora_bc.restype = ctypes.POINTER(ctypes.c_int64)
        L.ora_bc.argtypes = [vp, ctypes.c_int, i64p, dp]
        L.ora_calc_dt_hydro.restype = ctypes.c_double
        L.ora_calc_dt_hydro.argtypes = [ctypes.c_double] * 4
        L.ora_calc_global_dt.restype = ctypes.c_double
        L.ora_calc_global_dt.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int] + [ctypes.c_double] * 3
        L.ora_write_xy.argtypes = [vp, ctypes.c_char_p]
        _lib = L
    return _lib


FIELDS = ["px", "pxp", "px0", "pu", "pu0", "pf", "pap", "pmaswt", "zx", "zxp", "zuc", "zarea",
          "zvol", "zareap", "zvolp", "zvol0", "zdl", "zm", "zr", "zrp", "ze", "zetot", "zw",
          "zwrate", "zp", "zss", "zdu", "ztmp", "ex", "exp", "ssurfp", "sfp", "sfq", "sft",
          "cqe1", "cqe2", "sarea", "svol", "sareap", "svolp", "elen", "smf", "carea", "cevol",
          "cdu", "cdiv", "ccos", "crmu", "cw"]

SCALAR_KEYS = ("time", "dt", "dtrec", "cycle", "dtnew", "dvovmax", "numsbad_pred", "numsbad_corr")


class OracleError(RuntimeError):
    pass


class Oracle:
    def __init__(self, deck_text: str, omp: bool = False):
        L = lib()
        err = ctypes.create_string_buffer(512)
        h = L.ora_create(deck_text.encode(), err, len(err))
        if not h:
            raise OracleError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        sz = (ctypes.c_int64 * 3)()
        L.ora_sizes(self._h, sz)
        self.nump, self.numz, self.nums = (int(v) for v in sz)
        if omp:
            L.ora_set_omp(self._h, 1)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ora_destroy(h)
            self._h = None

    def run(self, maxcycles: int) -> int:
        n = lib().ora_run(self._h, int(maxcycles))
        if n < 0:
            raise OracleError(lib().ora_error(self._h).decode())
        return n

    def done(self) -> bool:
        return bool(lib().ora_done(self._h))

    def begin_cycle(self) -> float:
        return lib().ora_begin_cycle(self._h)

    def step(self, k: int) -> int:
        return lib().ora_step(self._h, int(k))

    def end_cycle(self):
        lib().ora_end_cycle(self._h)

    def scalars(self) -> dict:
        out = (ctypes.c_double * 8)()
        lib().ora_scalars(self._h, out)
        return dict(zip(SCALAR_KEYS, list(out)))

    def field(self, name: str) -> np.ndarray:
        """Live view (no copy) of an oracle field; shape (n,) or (n, 2)."""
        n = ctypes.c_int64()
        c = ctypes.c_int()
        p = lib().ora_field(self._h, name.encode(), ctypes.byref(n), ctypes.byref(c))
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(n.value * c.value,))
        return a.reshape(n.value, 2) if c.value == 2 else a

    def get(self, name: str) -> np.ndarray:
        return self.field(name).copy()

    def map(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        w = ctypes.c_int()
        p = lib().ora_map(self._h, name.encode(), ctypes.byref(n), ctypes.byref(w))
        if not p:
            raise KeyError(name)
        ct = ctypes.c_int64 if w.value == 8 else ctypes.c_int32
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(n.value,)).copy()

    def bcs(self):
        out = []
        for i in range(lib().ora_num_bcs(self._h)):
            n = ctypes.c_int64()
            v = (ctypes.c_double * 2)()
            p = lib().ora_bc(self._h, i, ctypes.byref(n), v)
            pts = np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int64)
            out.append((pts, (v[0], v[1])))
        return out

    def write_xy(self, path):
        lib().ora_write_xy(self._h, str(path).encode())


def calc_dt_hydro(cflv, dtlast, dtnew, dvovmax):
    return lib().ora_calc_dt_hydro(cflv, dtlast, dtnew, dvovmax)


def calc_global_dt(dtmax, dtinit, dtfac, tstop, cycle, time, dtlast, dtrec):
    return lib().ora_calc_global_dt(dtmax, dtinit, dtfac, tstop, cycle, time, dtlast, dtrec)
