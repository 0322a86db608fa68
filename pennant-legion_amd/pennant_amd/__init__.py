"""Python binding of the gfx950 PENNANT hydro library (libpennant_hip.so).

The library's C ABI (include/pennant_hip.h) is the product: task entry points
``pennant_hip_<task>`` mirroring the reference's Legion task IDs, and a
whole-problem driver ``pennant_*`` mirroring Driver/Hydro (src/Driver.cc,
src/Hydro.cc).  This module only loads the shared library with ctypes and
wraps the driver calls; it has no compute of its own and no fallback: if the
library is missing, importing :class:`Problem` or calling :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpennant_hip.so"

MODE_TASKS = 0
MODE_FUSED = 1
MODE_MIXED = 2     # fused, FP32 side/corner/QCS math, FP64 state, volumes, energy and dt
MODE_LEARNED = 3   # the mixed form with the learned QCS coefficient (Problem.set_closure)
EBADSIDES = -100
EINVAL = -22

# doCycle step numbering shared with the oracle (src/Hydro.cc:342-983 order)
NSTEPS = 30

_lib = None


class PennantError(RuntimeError):
    pass


class _Vec(ctypes.Structure):
    """pennant_vec / pennant_cvec: element i is (x[i*stride], y[i*stride])."""
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("stride", ctypes.c_int64)]


class Mesh(ctypes.Structure):
    """pennant_mesh: device connectivity (include/pennant_hip.h)."""
    _fields_ = [("nump", ctypes.c_int32), ("numz", ctypes.c_int32), ("nums", ctypes.c_int32),
                ("maxznump", ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in ("mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4", "znump",
                                       "zsbeg", "pcbeg", "pcside")]


def vec(x, y, stride: int = 1):
    """Vector-field argument: components at device addresses x and y, element
    i at x[i*stride], y[i*stride] (two arrays: stride 1)."""
    return _Vec(ctypes.c_void_p(int(x)), ctypes.c_void_p(int(y)), int(stride))


def vec_interleaved(base):
    """Vector-field argument for one interleaved double2 array at device
    address `base` (the reference's field instance): x = base, y = base + 8
    bytes, stride 2."""
    return _Vec(ctypes.c_void_p(int(base)), ctypes.c_void_p(int(base) + 8), 2)


# pennant_report_fn: void (*)(const char* text, void* user)
REPORT_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)

# argument kinds used to declare the task entry points
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int32
_M = ctypes.POINTER(Mesh)
_V = _Vec

TASK_SIGNATURES = {
    "pennant_hip_calc_ctrs": [_M, _V, _V, _V, _P],
    "pennant_hip_calc_vols": [_M, _V, _V, _P, _P, _P, _P, _P, _P, _P],
    "pennant_hip_calc_side_fracs": [_M, _P, _P, _P, _P],
    "pennant_hip_calc_surf_vecs": [_M, _V, _V, _V, _P],
    "pennant_hip_calc_edge_len": [_M, _V, _P, _P],
    "pennant_hip_calc_char_len": [_M, _P, _P, _P, _P],
    "pennant_hip_adv_pos_half": [_M, _V, _V, _V, _P, _P],
    "pennant_hip_calc_rho": [_M, _P, _P, _P, _P],
    "pennant_hip_calc_crnr_mass": [_M, _P, _P, _P, _P, _P],
    "pennant_hip_sum_crnr_force": [_M, _V, _V, _V, _V, _P],
    "pennant_hip_calc_accel": [_M, _V, _P, _V, _P],
    "pennant_hip_adv_pos_full": [_M, _V, _V, _V, _V, _V, _P, _P],
    "pennant_hip_calc_work": [_M, _V, _V, _V, _V, _V, _P, _P, _P, _P],
    "pennant_hip_calc_work_rate": [_M, _P, _P, _P, _P, _P, _P, _P],
    "pennant_hip_calc_energy": [_M, _P, _P, _P, _P],
    "pennant_hip_calc_dt_new": [_M, _P, _P, _P, _D, _P, _P, _P],
    "pennant_hip_calc_dt_new_noalloc": [_M, _P, _P, _P, _D, _P, _P],
    "pennant_hip_calc_dvol": [_M, _P, _P, _P, _P, _P],
    "pennant_hip_calc_dt": [_P, _P, _P, _D, _P, _P],
    "pennant_hip_calc_state_half": [_M, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P],
    "pennant_hip_calc_force_pgas": [_M, _V, _P, _V, _P],
    "pennant_hip_calc_force_tts": [_M, _P, _P, _V, _P, _P, _P, _V, _D, _D, _P],
    "pennant_hip_set_corner_div": [_M, _V, _P, _V, _V, _V, _V, _P, _P, _P, _P, _P, _P],
    "pennant_hip_set_qcn_force": [_M, _P, _P, _P, _P, _P, _P, _V, _P, _V, _V, _D, _D, _D, _P],
    "pennant_hip_set_force_qcs": [_M, _P, _V, _V, _P, _P, _P, _V, _P],
    "pennant_hip_set_vel_diff": [_M, _P, _P, _V, _V, _P, _P, _D, _D, _P],
    "pennant_hip_apply_fixed_bc": [_P, _I, _D, _D, _V, _V, _P],
    "pennant_hip_calc_global_dt": [_P, _P, _P, _I, _D, _D, _D, _D, _P, _P],
}

DRIVER_SIGNATURES = {
    "pennant_create": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int], ctypes.c_void_p),
    "pennant_destroy": ([_P], None),
    "pennant_launch_cycles": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pennant_sync": ([_P], ctypes.c_int),
    "pennant_run": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pennant_error": ([_P], ctypes.c_char_p),
    "pennant_drive": ([_P, ctypes.c_int, REPORT_FN, _P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_sizes": ([_P, ctypes.POINTER(ctypes.c_int64)], None),
    "pennant_scalars": ([_P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_set_scalars": ([_P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_field_info": ([_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "pennant_get_field": ([_P, ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "pennant_set_field": ([_P, ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "pennant_field_device": ([_P, ctypes.c_char_p, ctypes.c_int], ctypes.c_void_p),
    "pennant_get_map": ([_P, ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)],
                        ctypes.c_int),
    "pennant_mesh_desc": ([_P], ctypes.POINTER(Mesh)),
    "pennant_stream": ([_P], ctypes.c_void_p),
    "pennant_begin_cycle": ([_P], ctypes.c_int),
    "pennant_step": ([_P, ctypes.c_int], ctypes.c_int),
    "pennant_end_cycle": ([_P], ctypes.c_int),
    "pennant_write_xy": ([_P, ctypes.c_char_p], ctypes.c_int),
    "pennant_write_gold": ([_P, ctypes.c_char_p], ctypes.c_int),
    "pennant_mlp_num_params": ([], ctypes.c_int),
    "pennant_mlp_create": ([ctypes.c_uint64], ctypes.c_void_p),
    "pennant_mlp_destroy": ([_P], None),
    "pennant_mlp_set_params": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "pennant_mlp_get_params": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "pennant_mlp_set_envelope": ([_P, _P, _P], ctypes.c_int),
    "pennant_mlp_get_envelope": ([_P, _P, _P], ctypes.c_int),
    "pennant_mlp_forward": ([_P, _P, ctypes.c_int64, _P, ctypes.c_int], ctypes.c_int),
    "pennant_mlp_loss_grad": ([_P, _P, ctypes.c_int64, _P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "pennant_mlp_train": ([_P, _P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                           ctypes.c_uint64, _P], ctypes.c_int),
    "pennant_set_closure": ([_P, _P], ctypes.c_int),
    "pennant_closure_stats": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_set_capture": ([_P, _P, ctypes.c_int64], ctypes.c_int),
    "pennant_capture_count": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_energy": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_batch_create": ([ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
                             ctypes.c_void_p),
    "pennant_batch_destroy": ([_P], None),
    "pennant_batch_count": ([_P], ctypes.c_int),
    "pennant_batch_sizes": ([_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_batch_error": ([_P], ctypes.c_char_p),
    "pennant_batch_launch_cycles": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pennant_batch_sync": ([_P], ctypes.c_int),
    "pennant_batch_active": ([_P], ctypes.c_int),
    "pennant_batch_time_cycles": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                                  ctypes.c_int),
    "pennant_batch_scalars": ([_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_batch_get_field": ([_P, ctypes.c_int, ctypes.c_char_p, _P], ctypes.c_int),
    "pennant_batch_set_closure": ([_P, _P], ctypes.c_int),
    "pennant_batch_closure_stats": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_batch_set_capture": ([_P, _P, ctypes.c_int64], ctypes.c_int),
    "pennant_batch_capture_count": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_batch_set_tuning": ([_P, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "pennant_batch_get_tuning": ([_P, ctypes.c_char_p], ctypes.c_int),
    "pennant_batch_autotune": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_ensemble_params": ([ctypes.c_uint64, ctypes.c_int, _P, _P], ctypes.c_int),
    "pennant_write_gold_host": ([ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, _P, _P, _P, _P,
                                 ctypes.c_int32, _P, _P, ctypes.c_int32, _P], ctypes.c_int),
    "pennant_hip_reduce_workspace_bytes": ([], ctypes.c_size_t),
    "pennant_host_mesh_create": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int], ctypes.c_void_p),
    "pennant_host_mesh_get": ([_P, ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)],
                              ctypes.c_int),
    "pennant_host_mesh_destroy": ([_P], None),
    "pennant_time_cycles": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_set_tuning": ([_P, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "pennant_autotune": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_bench_copy": ([ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                           ctypes.c_int),
    "pennant_get_tuning": ([_P, ctypes.c_char_p], ctypes.c_int),
    "pennant_lookahead_redos": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_pieces_create": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int], ctypes.c_void_p),
    "pennant_pieces_destroy": ([_P], None),
    "pennant_pieces_plan": ([ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64,
                             ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "pennant_pieces_error": ([_P], ctypes.c_char_p),
    "pennant_pieces_count": ([_P], ctypes.c_int),
    "pennant_pieces_layout": ([_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "pennant_pieces_run": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pennant_pieces_scalars": ([_P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "pennant_pieces_time_cycles": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                                   ctypes.c_int),
    "pennant_pieces_field_info": ([_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "pennant_pieces_get_field": ([_P, ctypes.c_char_p, _P], ctypes.c_int),
}


def lib() -> ctypes.CDLL:
    """Loads libpennant_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PennantError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C pennant-legion_amd/csrc`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, args in TASK_SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        for name, (args, res) in DRIVER_SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(TASK_SIGNATURES) + list(DRIVER_SIGNATURES)


class HostMesh:
    """Deck parsing and mesh generation without a device (host logic only)."""

    def __init__(self, deck_text: str):
        L = lib()
        err = ctypes.create_string_buffer(512)
        h = L.pennant_host_mesh_create(deck_text.encode(), err, len(err))
        if not h:
            raise PennantError(err.value.decode())
        self._h = ctypes.c_void_p(h)

    def get(self, name: str) -> np.ndarray:
        L = lib()
        n = ctypes.c_int64()
        if L.pennant_host_mesh_get(self._h, name.encode(), None, ctypes.byref(n)) != 0:
            raise KeyError(name)
        out = np.empty(n.value, dtype=np.float64 if name in ("px", "py") else np.int32)
        L.pennant_host_mesh_get(self._h, name.encode(), out.ctypes.data, ctypes.byref(n))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pennant_host_mesh_destroy(h)
            self._h = None


class Problem:
    """One PENNANT problem resident on the current HIP device."""

    def __init__(self, deck_text: str):
        L = lib()
        err = ctypes.create_string_buffer(512)
        h = L.pennant_create(deck_text.encode(), err, len(err))
        if not h:
            raise PennantError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        sz = (ctypes.c_int64 * 4)()
        L.pennant_sizes(self._h, sz)
        self.nump, self.numz, self.nums, self.nbcs = (int(v) for v in sz)

    @classmethod
    def from_file(cls, path) -> "Problem":
        return cls(Path(path).read_text())

    def close(self):
        if getattr(self, "_h", None):
            lib().pennant_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0 and rc == EBADSIDES:
            raise PennantError(lib().pennant_error(self._h).decode())
        if rc != 0:
            raise PennantError(f"HIP/pennant error {rc}")

    # -- time loop ---------------------------------------------------------
    def run(self, ncycles: int, mode: int = MODE_FUSED) -> int:
        rc = lib().pennant_run(self._h, int(ncycles), int(mode))
        if rc < 0:
            self._check(rc)
        return rc

    def drive(self, mode: int = MODE_FUSED, echo: bool = False):
        """Driver::run: runs to the stop condition with the reference's report
        lines and end-of-run summary (pennant_drive).  Returns (cycles run,
        the report text, hydro cycle run time in us); echo also prints it."""
        parts = []

        def _cb(text, _user):
            t = text.decode()
            parts.append(t)
            if echo:
                print(t, end="", flush=True)

        cb = REPORT_FN(_cb)
        us = ctypes.c_double()
        rc = lib().pennant_drive(self._h, int(mode), cb, None, ctypes.byref(us))
        if rc < 0:
            self._check(rc)
        return rc, "".join(parts), us.value

    def launch(self, ncycles: int, mode: int = MODE_FUSED):
        self._check(lib().pennant_launch_cycles(self._h, int(ncycles), int(mode)))

    def sync(self):
        self._check(lib().pennant_sync(self._h))

    def time_cycles(self, ncycles: int, mode: int = MODE_FUSED, phases: bool = True):
        """(total_ms, phase_ms or None) for ncycles cycles timed with HIP
        events on the problem's stream; phase_ms holds the summed durations of
        the zone-predictor, point, zone-corrector and finalize kernels and of
        the learned pass (for the pipelined mixed cycle: zone pass, point
        kernel, 0, finalize with the predictor redo, learned pass).  Without phases the cycles run as
        pennant_launch_cycles runs them (captured graphs)."""
        tot = ctypes.c_double()
        ph = (ctypes.c_double * 5)() if phases else None
        self._check(lib().pennant_time_cycles(self._h, int(ncycles), int(mode), ctypes.byref(tot), ph))
        return tot.value, (list(ph) if phases and mode in (MODE_FUSED, MODE_MIXED, MODE_LEARNED) else None)

    def set_tuning(self, key: str, value: int):
        if lib().pennant_set_tuning(self._h, key.encode(), int(value)) != 0:
            raise ValueError(f"bad tuning {key}={value}")

    def autotune(self, mode: int = MODE_FUSED, cycles: int = 4) -> float:
        """Picks the fastest launch shape for `mode` (the state is restored)."""
        ms = ctypes.c_double()
        self._check(lib().pennant_autotune(self._h, int(mode), int(cycles), ctypes.byref(ms)))
        return ms.value

    def get_tuning(self, key: str) -> int:
        return lib().pennant_get_tuning(self._h, key.encode())

    def lookahead_redos(self) -> int:
        """Look-ahead predictors the pipelined mixed cycle has redone."""
        n = ctypes.c_int64()
        self._check(lib().pennant_lookahead_redos(self._h, ctypes.byref(n)))
        return int(n.value)

    def scalars(self) -> dict:
        out = (ctypes.c_double * 10)()
        self._check(lib().pennant_scalars(self._h, out))
        keys = ("time", "dt", "dtrec", "cycle", "dtnew", "dvovmax", "numsbad_pred",
                "numsbad_corr", "dtnext", "done")
        return dict(zip(keys, list(out)))

    def set_scalars(self, vals):
        arr = (ctypes.c_double * 10)(*vals)
        self._check(lib().pennant_set_scalars(self._h, arr))

    # -- fields ------------------------------------------------------------
    def field_info(self, name: str):
        n = ctypes.c_int64()
        c = ctypes.c_int()
        self._check(lib().pennant_field_info(self._h, name.encode(), ctypes.byref(n), ctypes.byref(c)))
        return n.value, c.value

    def get(self, name: str) -> np.ndarray:
        n, c = self.field_info(name)
        out = np.empty(n * c, dtype=np.float64)
        self._check(lib().pennant_get_field(self._h, name.encode(), out.ctypes.data))
        return out.reshape(n, 2) if c == 2 else out

    def set(self, name: str, values):
        n, c = self.field_info(name)
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if arr.size != n * c:
            raise ValueError(f"{name}: expected {n * c} values, got {arr.size}")
        self._check(lib().pennant_set_field(self._h, name.encode(), arr.ctypes.data))

    def device_ptr(self, name: str, comp: int = 0) -> int:
        p = lib().pennant_field_device(self._h, name.encode(), comp)
        if not p:
            raise PennantError(f"unknown field {name}[{comp}]")
        return int(p)

    def get_map(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        self._check(lib().pennant_get_map(self._h, name.encode(), None, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.int32)
        self._check(lib().pennant_get_map(self._h, name.encode(), out.ctypes.data, ctypes.byref(n)))
        return out

    def mesh_desc(self):
        return lib().pennant_mesh_desc(self._h)

    def stream(self) -> int:
        return int(lib().pennant_stream(self._h) or 0)

    # -- step-by-step (task mode) -----------------------------------------
    def step(self, k: int) -> int:
        return lib().pennant_step(self._h, int(k))

    def end_cycle(self):
        self._check(lib().pennant_end_cycle(self._h))

    def write_xy(self, path):
        self._check(lib().pennant_write_xy(self._h, str(path).encode()))

    def set_closure(self, mlp):
        """Attaches a closure network (pennant_amd.closure.Mlp) for MODE_LEARNED
        cycles; None detaches.  The problem keeps a reference to it."""
        self._closure = mlp
        self._check(lib().pennant_set_closure(self._h, mlp._h if mlp is not None else None))

    def closure_stats(self):
        """(zones evaluated by the network, zones that fell back to exact QCS)."""
        out = (ctypes.c_int64 * 2)()
        self._check(lib().pennant_closure_stats(self._h, out))
        return int(out[0]), int(out[1])

    def set_capture(self, buf, cap: int):
        """Records training tuples of MODE_MIXED cycles into a device buffer
        (pennant_amd.device.DeviceBuffer of cap x 12 float32); None stops."""
        self._capture = buf
        self._check(lib().pennant_set_capture(self._h, buf.ptr if buf is not None else None,
                                              int(cap) if buf is not None else 0))

    def capture_count(self) -> int:
        n = ctypes.c_int64()
        self._check(lib().pennant_capture_count(self._h, ctypes.byref(n)))
        return n.value

    def energy(self):
        """(internal, kinetic) energy of the current state."""
        ei, ek = ctypes.c_double(), ctypes.c_double()
        self._check(lib().pennant_energy(self._h, ctypes.byref(ei), ctypes.byref(ek)))
        return ei.value, ek.value

    def write_gold(self, base):
        """EnSight Gold files <base>.case/.geo/.zr/.ze/.zp of the current state."""
        self._check(lib().pennant_write_gold(self._h, str(base).encode()))


def write_gold_host(base, cycle: int, time: float, zr, ze, zp, znump, px, py, mapsp1):
    """The EnSight Gold writer on host arrays (no device needed)."""
    f64 = [np.ascontiguousarray(a, dtype=np.float64) for a in (zr, ze, zp, px, py)]
    i32 = [np.ascontiguousarray(a, dtype=np.int32) for a in (znump, mapsp1)]
    rc = lib().pennant_write_gold_host(str(base).encode(), int(cycle), float(time),
                                       f64[0].ctypes.data, f64[1].ctypes.data, f64[2].ctypes.data,
                                       i32[0].ctypes.data, int(i32[0].size), f64[3].ctypes.data,
                                       f64[4].ctypes.data, int(f64[3].size), i32[1].ctypes.data)
    if rc != 0:
        raise PennantError(f"cannot write {base}.*")


class Batch:
    """Many independent problems advanced together (one launch per phase
    over all of them); each keeps its own dt, parameters and stop test."""

    def __init__(self, deck_texts):
        texts = [t.encode() for t in deck_texts]
        arr = (ctypes.c_char_p * len(texts))(*texts)
        err = ctypes.create_string_buffer(1024)
        h = lib().pennant_batch_create(arr, len(texts), err, len(err))
        if not h:
            raise PennantError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        self.count = len(texts)

    def _check(self, rc):
        if rc == EBADSIDES:
            raise PennantError(lib().pennant_batch_error(self._h).decode())
        if rc != 0:
            raise PennantError(f"HIP/pennant error {rc}")

    def sizes(self, i: int):
        """(nump, numz, nums) of problem i."""
        out = (ctypes.c_int64 * 3)()
        self._check(lib().pennant_batch_sizes(self._h, int(i), out))
        return int(out[0]), int(out[1]), int(out[2])

    def launch(self, ncycles: int, mode: int = MODE_MIXED):
        self._check(lib().pennant_batch_launch_cycles(self._h, int(ncycles), int(mode)))

    def sync(self, raise_on_error: bool = True):
        """Waits.  A problem that stopped on negative side volumes raises
        PennantError unless raise_on_error is False (the other problems run
        on regardless; stopped() lists them)."""
        rc = lib().pennant_batch_sync(self._h)
        if rc == EBADSIDES and not raise_on_error:
            return
        self._check(rc)

    def stopped(self):
        """Indices of the problems that stopped on negative side volumes."""
        out = []
        for i in range(self.count):
            s = self.scalars(i)
            if s["numsbad_pred"] > 0 or s["numsbad_corr"] > 0:
                out.append(i)
        return out

    def active(self) -> int:
        n = lib().pennant_batch_active(self._h)
        if n < 0:
            self._check(n)
        return n

    def run(self, ncycles: int, mode: int = MODE_MIXED, raise_on_error: bool = True):
        self.launch(ncycles, mode)
        self.sync(raise_on_error)

    def run_to_end(self, mode: int = MODE_MIXED, chunk: int = 256, max_cycles: int = 10 ** 7):
        done = 0
        while done < max_cycles and self.active() > 0:
            self.run(chunk, mode)
            done += chunk
        return done

    def time_cycles(self, ncycles: int, mode: int = MODE_MIXED, raise_on_error: bool = True) -> float:
        ms = ctypes.c_double()
        rc = lib().pennant_batch_time_cycles(self._h, int(ncycles), int(mode), ctypes.byref(ms))
        if not (rc == EBADSIDES and not raise_on_error):
            self._check(rc)
        return ms.value

    def scalars(self, i: int) -> dict:
        out = (ctypes.c_double * 10)()
        self._check(lib().pennant_batch_scalars(self._h, int(i), out))
        keys = ("time", "dt", "dtrec", "cycle", "dtnew", "dvovmax", "numsbad_pred", "numsbad_corr",
                "dtnext", "done")
        d = dict(zip(keys, list(out)))
        d["cycle"] = int(d["cycle"])
        d["done"] = int(d["done"])
        return d

    def get(self, i: int, name: str) -> np.ndarray:
        nump, numz, _ = self.sizes(i)
        if name in ("px", "pu"):
            out = np.empty((nump, 2), dtype=np.float64)
        else:
            out = np.empty(numz, dtype=np.float64)
        self._check(lib().pennant_batch_get_field(self._h, int(i), name.encode(), out.ctypes.data))
        return out

    def set_closure(self, mlp):
        self._closure = mlp
        self._check(lib().pennant_batch_set_closure(self._h, mlp._h if mlp is not None else None))

    def closure_stats(self):
        out = (ctypes.c_int64 * 2)()
        self._check(lib().pennant_batch_closure_stats(self._h, out))
        return int(out[0]), int(out[1])

    def set_capture(self, buf, cap: int):
        self._capture = buf
        self._check(lib().pennant_batch_set_capture(self._h, buf.ptr if buf is not None else None,
                                                    int(cap) if buf is not None else 0))

    def capture_count(self) -> int:
        n = ctypes.c_int64()
        self._check(lib().pennant_batch_capture_count(self._h, ctypes.byref(n)))
        return n.value

    def set_tuning(self, key: str, value: int):
        if lib().pennant_batch_set_tuning(self._h, key.encode(), int(value)) != 0:
            raise ValueError(f"bad tuning {key}={value}")

    def get_tuning(self, key: str) -> int:
        return lib().pennant_batch_get_tuning(self._h, key.encode())

    def autotune(self, mode: int = MODE_MIXED, cycles: int = 8) -> float:
        ms = ctypes.c_double()
        self._check(lib().pennant_batch_autotune(self._h, int(mode), int(cycles), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            lib().pennant_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


class Pieces:
    """One mesh run as numpcs pieces on the current HIP device (the
    reference's -n <numpcs>): private, master and ghost points per piece,
    corners at ghost points summed by their owner in one-piece side order,
    the owners' point state copied back to the ghosts every cycle."""

    def __init__(self, deck_text: str, numpcs: int):
        L = lib()
        err = ctypes.create_string_buffer(512)
        h = L.pennant_pieces_create(deck_text.encode(), int(numpcs), err, len(err))
        if not h:
            raise PennantError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        self.count = int(L.pennant_pieces_count(self._h))

    def _check(self, rc):
        if rc == EBADSIDES:
            raise PennantError(lib().pennant_pieces_error(self._h).decode())
        if rc != 0:
            raise PennantError(f"pennant pieces call failed ({rc})")
        return rc

    def layout(self, i: int) -> dict:
        out = (ctypes.c_int64 * 8)()
        if lib().pennant_pieces_layout(self._h, int(i), out) != 0:
            raise IndexError(i)
        keys = ("piece", "zones", "sides", "private", "master", "ghost", "received")
        d = dict(zip(keys, (int(v) for v in out[:7])))
        d["numpcx"], d["numpcy"] = int(out[7]) >> 16, int(out[7]) & 0xFFFF
        return d

    def run(self, ncycles: int, mode: int = MODE_FUSED) -> int:
        rc = lib().pennant_pieces_run(self._h, int(ncycles), int(mode))
        if rc < 0:
            self._check(rc)
        return rc

    def time_cycles(self, ncycles: int, mode: int = MODE_FUSED) -> float:
        """Milliseconds for ncycles cycles, HIP events on the pieces' stream."""
        ms = ctypes.c_double()
        self._check(lib().pennant_pieces_time_cycles(self._h, int(ncycles), int(mode), ctypes.byref(ms)))
        return ms.value

    def scalars(self) -> dict:
        out = (ctypes.c_double * 10)()
        self._check(lib().pennant_pieces_scalars(self._h, out))
        keys = ("time", "dt", "dtrec", "cycle", "dtnew", "dvovmax", "numsbad_pred",
                "numsbad_corr", "dtnext", "done")
        return dict(zip(keys, list(out)))

    def get(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        c = ctypes.c_int()
        self._check(lib().pennant_pieces_field_info(self._h, name.encode(), ctypes.byref(n), ctypes.byref(c)))
        out = np.empty(n.value * c.value, dtype=np.float64)
        self._check(lib().pennant_pieces_get_field(self._h, name.encode(), out.ctypes.data))
        return out.reshape(n.value, 2) if c.value == 2 else out

    def close(self):
        if getattr(self, "_h", None):
            lib().pennant_pieces_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def pieces_plan(deck_text: str, numpcs: int) -> dict:
    """The partition of a deck's mesh into numpcs pieces, computed on the host
    (no device): numpcx, numpcy and per non-empty piece its number, zones,
    sides, private, master and ghost points and received corners."""
    n = 3 + 7 * max(int(numpcs), 1)
    out = (ctypes.c_int64 * n)()
    err = ctypes.create_string_buffer(512)
    rc = lib().pennant_pieces_plan(deck_text.encode(), int(numpcs), out, n, err, len(err))
    if rc < 0:
        raise PennantError(err.value.decode())
    keys = ("piece", "zones", "sides", "private", "master", "ghost", "received")
    pieces = [dict(zip(keys, (int(v) for v in out[3 + 7 * i: 10 + 7 * i]))) for i in range(int(out[2]))]
    return {"numpcx": int(out[0]), "numpcy": int(out[1]), "pieces": pieces}
