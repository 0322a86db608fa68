"""Input decks (.pnt text, src/InputFile.cc syntax) for the benchmark and test
workloads.  Parameters follow the reference decks (test/*/*.pnt) scaled to the
mesh sizes of the measured configurations.
"""
from __future__ import annotations

import math


def _deck(lines: dict) -> str:
    out = []
    for k, v in lines.items():
        if v is None:
            continue
        if isinstance(v, (list, tuple)):
            v = " ".join(repr(float(x)) if not isinstance(x, int) else str(x) for x in v)
        out.append(f"{k} {v}")
    return "\n".join(out) + "\n"


def sedov_rect(n: int, *, length: float = 1.125, energy: float = 0.2468, cstop: int | None = None,
               tstop: float = 1.0, bc_far: bool = False, dtinit: float | None = None) -> str:
    """Sedov blast on an n x n rect mesh of [0,length]^2: one hot zone at the
    origin holding total energy `energy` (RZ volume 2*pi*dx^3/2, the scaling
    behind the reference decks' einitsub 40.222 at 9^2 and 5027.7 at 45^2),
    reflecting x=0 and y=0 planes (and x=y=length when bc_far), test/sedov
    parameters (ssmin 0.1, q1 0.1, q2 1.0).  dtinit defaults to half the
    Courant dt of the hot zone."""
    dx = length / n
    einitsub = energy / (2.0 * math.pi * dx ** 3 / 2.0)
    if dtinit is None:
        c = math.sqrt(5.0 / 3.0 * (2.0 / 3.0) * einitsub)
        dtinit = float(f"{0.5 * 0.6 * dx / c:.3g}")
    return _deck({
        "cstop": cstop, "tstop": tstop, "meshtype": "rect",
        "meshparams": f"{n} {n} {length!r} {length!r}",
        "subregion": f"0.0 {dx!r} 0.0 {dx!r}",
        "rinit": 1.0, "einit": 0.0, "rinitsub": 1.0, "einitsub": f"{einitsub:.6g}",
        "bcx": f"0.0 {length!r}" if bc_far else "0.0",
        "bcy": f"0.0 {length!r}" if bc_far else "0.0",
        "ssmin": 0.1, "q1": 0.1, "q2": 1.0, "alfa": 0.5, "cfl": 0.6, "cflv": 0.1,
        "dtinit": f"{dtinit:.6g}",
    })


def noh_pie(nr: int, nth: int | None = None, *, speed: float = 1.0, q1: float = 1.0,
            cstop: int | None = None, tstop: float = 0.6, dtinit: float = 1.e-7) -> str:
    """Noh implosion on a 90-degree pie mesh of radius 1 (nth angular x nr
    radial zones; triangles at the centre), inflow `speed` radially inward."""
    nth = nr if nth is None else nth
    return _deck({
        "cstop": cstop, "tstop": tstop, "meshtype": "pie", "meshparams": f"{nth} {nr}",
        "rinit": 1.0, "einit": 0.0, "uinitradial": -float(speed), "bcx": "0.0", "bcy": "0.0",
        "ssmin": 0.1, "q1": q1, "q2": 1.0, "alfa": 0.5, "cfl": 0.6, "dtinit": f"{dtinit:.6g}",
    })


def leblanc_rect(nx: int, ny: int, *, cstop: int | None = None, tstop: float = 6.0,
                 dtinit: float = 1.e-4) -> str:
    """LeBlanc shock tube along x on [0,9]x[0,0.5625]: rho=1, e=0.1 left of
    x=3; rho=1e-3, e=1e-7 right (test/leblanc parameters, tube turned to x)."""
    return _deck({
        "cstop": cstop, "tstop": tstop, "meshtype": "rect", "meshparams": f"{nx} {ny} 9.0 0.5625",
        "subregion": "3.0 9.0 0.0 0.5625", "rinit": 1.0, "einit": 0.1, "rinitsub": 1.e-3,
        "einitsub": 1.e-7, "bcx": "0.0 9.0", "bcy": "0.0 0.5625", "ssmin": 0.1, "q1": 0.3,
        "q2": 2.0, "dtinit": f"{dtinit:.6g}",
    })


# Measured configurations (sizes from the benchmark definition)
def sedov_1440(cstop: int | None = None) -> str:
    """Sedov 1440 x 1440 rect: 2,073,600 zones, 8,294,400 sides."""
    return sedov_rect(1440, cstop=cstop)


def noh_pie_1024(cstop: int | None = None) -> str:
    """Noh 1024 x 1024 pie: 1,048,576 zones, 4,193,280 sides."""
    return noh_pie(1024, 1024, cstop=cstop)


def leblanc_4096x256(cstop: int | None = None) -> str:
    return leblanc_rect(4096, 256, cstop=cstop)


def sedov_90(cstop: int | None = None) -> str:
    """The CPU-sized Sedov case: 90 x 90, einitsub 40222.1, dtinit 2.5e-5."""
    return sedov_rect(90, cstop=cstop, dtinit=2.5e-5)


def ensemble(count: int = 256, seed: int = 20261202, cstop: int | None = None):
    """The batched ensemble: for k even a Sedov blast on an n x n rect mesh
    with energy E, for k odd a Noh implosion on an n x n pie mesh with inflow
    speed s; n log-uniform in [32, 512], E in [0.1, 1.0], s in [0.5, 2.0]
    (std::mt19937_64(seed), pennant_ensemble_params).  Returns (decks, n, param)."""
    import ctypes

    import numpy as np

    from . import lib
    n = np.zeros(count, dtype=np.int32)
    par = np.zeros(count, dtype=np.float64)
    if lib().pennant_ensemble_params(ctypes.c_uint64(seed), int(count), n.ctypes.data, par.ctypes.data) != 0:
        raise ValueError("bad ensemble size")
    decks = []
    for k in range(count):
        if k % 2 == 0:
            decks.append(sedov_rect(int(n[k]), energy=float(par[k]), cstop=cstop))
        else:
            decks.append(noh_pie(int(n[k]), int(n[k]), speed=float(par[k]), cstop=cstop))
    return decks, n, par
