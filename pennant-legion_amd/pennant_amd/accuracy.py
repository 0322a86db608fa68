"""Error norms used to state the accuracy of the mixed-precision form
against the plain FP64 form (which reproduces the reference bit for bit)."""
from __future__ import annotations

import numpy as np


def rel_l1(got, want) -> float:
    """sum |got - want| / sum |want| over all zones (unweighted)."""
    got = np.asarray(got, dtype=np.float64).reshape(-1)
    want = np.asarray(want, dtype=np.float64).reshape(-1)
    den = float(np.sum(np.abs(want)))
    return float(np.sum(np.abs(got - want))) / den if den > 0 else float(np.sum(np.abs(got)))


def rel_linf(got, want) -> float:
    """max |got - want| / max |want|."""
    got = np.asarray(got, dtype=np.float64).reshape(-1)
    want = np.asarray(want, dtype=np.float64).reshape(-1)
    den = float(np.max(np.abs(want))) if want.size else 0.0
    d = float(np.max(np.abs(got - want))) if want.size else 0.0
    return d / den if den > 0 else d


def field_errors(got: dict, want: dict, names=("zr", "ze", "zp")) -> dict:
    """{name: (rel_l1, rel_linf)} for the named zone fields."""
    return {n: (rel_l1(got[n], want[n]), rel_linf(got[n], want[n])) for n in names}
