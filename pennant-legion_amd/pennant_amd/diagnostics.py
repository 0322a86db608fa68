"""Physics diagnostics used to compare the forms of the cycle (plain FP64,
mixed, learned closure): zone centroids, the shock radius and the total
energy of a state."""
from __future__ import annotations

import numpy as np


def zone_centroids(znump: np.ndarray, mapsp1: np.ndarray, px: np.ndarray) -> np.ndarray:
    """Mean of each zone's corner points (zones' sides are contiguous)."""
    znump = np.asarray(znump)
    starts = np.concatenate([[0], np.cumsum(znump)[:-1]])
    out = np.zeros((znump.size, 2))
    for n in np.unique(znump):
        z = np.nonzero(znump == n)[0]
        idx = starts[z][:, None] + np.arange(n)[None, :]
        out[z] = px[mapsp1[idx]].mean(axis=1)
    return out


def shock_radius(zr: np.ndarray, centroids: np.ndarray) -> float:
    """Radius (from the origin) of the densest zone: the position of the
    density peak behind a Sedov blast wave or a Noh shock."""
    z = int(np.argmax(zr))
    return float(np.hypot(centroids[z, 0], centroids[z, 1]))


def radial_profile(zr: np.ndarray, centroids: np.ndarray, nbins: int = 64, rmax: float | None = None):
    """Mean density in radial bins: (bin centres, mean density)."""
    r = np.hypot(centroids[:, 0], centroids[:, 1])
    rmax = float(r.max()) if rmax is None else rmax
    edges = np.linspace(0.0, rmax, nbins + 1)
    k = np.clip(np.digitize(r, edges) - 1, 0, nbins - 1)
    s = np.bincount(k, weights=zr, minlength=nbins)
    c = np.bincount(k, minlength=nbins)
    return 0.5 * (edges[1:] + edges[:-1]), s / np.maximum(c, 1)
