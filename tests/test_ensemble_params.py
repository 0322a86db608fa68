"""The batched ensemble's problem parameters (pennant_ensemble_params):
std::mt19937_64(seed), for each k: n = round(exp(U(ln 32, ln 512))), then
the Sedov energy U[0.1, 1.0] (k even) or the Noh speed U[0.5, 2.0] (k odd).

Checked against an independent Python restatement of mt19937_64 (pinned by
the C++ standard's check value: the 10000th output of the default-seeded
engine is 9981545732273789042) and of uniform_real_distribution<double> on a
64-bit engine (one draw, x / 2^64, scaled to [a, b)).
"""
import math

import numpy as np

import pennant_amd as pa
from pennant_amd import decks as wd

M64 = (1 << 64) - 1


class MT19937_64:
    N, M = 312, 156

    def __init__(self, seed):
        self.mt = [0] * self.N
        self.mt[0] = seed & M64
        for i in range(1, self.N):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & M64
        self.i = self.N

    def next(self):
        if self.i >= self.N:
            up, lo = 0xFFFFFFFF80000000, 0x7FFFFFFF
            for k in range(self.N):
                x = (self.mt[k] & up) | (self.mt[(k + 1) % self.N] & lo)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[k] = self.mt[(k + self.M) % self.N] ^ xa
            self.i = 0
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & M64

    def uniform(self, a, b):
        u = self.next() / 18446744073709551616.0
        if u >= 1.0:
            u = np.nextafter(1.0, 0.0)
        return a + (b - a) * u


def test_mt19937_64_check_value():
    g = MT19937_64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


def test_ensemble_params_match_restatement():
    count, seed = 256, 20261202
    n = np.zeros(count, dtype=np.int32)
    par = np.zeros(count)
    assert pa.lib().pennant_ensemble_params(seed, count, n.ctypes.data, par.ctypes.data) == 0
    g = MT19937_64(seed)
    for k in range(count):
        x = math.exp(g.uniform(math.log(32.0), math.log(512.0)))
        assert n[k] == int(math.floor(x + 0.5)), k
        p = g.uniform(0.1, 1.0) if k % 2 == 0 else g.uniform(0.5, 2.0)
        assert par[k] == p, k
    assert n.min() >= 32 and n.max() <= 512
    assert np.all(par[0::2] >= 0.1) and np.all(par[0::2] < 1.0)
    assert np.all(par[1::2] >= 0.5) and np.all(par[1::2] < 2.0)


def test_ensemble_decks():
    decks, n, par = wd.ensemble(8)
    assert len(decks) == 8
    for k, text in enumerate(decks):
        h = pa.HostMesh(text)
        nz = h.get("znump").size
        assert nz == n[k] * n[k]
        assert ("meshtype rect" in text) == (k % 2 == 0)
    # the whole ensemble is about 12-14 million zones
    _, n256, _ = wd.ensemble(256)
    tot = int((n256.astype(np.int64) ** 2).sum())
    assert 10_000_000 < tot < 16_000_000
