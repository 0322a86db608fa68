"""The partition of a mesh into pieces (pennant_pieces_plan, host only), against
an independent numpy brute force of the reference's rules:

  * piece grid: numpcx * numpcy = numpcs, blocks as close to square as the
    two divisors of numpcs nearest to sqrt(numpcs * short / long) allow
    (src/Mesh.hh:760-791);
  * zone (i, j) belongs to block (i // zpx, j // zpy) with zpx = ceil(nzx /
    numpcx), zpy = ceil(nzy / numpcy) (src/GenMesh.cc:898-975);
  * point (i, j) belongs to the block of zone (min(i, nzx-1), min(j, nzy-1));
    the pie origin to piece 0 (src/GenMesh.cc:790-873);
  * a piece's private points are its owned points no other piece reaches,
    master points its owned points another piece reaches, ghost points the
    points it reaches and does not own (src/Mesh.cc:432-617); a master point
    receives every corner another piece has at it.
"""
import math

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd


def piece_grid(numpcs, nzx, nzy):
    nx, ny = float(nzx), float(nzy)
    swapped = nx > ny
    if swapped:
        nx, ny = ny, nx
    ideal = math.sqrt(numpcs * nx / ny)
    lo = max(1, math.floor(ideal + 1e-12))
    while numpcs % lo:
        lo -= 1
    hi = math.ceil(ideal - 1e-12)
    while numpcs % hi:
        hi += 1
    side = lambda k: max(nx / k, ny / (numpcs // k))
    px = lo if side(lo) <= side(hi) else hi
    py = numpcs // px
    return (py, px) if swapped else (px, py)


def brute_force(text, numpcs):
    kv = {ln.split()[0]: ln.split()[1:] for ln in text.splitlines() if ln.split()}
    pie = kv["meshtype"][0] == "pie"
    nzx, nzy = int(kv["meshparams"][0]), int(kv["meshparams"][1])
    h = pa.HostMesh(text)
    mapsp1, mapsz = h.get("mapsp1"), h.get("mapsz")
    numz, nump = nzx * nzy, h.get("px").size
    npcx, npcy = piece_grid(numpcs, nzx, nzy)
    zpx, zpy = -(-nzx // npcx), -(-nzy // npcy)
    z = np.arange(numz)
    zpiece = (z // nzx) // zpy * npcx + (z % nzx) // zpx
    p = np.arange(nump)
    if pie:
        i, j = (p - 1) % (nzx + 1), (p - 1) // (nzx + 1) + 1
    else:
        i, j = p % (nzx + 1), p // (nzx + 1)
    own = np.minimum(j, nzy - 1) // zpy * npcx + np.minimum(i, nzx - 1) // zpx
    if pie:
        own[0] = 0
    spiece = zpiece[mapsz]
    reached_by_other = np.zeros(nump, bool)
    reached_by_other[mapsp1[spiece != own[mapsp1]]] = True
    out = []
    for c in range(npcx * npcy):
        sides = np.nonzero(spiece == c)[0]
        if sides.size == 0:
            continue
        reach = np.zeros(nump, bool)
        reach[mapsp1[sides]] = True
        owned = reach & (own == c)
        master = owned & reached_by_other
        ghost = reach & (own != c)
        # corners other pieces have at this piece's master points
        recv = int(np.count_nonzero(master[mapsp1] & (spiece != c)))
        out.append({"piece": c, "zones": int(np.count_nonzero(zpiece == c)), "sides": int(sides.size),
                    "private": int(np.count_nonzero(owned & ~master)), "master": int(np.count_nonzero(master)),
                    "ghost": int(np.count_nonzero(ghost)), "received": recv})
    return npcx, npcy, out


CASES = [
    (wd.sedov_rect(30, bc_far=True), (1, 2, 3, 4, 6, 8, 9, 12, 16)),
    (wd.sedov_rect(10), (7, 5)),                  # block rounding leaves pieces empty
    (wd.leblanc_rect(64, 8), (4, 8, 6)),
    (wd.noh_pie(20, 24), (2, 4, 5, 8, 12)),
    (wd.noh_pie(8, 96), (4, 6)),
    (wd.noh_pie(33, 7), (3, 10)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plan_matches_brute_force(case):
    text, counts = CASES[case]
    for n in counts:
        got = pa.pieces_plan(text, n)
        npcx, npcy, want = brute_force(text, n)
        assert (got["numpcx"], got["numpcy"]) == (npcx, npcy), n
        assert npcx * npcy == n
        assert got["pieces"] == want, n


def test_plan_invariants():
    text = wd.noh_pie(20, 24)
    h = pa.HostMesh(text)
    nump, nums = h.get("px").size, h.get("mapsp1").size
    for n in (1, 2, 4, 8, 12):
        pl = pa.pieces_plan(text, n)["pieces"]
        assert sum(d["sides"] for d in pl) == nums
        assert sum(d["private"] + d["master"] for d in pl) == nump
        if n == 1:
            assert pl[0]["master"] == pl[0]["ghost"] == pl[0]["received"] == 0
        else:
            # every ghost point is some other piece's master point
            assert sum(d["ghost"] for d in pl) >= sum(d["master"] for d in pl) > 0


def test_plan_square_counts_and_errors():
    assert (pa.pieces_plan(wd.sedov_rect(40), 16)["numpcx"], pa.pieces_plan(wd.sedov_rect(40), 16)["numpcy"]) == (4, 4)
    assert pa.pieces_plan(wd.leblanc_rect(64, 8), 8)["numpcx"] == 8      # long along x
    with pytest.raises(pa.PennantError):
        pa.pieces_plan(wd.sedov_rect(10), 0)
    with pytest.raises(pa.PennantError):
        pa.pieces_plan("meshtype hex\nmeshparams 4 4 1.0 1.0\n", 2)
