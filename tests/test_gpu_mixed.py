"""Accuracy of the mixed-precision fused cycle (PENNANT_MODE_MIXED).

The mixed form keeps positions, volumes, the zone EOS, energy, work and dt in
FP64 and computes side forces, corner forces and masses and the whole QCS
chain in FP32.  It is not bit-exact; these tests bound its error:

  * against the CPU oracle on the reference's decks (Sedov 45^2, Noh pie,
    LeBlanc; 10 cycles each), on rect and pie decks that set the parameters
    no reference deck sets, and against the reference-held golden
    test/sedov/sedov.xy;
  * against the plain FP64 fused form -- which the parity tests prove
    bit-identical to the oracle -- on runs too long for the oracle: Sedov
    90^2 to t = 1, LeBlanc 256x16 to t = 6, Noh pie 256^2 for 3000 cycles,
    Sedov 1440^2 for 50 cycles;
  * launch shapes and repeated runs give bit-identical mixed results (the
    mixed form is deterministic too).

Error norms (pennant_amd.accuracy): rel_l1 = sum|got-want| / sum|want| and
rel_linf = max|got-want| / max|want| over all zones.  The bound asked of the
mixed form is rel_l1(zr) <= 1e-2; the tolerances below are much tighter,
set about 10x above the errors measured on MI355X.
"""
import sys

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from pennant_amd.accuracy import rel_l1, rel_linf
from oracle_lib import Oracle
from golden_decks import HERE, PARAMS, override, ref_deck

pytestmark = pytest.mark.gpu

ZF = ("zr", "ze", "zp")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_mode(text, mode, ncyc=10 ** 7):
    with pa.Problem(text) as pb:
        n = pb.run(ncyc, mode)
        return n, pb.scalars(), {f: pb.get(f) for f in ZF + ("px",)}


# (deck, cycles, tolerance on rel_l1 of zr/ze/zp, on their rel_linf, on the
# relative error of the simulated time).  The Noh pie's dt comes from the
# zones at the origin, whose state carries the largest error (1e-5).
ORACLE_DECKS = {
    "sedov45": (lambda: ref_deck("sedov"), 10, 1e-7, 1e-6, 1e-6),
    "noh_pie30x100": (lambda: ref_deck("noh"), 10, 2e-4, 5e-4, 5e-5),
    "leblanc10x90": (lambda: ref_deck("leblanc"), 10, 1e-7, 1e-6, 1e-6),
    # parameters no reference deck sets (see golden_decks.PARAMS)
    "sedov40_params": (lambda: override(wd.sedov_rect(40, cstop=40), **PARAMS), 40, 2e-7, 2e-7, 1e-7),
    "noh_pie20x16_params": (lambda: override(wd.noh_pie(20, 16, cstop=25, dtinit=1e-4),
                                             **dict(PARAMS, q1=0.5, q2=1.2, qgamma=1.2)), 25, 1e-5, 2e-5, 1e-6),
}


@pytest.mark.parametrize("deck", list(ORACLE_DECKS))
def test_mixed_against_oracle(deck):
    make, ncyc, tol1, tolinf, toltime = ORACLE_DECKS[deck]
    text = make()
    say(f"[mixed vs oracle] {deck}")
    o = Oracle(text)
    assert o.run(1000) == ncyc
    n, sc, got = run_mode(text, pa.MODE_MIXED)
    assert n == ncyc and sc["done"] == 1
    for f in ZF:
        e1, ei = rel_l1(got[f], o.get(f)), rel_linf(got[f], o.get(f))
        say(f"  {f}: rel_l1 {e1:.2e} rel_linf {ei:.2e}")
        assert e1 <= tol1 and ei <= tolinf, (f, e1, ei)
    # the simulated time (the whole dt chain)
    terr = abs(sc["time"] - o.scalars()["time"]) / o.scalars()["time"]
    say(f"  time: rel err {terr:.2e}")
    assert terr <= toltime
    # positions: to 1e-7 of the domain size
    assert np.max(np.abs(got["px"] - o.get("px"))) <= 1e-7 * np.max(np.abs(o.get("px")))


def test_mixed_against_sedov_golden():
    """Mixed form vs the reference's own sedov.xy (9 significant digits) in the
    blast region (zone rows 0-14; the golden's far field is excluded the same
    way as for the oracle): every significant density value within 1e-6."""
    n, sc, got = run_mode(ref_deck("sedov"), pa.MODE_MIXED)
    assert n == 10
    vals = []
    for line in open(HERE / "golden" / "sedov" / "sedov.xy"):
        if not line.startswith("#"):
            vals.append(float(line.split()[1]))
    nz = 45 * 45
    gold_zr = np.array(vals[:nz])
    rows = np.arange(nz) // 45
    sel = (rows < 15) & (np.abs(gold_zr) >= 1e-30)
    rel = np.abs(got["zr"][sel] - gold_zr[sel]) / np.abs(gold_zr[sel])
    assert sel.sum() >= 600
    assert rel.max() <= 1e-6, rel.max()


# (deck, tolerance rel_l1 for zr, ze, zp)
LONG_DECKS = {
    "sedov90_t1": (lambda: wd.sedov_90(), 2e-4),
    "leblanc256x16_t6": (lambda: wd.leblanc_rect(256, 16), 2e-6),
    "noh_pie256_c3000": (lambda: wd.noh_pie(256, 256, dtinit=1e-5, cstop=3000), 2e-4),
    "sedov360_c2000": (lambda: wd.sedov_rect(360, cstop=2000), 2e-3),
}


@pytest.mark.parametrize("deck", list(LONG_DECKS))
def test_mixed_long_runs_against_plain_fp64(deck):
    make, tol = LONG_DECKS[deck]
    text = make()
    say(f"[mixed vs fp64, long run] {deck}")
    n64, s64, ref = run_mode(text, pa.MODE_FUSED)
    nmx, smx, got = run_mode(text, pa.MODE_MIXED)
    assert nmx == n64 and smx["done"] == 1
    for f in ZF:
        e1 = rel_l1(got[f], ref[f])
        say(f"  {f}: rel_l1 {e1:.2e} rel_linf {rel_linf(got[f], ref[f]):.2e}  ({n64} cycles)")
        assert e1 <= tol, (f, e1)
    assert rel_l1(got["zr"], ref["zr"]) <= 1e-2


def test_mixed_full_size_sedov1440():
    text = wd.sedov_1440(cstop=50)
    say("[mixed vs fp64] sedov 1440^2, 50 cycles")
    n64, s64, ref = run_mode(text, pa.MODE_FUSED)
    nmx, smx, got = run_mode(text, pa.MODE_MIXED)
    assert n64 == nmx == 50
    for f in ZF:
        e1 = rel_l1(got[f], ref[f])
        say(f"  {f}: rel_l1 {e1:.2e}")
        assert e1 <= 1e-6, (f, e1)
    # mass is the FP64 zm in both forms; the mixed form conserves it exactly
    assert abs(smx["time"] - s64["time"]) <= 1e-6 * s64["time"]


@pytest.mark.parametrize("maxb,xcd", [(1, 1), (7, 0), (4096, 1)])
def test_mixed_deterministic_across_launch_shapes(maxb, xcd):
    for text in (wd.noh_pie(40, 24, cstop=30, dtinit=1e-4), wd.sedov_rect(37, cstop=30)):
        _, _, base = run_mode(text, pa.MODE_MIXED)
        with pa.Problem(text) as pb:
            pb.set_tuning("max_blocks", maxb)
            pb.set_tuning("xcd", xcd)
            pb.run(100, pa.MODE_MIXED)
            for f in ZF + ("px",):
                assert np.array_equal(pb.get(f).view(np.uint64), base[f].view(np.uint64)), f


def test_mixed_and_plain_interleave_on_one_problem():
    """Switching forms between cycles on one problem: the FP32 buffers are
    rebuilt each cycle from the FP64 state, so a plain cycle after mixed
    cycles starts from that state exactly."""
    text = wd.sedov_rect(30, cstop=40)
    with pa.Problem(text) as a, pa.Problem(text) as b:
        a.run(20, pa.MODE_MIXED)
        b.run(20, pa.MODE_MIXED)
        a.run(20, pa.MODE_FUSED)
        b.run(20, pa.MODE_TASKS)
        for f in ZF + ("px", "pu"):
            assert np.array_equal(a.get(f).view(np.uint64), b.get(f).view(np.uint64)), f


@pytest.mark.parametrize("mode", [pa.MODE_MIXED, pa.MODE_LEARNED])
def test_mixed_zr_ze_are_the_corrector_quotients(mode):
    """The mixed corrector does not store zr and ze; its predictors derive
    them from zm, zvol and zetot, and reading the fields materialises them
    with the FP64 corrector's quotients, bit for bit."""
    from pennant_amd.closure import Mlp
    from pennant_amd.train_closure import DEFAULT_OUT
    m = Mlp.load(DEFAULT_OUT) if mode == pa.MODE_LEARNED else None
    for text in (wd.sedov_rect(40, cstop=60), wd.noh_pie(30, 20, cstop=50, dtinit=1e-4)):
        with pa.Problem(text) as pb:
            if m is not None:
                pb.set_closure(m)
            pb.run(37, mode)
            zm, zvol, zetot = pb.get("zm"), pb.get("zvol"), pb.get("zetot")
            assert np.array_equal(pb.get("zr").view(np.uint64), (zm / zvol).view(np.uint64))
            assert np.array_equal(pb.get("ze").view(np.uint64),
                                  (zetot / (zm + 1e-99)).view(np.uint64))
    if m is not None:
        m.close()


def test_written_zr_drives_the_next_mixed_predictor():
    """A density written between mixed cycles is what the next predictor's
    EOS reads (calcStateHalf reads ZR, src/PolyGas.cc:63-108): writing back
    the materialised values changes nothing, a changed density changes the
    next cycle's pressure in those zones."""
    text = wd.sedov_rect(32, cstop=200)
    with pa.Problem(text) as a, pa.Problem(text) as b, pa.Problem(text) as c:
        for pb in (a, b, c):
            pb.run(25, pa.MODE_MIXED)
        zr = a.get("zr")
        b.set("zr", zr)
        zr2 = zr.copy()
        zr2[::5] *= 1.0 + 1e-3
        c.set("zr", zr2)
        for pb in (a, b, c):
            pb.run(1, pa.MODE_MIXED)
        for f in ZF + ("px", "pu", "zetot", "zvol"):
            assert np.array_equal(a.get(f).view(np.uint64), b.get(f).view(np.uint64)), f
        zp_a, zp_c = a.get("zp")[::5], c.get("zp")[::5]
        hot = zp_a != 0.0          # zero internal energy gives zero pressure at any density
        assert hot.sum() > 10
        assert np.all(zp_a[hot] != zp_c[hot])


def test_written_smf_drives_the_fused_cycles():
    """The fused cycles read FP32 copies of smf and per-zone codes that mark
    uniform side fractions; a write to smf (set_field, or a store through its
    device pointer) rebuilds both.  Perturbed fractions on a few zones: the
    mixed form then follows the FP64 form on the perturbed mesh (to its usual
    accuracy) and departs from the unperturbed run."""
    text = wd.sedov_rect(48, cstop=40)

    def run(mode, perturb):
        with pa.Problem(text) as pb:
            if perturb:
                smf = pb.get("smf")
                z = np.arange(0, smf.size // 4, 7)          # every 7th zone
                s = 4 * z
                smf[s] += 0.01                               # side 0 of those zones
                smf[s + 2] -= 0.01                           # side 2
                pb.set("smf", smf)
            pb.run(40, mode)
            return {f: pb.get(f) for f in ZF}

    mix_p, fp_p, mix_0 = run(pa.MODE_MIXED, True), run(pa.MODE_FUSED, True), run(pa.MODE_MIXED, False)
    for f in ZF:
        e = rel_l1(mix_p[f], fp_p[f])
        d = rel_l1(mix_p[f], mix_0[f])
        say(f"[smf write] {f}: mixed vs FP64 (perturbed) {e:.2e}, perturbed vs not {d:.2e}")
        assert e < 1e-3, f
        assert d > 10 * e, f


def test_smf_written_through_device_pointer_matches_set_field():
    """A store through the device pointer handed out for smf reaches the
    mixed form's FP32 copy before the next cycle, as set_field does at once:
    both runs give the same bits."""
    import hip_rt
    text = wd.sedov_rect(40, cstop=30)

    def run(via_ptr):
        with pa.Problem(text) as pb:
            smf = np.ascontiguousarray(pb.get("smf"))
            s = 4 * np.arange(0, smf.size // 4, 5)
            smf[s] += 0.01
            smf[s + 2] -= 0.01
            if via_ptr:
                p = pb.device_ptr("smf")
                assert hip_rt.hip().hipMemcpy(p, smf.ctypes.data, smf.nbytes, hip_rt.H2D) == 0
            else:
                pb.set("smf", smf)
            pb.run(30, pa.MODE_MIXED)
            return {f: pb.get(f) for f in ZF}

    a, b = run(True), run(False)
    for f in ZF:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f
