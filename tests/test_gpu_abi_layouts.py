"""Every task entry point of the C ABI, called the way a Legion task wrapper
calls it: on device buffers the caller allocated (hipMalloc), with a mesh
descriptor whose maps the caller built and uploaded itself, and with vector
fields in each layout a caller may hold:

  * "soa"          two arrays per vector field (stride 1);
  * "interleaved"  one double2 array per field -- the reference's field
                   instance, AccessorRO<double2>::ptr(rect) -- passed as
                   x = ptr, y = ptr + 1, stride 2;
  * "stride3"      records of three doubles {x, y, pad}, stride 3 (the generic
                   strided path); the pad must come back untouched.

For each of the 30 doCycle steps of a cycle the oracle's state before the step
is loaded into the caller's buffers, the step's entry point(s) run, and every
field the step writes is compared bit for bit with the oracle after the step.
calcSideFracs (init only) and calcGlobalDt are checked on their own.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import pennant_amd as pa
from pennant_amd import decks as wd
from oracle_lib import FIELDS, Oracle, calc_global_dt
from golden_decks import PARAMS, override, ref_deck

pytestmark = pytest.mark.gpu

VECTOR = {"px", "pxp", "px0", "pu", "pu0", "pf", "pap", "zx", "zxp", "zuc", "ex", "exp", "ssurfp",
          "sfp", "sfq", "sft", "cqe1", "cqe2"}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def deck_params(text):
    """The deck's physics and dt parameters with the reference's defaults
    (src/Driver.cc:78-87, src/Hydro.cc:206-214, src/PolyGas.cc:58-59,
    src/TTS.cc:46-47, src/QCS.cc:84-86)."""
    kv = {ln.split()[0]: ln.split()[1:] for ln in text.splitlines() if ln.split()}

    def g(k, d):
        return float(kv[k][0]) if k in kv else d
    return {"gamma": g("gamma", 5 / 3), "ssmin": g("ssmin", 0.0), "alfa": g("alfa", 0.5),
            "qgamma": g("qgamma", 5 / 3), "q1": g("q1", 0.0), "q2": g("q2", 2.0), "cfl": g("cfl", 0.6),
            "cflv": g("cflv", 0.1), "dtmax": g("dtmax", 1e99), "dtinit": g("dtinit", 1e99),
            "dtfac": g("dtfac", 1.2), "tstop": g("tstop", 1e99)}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bits(name, got, want):
    g, w = bits(got).reshape(-1), bits(want).reshape(-1)
    if g.shape != w.shape or not np.array_equal(g, w):
        bad = np.nonzero(g != w)[0]
        raise AssertionError(f"{name}: {bad.size} of {g.size} values differ, first at {bad[:4]}")


class CallerField:
    """A field in caller-owned device memory, in one layout."""

    PAD = np.float64(-7.25e300)

    def __init__(self, values, layout, vector):
        from hip_rt import DeviceArray
        self.vector, self.layout = vector, layout
        v = np.asarray(values, dtype=np.float64)
        if not vector:
            self.arrs = [DeviceArray(v.reshape(-1).copy())]
        elif layout == "soa":
            self.arrs = [DeviceArray(v[:, 0].copy()), DeviceArray(v[:, 1].copy())]
        elif layout == "interleaved":
            self.arrs = [DeviceArray(np.ascontiguousarray(v.reshape(-1, 2)))]
        else:
            rec = np.full((v.shape[0], 3), self.PAD)
            rec[:, :2] = v
            self.arrs = [DeviceArray(rec)]

    def arg(self):
        if not self.vector:
            return self.arrs[0].ptr
        if self.layout == "soa":
            return pa.vec(self.arrs[0].ptr, self.arrs[1].ptr, 1)
        if self.layout == "interleaved":
            return pa.vec_interleaved(self.arrs[0].ptr)
        return pa.vec(self.arrs[0].ptr, self.arrs[0].ptr + 8, 3)

    def value(self):
        if not self.vector:
            return self.arrs[0].numpy()
        if self.layout == "soa":
            return np.stack([self.arrs[0].numpy(), self.arrs[1].numpy()], axis=1)
        if self.layout == "interleaved":
            return self.arrs[0].numpy().reshape(-1, 2)
        rec = self.arrs[0].numpy()
        assert np.all(bits(rec[:, 2]) == bits(np.full(rec.shape[0], self.PAD))), "pad overwritten"
        return rec[:, :2]


def caller_mesh(o):
    """The mesh descriptor built by the caller from the oracle's maps: int32
    maps, zone offsets, and the point-corner CSR (sides grouped by their first
    point, ascending)."""
    from hip_rt import DeviceArray
    maps = {n: o.map(n).astype(np.int32) for n in ("mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4",
                                                    "znump")}
    znump = maps["znump"]
    maps["zsbeg"] = np.concatenate([[0], np.cumsum(znump)]).astype(np.int32)
    p1 = maps["mapsp1"]
    order = np.argsort(p1, kind="stable").astype(np.int32)
    maps["pcside"] = order
    maps["pcbeg"] = np.concatenate([[0], np.cumsum(np.bincount(p1, minlength=o.nump))]).astype(np.int32)
    dev = {k: DeviceArray(v) for k, v in maps.items()}
    m = pa.Mesh()
    m.nump, m.numz, m.nums = o.nump, o.numz, o.nums
    m.maxznump = int(znump.max())
    for k in ("mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4", "znump", "zsbeg", "pcbeg", "pcside"):
        setattr(m, k, dev[k].ptr)
    return m, dev


# step -> [(entry point, argument spec)], in the order of the task_step
# driver (src/Hydro.cc:342-983).  Argument kinds: "m" mesh; ("in", f) read,
# ("out", f) written, ("io", f) read and written; "dt" the cycle's dt on the
# device; ("d", key) a deck parameter; "ws" a reduction workspace;
# ("iout", key) a device int, ("sout", key) / ("sin", key) a device double
# (compared with / taken from the oracle's scalars).
STEPS = {
    1: [("pennant_hip_adv_pos_half", ["m", ("in", "px0"), ("in", "pu0"), ("out", "pxp"), "dt"])],
    2: [("pennant_hip_calc_ctrs", ["m", ("in", "pxp"), ("out", "exp"), ("out", "zxp")])],
    3: [("pennant_hip_calc_vols", ["m", ("in", "pxp"), ("in", "zxp"), ("out", "sareap"), ("out", "svolp"),
                                    ("out", "zareap"), ("out", "zvolp"), ("iout", "numsbad_pred"), "ws"])],
    4: [("pennant_hip_calc_surf_vecs", ["m", ("in", "exp"), ("in", "zxp"), ("out", "ssurfp")])],
    5: [("pennant_hip_calc_edge_len", ["m", ("in", "pxp"), ("out", "elen")])],
    6: [("pennant_hip_calc_char_len", ["m", ("in", "sareap"), ("in", "elen"), ("out", "zdl")])],
    7: [("pennant_hip_calc_rho", ["m", ("in", "zm"), ("in", "zvolp"), ("out", "zrp")])],
    8: [("pennant_hip_calc_crnr_mass", ["m", ("in", "smf"), ("in", "zrp"), ("in", "zareap"),
                                         ("io", "pmaswt")])],
    9: [("pennant_hip_calc_state_half", ["m", ("in", "zr"), ("in", "zvolp"), ("in", "zvol0"), ("in", "ze"),
                                          ("in", "zwrate"), ("in", "zm"), ("out", "zp"), ("out", "zss"),
                                          ("d", "gamma"), ("d", "ssmin"), "dt"])],
    10: [("pennant_hip_calc_force_pgas", ["m", ("in", "ssurfp"), ("in", "zp"), ("out", "sfp")])],
    11: [("pennant_hip_calc_force_tts", ["m", ("in", "sareap"), ("in", "smf"), ("in", "ssurfp"),
                                         ("in", "zareap"), ("in", "zrp"), ("in", "zss"), ("out", "sft"),
                                         ("d", "alfa"), ("d", "ssmin")])],
    12: [("pennant_hip_set_corner_div", ["m", ("in", "exp"), ("in", "elen"), ("in", "zxp"), ("in", "pxp"),
                                         ("in", "pu0"), ("out", "zuc"), ("out", "carea"), ("out", "ccos"),
                                         ("out", "cdiv"), ("out", "cevol"), ("out", "cdu")])],
    13: [("pennant_hip_set_qcn_force", ["m", ("in", "elen"), ("in", "cdiv"), ("in", "cdu"), ("in", "cevol"),
                                        ("in", "zrp"), ("in", "zss"), ("in", "pu0"), ("out", "crmu"),
                                        ("out", "cqe1"), ("out", "cqe2"), ("d", "qgamma"), ("d", "q1"),
                                        ("d", "q2")])],
    14: [("pennant_hip_set_force_qcs", ["m", ("in", "carea"), ("in", "cqe1"), ("in", "cqe2"), ("in", "elen"),
                                        ("io", "ccos"), ("out", "cw"), ("out", "sfq")])],
    15: [("pennant_hip_set_vel_diff", ["m", ("in", "elen"), ("in", "zss"), ("in", "pxp"), ("in", "pu0"),
                                       ("out", "ztmp"), ("out", "zdu"), ("d", "q1"), ("d", "q2")])],
    16: [("pennant_hip_sum_crnr_force", ["m", ("in", "sfp"), ("in", "sfq"), ("in", "sft"), ("io", "pf")])],
    19: [("pennant_hip_calc_accel", ["m", ("in", "pf"), ("in", "pmaswt"), ("out", "pap")]),
         ("pennant_hip_adv_pos_full", ["m", ("in", "px0"), ("in", "pu0"), ("in", "pap"), ("out", "px"),
                                       ("out", "pu"), "dt"])],
    20: [("pennant_hip_calc_ctrs", ["m", ("in", "px"), ("out", "ex"), ("out", "zx")])],
    21: [("pennant_hip_calc_vols", ["m", ("in", "px"), ("in", "zx"), ("out", "sarea"), ("out", "svol"),
                                     ("out", "zarea"), ("out", "zvol"), ("iout", "numsbad_corr"), "ws"])],
    22: [("pennant_hip_calc_work", ["m", ("in", "sfp"), ("in", "sfq"), ("in", "pu"), ("in", "pu0"),
                                    ("in", "pxp"), ("out", "zw"), ("io", "zetot"), "dt"])],
    23: [("pennant_hip_calc_work_rate", ["m", ("in", "zvol0"), ("in", "zvol"), ("in", "zw"), ("in", "zp"),
                                         ("out", "zwrate"), "dt"])],
    24: [("pennant_hip_calc_energy", ["m", ("in", "zetot"), ("in", "zm"), ("out", "ze")])],
    25: [("pennant_hip_calc_rho", ["m", ("in", "zm"), ("in", "zvol"), ("out", "zr")])],
    26: [("pennant_hip_calc_dt_new", ["m", ("in", "zdl"), ("in", "zdu"), ("in", "zss"), ("d", "cfl"),
                                      ("sout", "dtnew"), "ws"])],
    27: [("pennant_hip_calc_dvol", ["m", ("in", "zvol"), ("in", "zvol0"), ("sout", "dvovmax"), "ws"])],
    28: [("pennant_hip_calc_dt", [("sin", "dt"), ("sin", "dtnew"), ("sin", "dvovmax"), ("d", "cflv"),
                                  ("sout", "dtrec")])],
}

LAYOUT_DECKS = {
    "sedov32_params": lambda: override(wd.sedov_rect(32, cstop=10 ** 6), **PARAMS),
    "noh_pie12x40": lambda: wd.noh_pie(12, 40, cstop=10 ** 6, dtinit=1e-4),   # 40-corner origin
    "leblanc10x90": lambda: ref_deck("leblanc"),
}


@pytest.mark.parametrize("layout", ["soa", "interleaved", "stride3"])
@pytest.mark.parametrize("deck", list(LAYOUT_DECKS))
def test_every_task_entry_point_on_caller_buffers(deck, layout):
    from hip_rt import DeviceArray, synchronize
    L = pa.lib()
    text = LAYOUT_DECKS[deck]()
    prm = deck_params(text)
    say(f"[abi] {deck} layout {layout}")
    o = Oracle(text)
    o.run(3)
    dt = o.begin_cycle()
    mesh, _maps = caller_mesh(o)
    ws = DeviceArray.zeros(L.pennant_hip_reduce_workspace_bytes() // 8)
    dtbuf = DeviceArray(np.array([dt]))
    bcs = o.bcs()
    called = set()
    for k in range(pa.NSTEPS):
        pre = {f: o.get(f) for f in FIELDS}
        pre_sc = o.scalars()
        assert o.step(k) == 0
        post = {f: o.get(f) for f in FIELDS}
        post_sc = o.scalars()
        if k == 17:
            # applyFixedBC: one call per BC list, in order, on the same buffers
            pf = CallerField(pre["pf"], layout, True)
            pu0 = CallerField(pre["pu0"], layout, True)
            lists = [DeviceArray(pts.astype(np.int32)) for pts, _ in bcs]
            synchronize()
            for (pts, (vx, vy)), dl in zip(bcs, lists):
                assert L.pennant_hip_apply_fixed_bc(dl.ptr, int(pts.size), vx, vy, pf.arg(), pu0.arg(),
                                                    None) == 0
            synchronize()
            assert_bits(f"step 17 pf", pf.value(), post["pf"])
            assert_bits(f"step 17 pu0", pu0.value(), post["pu0"])
            called.add("pennant_hip_apply_fixed_bc")
            continue
        for name, spec in STEPS.get(k, []):
            args, outs, keep = [], [], []
            for a in spec:
                if a == "m":
                    args.append(ctypes.byref(mesh))
                elif a == "dt":
                    args.append(dtbuf.ptr)
                elif a == "ws":
                    args.append(ws.ptr)
                elif a[0] == "d":
                    args.append(prm[a[1]])
                elif a[0] in ("in", "io", "out"):
                    f = a[1]
                    init = post[f] if a[0] == "in" else pre[f]
                    if a[0] == "out":   # outputs start as garbage the kernel must overwrite
                        init = np.full_like(pre[f], np.nan)
                    cf = CallerField(init, layout, f in VECTOR)
                    keep.append(cf)
                    args.append(cf.arg())
                    if a[0] != "in":
                        outs.append((f, cf))
                elif a[0] == "iout":
                    d = DeviceArray(np.array([12345], dtype=np.int32))
                    keep.append(d)
                    args.append(d.ptr)
                    outs.append((("i", a[1]), d))
                elif a[0] == "sout":
                    d = DeviceArray(np.array([np.nan]))
                    keep.append(d)
                    args.append(d.ptr)
                    outs.append((("s", a[1]), d))
                elif a[0] == "sin":
                    d = DeviceArray(np.array([dt if a[1] == "dt" else post_sc[a[1]]]))
                    keep.append(d)
                    args.append(d.ptr)
            synchronize()
            rc = getattr(L, name)(*args, None)
            assert rc == 0, (k, name, rc)
            synchronize()
            for key, cf in outs:
                if isinstance(key, tuple):
                    got = cf.numpy()[0]
                    want = post_sc[key[1]]
                    if key[0] == "i":
                        assert int(got) == int(want), (k, name, key)
                    else:
                        assert_bits(f"step {k} {name} {key[1]}", got, want)
                else:
                    assert_bits(f"step {k} {name} {key}", cf.value(), post[key])
            called.add(name)
    o.end_cycle()
    # calcSideFracs on the initial geometry; calcGlobalDt for the next cycle
    o0 = Oracle(text)
    m0, _m0 = caller_mesh(o0)
    sa, za = DeviceArray(o0.get("sarea")), DeviceArray(o0.get("zarea"))
    smf = DeviceArray(np.full(o0.nums, np.nan))
    synchronize()
    assert L.pennant_hip_calc_side_fracs(ctypes.byref(m0), sa.ptr, za.ptr, smf.ptr, None) == 0
    synchronize()
    assert_bits("smf", smf.numpy(), o0.get("smf"))
    called.add("pennant_hip_calc_side_fracs")
    sc = o.scalars()
    dl, dr, tm, out = (DeviceArray(np.array([v])) for v in (sc["dt"], sc["dtrec"], sc["time"], np.nan))
    synchronize()
    assert L.pennant_hip_calc_global_dt(dl.ptr, dr.ptr, tm.ptr, int(sc["cycle"]), prm["dtmax"],
                                        prm["dtinit"], prm["dtfac"], prm["tstop"], out.ptr, None) == 0
    synchronize()
    want = calc_global_dt(prm["dtmax"], prm["dtinit"], prm["dtfac"], prm["tstop"], int(sc["cycle"]),
                          sc["time"], sc["dt"], sc["dtrec"])
    assert_bits("global dt", out.numpy()[0], want)
    called.add("pennant_hip_calc_global_dt")
    assert len(called) == 27, sorted(called)


def test_entry_points_reject_bad_descriptors():
    """Negative strides and untileable meshes come back as PENNANT_EINVAL,
    before any launch."""
    from hip_rt import DeviceArray
    L = pa.lib()
    o = Oracle(wd.sedov_rect(4, cstop=1))
    mesh, _maps = caller_mesh(o)
    px = DeviceArray(o.get("px"))
    ex = DeviceArray.zeros((o.nums, 2))
    zx = DeviceArray.zeros((o.numz, 2))
    bad = pa.vec(px.ptr, px.ptr + 8, -2)
    assert L.pennant_hip_calc_ctrs(ctypes.byref(mesh), bad, pa.vec_interleaved(ex.ptr),
                                   pa.vec_interleaved(zx.ptr), None) == pa.EINVAL
    m2 = pa.Mesh.from_buffer_copy(mesh)
    m2.maxznump = 9
    assert L.pennant_hip_calc_ctrs(ctypes.byref(m2), pa.vec_interleaved(px.ptr), pa.vec_interleaved(ex.ptr),
                                   pa.vec_interleaved(zx.ptr), None) == pa.EINVAL
