"""PENNANT hydro cycle benchmark on one MI355X.

Workload (default): Sedov blast on a 1440 x 1440 rect mesh (2,073,600 zones,
8,294,400 sides), fused cycle in mixed precision with the learned QCS
closure (trained weights loaded from the package; training is not timed),
reported against the plain FP64 fused cycle (bit-identical to the CPU
oracle) and the mixed form with the exact QCS on the same workload.  A step
is one hydro cycle over the whole mesh; the state is resident in HBM before
the timed region starts.

Each form is first advanced, untimed, to mid-problem (--mid-time, t = 0.5 of
the run to t = 1), so the timed cycles see the developed shock front that
the learned closure works on; after timing every form runs on to tstop, and
the forms are compared there, at equal simulated time.

Prints ONE JSON line on stdout:
  value       zone-cycles/s = zones * cycles / wall (HIP events on the
              library's stream around exactly --steps cycles)
  roofline    dominant kernel of the timed form: algorithmic bytes per launch
              / its average duration (HIP events between the launches, in
              further cycles right after the timed ones) against 8 TB/s;
              roofline_fp64: the same for the plain FP64 form
  accuracy    speedup over plain FP64; density/energy/pressure L1 and Linf
              errors, shock radius at tstop; the 1% density gate
  energy      total energy (internal + kinetic) at t = 0 and at tstop, and
              its relative drift, per form
  drop_in_tasks  the task-by-task FP64 cycle (one kernel launch per
              reference task ID, through the C ABI's per-task entry points) on
              the same workload from t = 0: us per cycle and the reference
              decomposition's bytes per cycle (688 S + 536 Z + 512 P) over that
              time, against 8 TB/s
  cpu_baseline  the CPU oracle (a serial C++ restatement of the reference's
              CPU task bodies, and its OpenMP variant) on a bounded sample of
              the same workload, on this host (lscpu model, sockets, cores)
Diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "pennant-legion_amd", ROOT / "tests"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
HEADLINE_METRIC = "zone-cycles/s on 1 MI355X (learned+mixed vs plain FP64) at \u22641% L1 density err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_deck(name: str, decks):
    big = 10 ** 8   # cstop large enough that the timed cycles never stop early
    if name == "sedov1440":
        return decks.sedov_1440(cstop=big), "sedov-rect-1440x1440"
    if name == "noh1024":
        return decks.noh_pie_1024(cstop=big), "noh-pie-1024x1024"
    if name == "leblanc4096":
        return decks.leblanc_4096x256(cstop=big), "leblanc-rect-4096x256"
    raise SystemExit(f"unknown workload {name}")


def kernel_bytes(S: int, Z: int, P: int, mixed: bool = False, pipelined: bool = False,
                 rect_maps: bool = False):
    """Algorithmic (compulsory) HBM bytes per launch of each fused kernel:
    every distinct input element read once, every output written once
    (int32 maps, FP64 values; FP32 side and corner arrays in the mixed form).
    With rect_maps the zone kernels compute the zones' points (mapsp1) and
    corner slots (pcpos) and read neither."""
    kb = _kernel_bytes(S, Z, P, mixed, pipelined)
    if rect_maps:
        maps = {"zone": 8, "zone_pred": 8, "zone_corr": 4}
        kb = {k: v - maps.get(k, 0) * S for k, v in kb.items()}
    return kb


def _kernel_bytes(S, Z, P, mixed, pipelined):
    if pipelined:
        return {
            # corrector and next predictor of each zone: mapsp1 4, pcpos 4,
            # smf 4, sfw 8 in + 8 out, corner record 12 out per side; zetot
            # zvol zm zpn 32 in, zwrate zetot zvol zp zpn 40 out per zone;
            # px 16, pu 16, (pu0 + pu) pxp.x 16 per point
            "zone": 40 * S + 72 * Z + 48 * P,
            "point": 12 * S + 85 * P,
        }
    if mixed:
        return {
            # mapsp1 4, pcpos 4, smf 4, cf 8, cmass 4, sfw 8 per side; zm zvol
            # zetot zwrate 32 in (zr and ze derived), zp 8 out per zone
            "zone_pred": 32 * S + 40 * Z + 32 * P,
            # cf 8, cmass 4 per corner; pcbeg 4, px/pu 32 in + 32 out, bc bits
            # 1, (pu0 + pu) pxp.x 16 out per point
            "point": 12 * S + 85 * P,
            # mapsp1 4, sfw 8 per side; zetot zvol zp 24 in, zwrate zetot zvol
            # 24 out per zone; px 16, (pu0 + pu) pxp.x 16 per point
            "zone_corr": 12 * S + 48 * Z + 32 * P,
        }
    return {
        # mapsp1 4, pcpos 4, smf 8, cf.x/cf.y/cmass/sfw.x/sfw.y 40 per side;
        # zm zr ze zwrate zvol 40 + zp 8 per zone; px, pu 32 per point
        "zone_pred": 56 * S + 48 * Z + 32 * P,
        # cmass 8, cf 16 per corner; pcbeg 4, px/pu 32 in + 32 out, bc bits 1,
        # vs 16 + pxp.x 8 out per point
        "point": 24 * S + 93 * P,
        # mapsp1 4, sfw 16 per side; zetot zvol zm zp 32 in, zwrate zetot ze
        # zr zvol 40 out per zone; px 16, vs 16, pxp.x 8 per point
        "zone_corr": 20 * S + 72 * Z + 40 * P,
    }


def bench_ensemble(args, pa, decks):
    """The batched ensemble: 256 Sedov (rect) and Noh (pie) problems of
    log-uniform sizes 32^2..512^2 advanced together, one launch per phase,
    16-cycle HIP graphs, autotuned grid.  value = zone-cycles actually
    advanced per second (problems that stop early stop counting)."""
    mode = {"learned": pa.MODE_LEARNED, "mixed": pa.MODE_MIXED, "fused": pa.MODE_FUSED}[args.mode]
    texts, n, _ = decks.ensemble(256, cstop=10 ** 7)
    t0 = time.perf_counter()
    b = pa.Batch(texts)
    log(f"ensemble setup {time.perf_counter() - t0:.1f}s: 256 problems, {int((n.astype('int64') ** 2).sum())} zones")
    if mode == pa.MODE_LEARNED:
        from pennant_amd.closure import Mlp
        from pennant_amd.train_closure import DEFAULT_OUT
        b.set_closure(Mlp.load(Path(args.closure) if args.closure else DEFAULT_OUT))
    if args.graph is not None:
        b.set_tuning("graph", args.graph)
    ms_tune = b.autotune(mode, cycles=4)
    log(f"autotune: max_blocks {b.get_tuning('max_blocks')} xcd {b.get_tuning('xcd')} ({ms_tune:.3f} ms / 4 cycles)")
    b.run(args.warmup, mode, raise_on_error=False)
    zones = [b.sizes(i)[1] for i in range(256)]
    c0 = [b.scalars(i)["cycle"] for i in range(256)]
    total_ms = b.time_cycles(args.steps, mode, raise_on_error=False)
    c1 = [b.scalars(i)["cycle"] for i in range(256)]
    zc = sum(z * (e - s) for z, s, e in zip(zones, c0, c1))
    value = zc / (total_ms * 1e-3)
    stopped = b.stopped()
    log(f"ensemble: {args.steps} cycles in {total_ms:.2f} ms, {zc:.4g} zone-cycles -> {value:.4e} zone-cycles/s; "
        f"stopped on bad sides: {stopped}")
    line = {"metric": "zone-cycles/s on 1 MI355X, batched ensemble of 256 Sedov/Noh problems",
            "value": round(value, 1), "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(total_ms / args.steps, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": {pa.MODE_LEARNED: "bf16+f32+f64", pa.MODE_MIXED: "f32+f64"}.get(mode, "f64"),
            "data": "synthetic",
            "config": {"workload": "ensemble-256 (std::mt19937_64 20261202)", "form": args.mode,
                       "zones": int(sum(zones)), "problems": 256, "stopped_on_bad_sides": len(stopped),
                       "max_blocks": b.get_tuning("max_blocks"), "xcd": b.get_tuning("xcd"), "graph": 1}}
    b.close()
    print(json.dumps(line), flush=True)


def bench_pieces(args, pa, text, wname):
    """The workload as --pieces pieces on the device against the one-piece
    run: both advance --warmup cycles from the start, then --steps cycles
    are timed, both replaying 16-cycle HIP graphs (--graph 0: the one-piece
    run launches directly; the run in pieces always replays its graph, in
    which each phase's piece kernels are parallel branches).  The end states
    are compared bit for bit."""
    import numpy as np
    mode = {"fused": pa.MODE_FUSED, "mixed": pa.MODE_MIXED}.get(args.mode)
    if mode is None:
        raise SystemExit("--pieces runs the fused or mixed mode")
    with pa.Problem(text) as one, pa.Pieces(text, args.pieces) as pp:
        lay = [pp.layout(i) for i in range(pp.count)]
        one.set_tuning("graph", 0 if args.graph == 0 else 1)
        one.run(args.warmup, mode)
        pp.run(args.warmup, mode)
        ms1, _ = one.time_cycles(args.steps, mode, phases=False)
        msp = pp.time_cycles(args.steps, mode)
        same = all(np.array_equal(one.get(f).view(np.uint64), pp.get(f).view(np.uint64))
                   for f in ("px", "pu", "zr", "ze", "zp", "zetot"))
        Z = one.numz
    log(f"one piece: {ms1 / args.steps * 1e3:.1f} us/cycle; {pp.count} pieces: {msp / args.steps * 1e3:.1f} us/cycle; "
        f"bit-identical: {same}")
    line = {"metric": f"zone-cycles/s on 1 MI355X, {wname} as {pp.count} pieces", "value": round(Z * args.steps / (msp * 1e-3), 1),
            "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(msp / args.steps, 5), "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f64" if mode == pa.MODE_FUSED else "f32+f64", "data": "synthetic",
            "config": {"workload": wname, "pieces": pp.count, "numpcx": lay[0]["numpcx"], "numpcy": lay[0]["numpcy"],
                       "ghost_points": sum(d["ghost"] for d in lay), "received_corners": sum(d["received"] for d in lay),
                       "one_piece_us_per_cycle": round(ms1 / args.steps * 1e3, 2), "bit_identical": same}}
    print(json.dumps(line), flush=True)


PROFILE_KERNELS = {
    "fp64": {"zone_pred": "k_pred_zone<4>", "point": "k_point<double>", "zone_corr": "k_corr<4, double>"},
    "mixed": {"zone_pred": "k_pred_mixed<4, false, false>", "point": "k_point<float>",
              "zone_corr": "k_corr<4, float>", "zone": "k_zone_pipe<4, false>"},
    "learned": {"zone_pred": "k_pred_mixed<4, true, false>", "point": "k_point<float>",
                "zone_corr": "k_corr<4, float>", "zone": "k_zone_pipe<4, true>"},
}


def profiled_traffic(workload: str, kernel: str, form: str = "fp64"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = ROOT / "profiles" / f"fused_{form}_{workload}.md"
    kmap = PROFILE_KERNELS[form]
    if not path.exists() or kernel not in kmap:
        return None
    name = f"`{kmap[kernel]}`"
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 8 and cells[0] == name:
            return int((float(cells[3]) + float(cells[4])) * 1e6)
    return None


def host_cpu():
    """The host's CPU as lscpu reports it: model, sockets, cores per socket,
    threads per core, and the logical CPUs this process may use."""
    info = {"nproc": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except Exception as ex:            # lscpu missing: report what os knows
        info["lscpu_error"] = str(ex)
    return info


def cpu_baseline(deck_text: str, numz: int, budget_s: float):
    """Times the oracle (serial, then OpenMP) on whole cycles of the same
    workload after setup, for about budget_s seconds each."""
    from oracle_lib import Oracle, lib
    out = {}
    for omp in (False, True):
        o = Oracle(deck_text, omp=omp)
        t0 = time.perf_counter()
        o.run(1)
        one = time.perf_counter() - t0
        n = max(1, min(50, int(budget_s / max(one, 1e-6))))
        t0 = time.perf_counter()
        o.run(n)
        dt = time.perf_counter() - t0
        out["omp" if omp else "serial"] = (numz * n / dt, n, dt)
        del o
    threads = lib().ora_omp_threads()
    return out, threads


def kernel_roofline(kb: dict, per_ms: dict, names) -> dict:
    """Roofline of the dominant (longest) kernel: its algorithmic bytes per
    launch over its average launch duration, against 8 TB/s."""
    dom = max(kb, key=lambda n: per_ms[n])
    achieved = kb[dom] / (per_ms[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": kb[dom],
            "avg_launch_us": round(per_ms[dom] * 1e3, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--phase-steps", type=int, default=20,
                    help="cycles after the timed ones, with events between the kernels (direct "
                         "launches), for the per-kernel durations (0 disables)")
    ap.add_argument("--mid-time", type=float, default=0.5,
                    help="simulated time each form is advanced to (untimed, graph replay) before "
                         "the warm-up, so that the timed cycles run mid-problem, where the shock "
                         "front and the learned closure's work are developed (0: time from the "
                         "start)")
    ap.add_argument("--to-end", type=int, default=1,
                    help="1: after timing, run every form on to tstop and compare them there, at "
                         "equal simulated time (density L1, shock radius, total-energy drift)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: pennant_autotune picks each form's launch shape (grid cap, XCD "
                         "mapping, K1 layout, pipelined cycle) after the warm-up; 0: defaults or "
                         "the explicit options")
    ap.add_argument("--pipeline", type=int, default=None,
                    help="mixed forms: 1 pipelined cycle, 0 separate predictor and corrector "
                         "passes (the default)")
    ap.add_argument("--pieces", type=int, default=0,
                    help="N > 1: also time the workload run as N pieces on the device (the reference's "
                         "-n N; fused and mixed modes) against the one-piece run, from the same start")
    ap.add_argument("--workload", default="sedov1440")
    ap.add_argument("--mode", default="learned", choices=["learned", "mixed", "fused", "tasks"],
                    help="learned: mixed precision with the learned QCS closure; mixed: mixed "
                         "precision with the exact QCS (both reported against the plain FP64 fused "
                         "cycle on the same workload); fused: plain FP64 fused cycle; tasks: one "
                         "launch per task")
    ap.add_argument("--closure", default=None,
                    help="trained closure (.npz from pennant_amd.train_closure); default: the "
                         "package's closure_ensemble.npz")
    ap.add_argument("--tasks-steps", type=int, default=20,
                    help="cycles of the task-by-task FP64 form timed for drop_in_tasks (0: skip)")
    ap.add_argument("--copy-mb", type=int, default=512,
                    help="size of the streaming-copy calibration (0 disables)")
    ap.add_argument("--k1", type=int, default=None, help="fused K1 layout: 0 zone/thread, 1 corner/lane")
    ap.add_argument("--max-blocks", type=int, default=None, help="grid cap of the fused zone kernels")
    ap.add_argument("--xcd", type=int, default=None, help="XCD band mapping of the zone kernels (0/1)")
    ap.add_argument("--rect-maps", type=int, default=None,
                    help="0: the zone kernels read mapsp1 and the slot map on rect meshes too")
    ap.add_argument("--graph", type=int, default=None,
                    help="0: launch every cycle's kernels directly instead of replaying 16-cycle HIP "
                         "graphs (for tools that trace kernel launches)")
    ap.add_argument("--cpu-budget", type=float, default=8.0,
                    help="seconds of CPU time per oracle variant (0 disables)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs "
              "on a single device)", file=sys.stderr)
        sys.exit(2)

    import pennant_amd as pa
    from pennant_amd import decks
    from pennant_amd.accuracy import field_errors
    from pennant_amd.diagnostics import shock_radius, zone_centroids

    if args.workload == "ensemble256":
        bench_ensemble(args, pa, decks)
        return
    text, wname = workload_deck(args.workload, decks)
    if args.pieces > 1:
        bench_pieces(args, pa, text, wname)
        return
    modes = {"learned": pa.MODE_LEARNED, "mixed": pa.MODE_MIXED, "fused": pa.MODE_FUSED,
             "tasks": pa.MODE_TASKS}
    mlp = None
    if args.mode == "learned":
        from pennant_amd.closure import Mlp
        from pennant_amd.train_closure import DEFAULT_OUT
        path = Path(args.closure) if args.closure else DEFAULT_OUT
        if not path.exists():
            raise SystemExit(f"no trained closure at {path}: run python -m pennant_amd.train_closure")
        mlp = Mlp.load(path)
    hm = pa.HostMesh(text)
    znump, mapsp1 = hm.get("znump"), hm.get("mapsp1")

    def measure(mode):
        """One form: set up, advance untimed to --mid-time, warm up, tune,
        time --steps cycles (then --phase-steps with per-kernel events), and
        with --to-end run on to tstop.  Returns the timings and the states."""
        t0 = time.perf_counter()
        pb = pa.Problem(text)
        if mode == pa.MODE_LEARNED:
            pb.set_closure(mlp)
        if args.k1 is not None:
            pb.set_tuning("k1_layout", args.k1)
        if args.max_blocks is not None:
            pb.set_tuning("max_blocks", args.max_blocks)
        if args.xcd is not None:
            pb.set_tuning("xcd", args.xcd)
        if args.pipeline is not None:
            pb.set_tuning("pipeline", args.pipeline)
        if args.graph is not None:
            pb.set_tuning("graph", args.graph)
        if args.rect_maps is not None:
            pb.set_tuning("rect_maps", args.rect_maps)
        e0 = sum(pb.energy())
        log(f"setup {time.perf_counter() - t0:.2f}s: zones {pb.numz} sides {pb.nums} points {pb.nump}")
        t0 = time.perf_counter()
        while pb.scalars()["time"] < args.mid_time and not pb.scalars()["done"]:
            pb.run(1024, mode)
        sc_mid = pb.scalars()
        if args.mid_time > 0:
            log(f"advanced to t={sc_mid['time']:.6f} (cycle {int(sc_mid['cycle'])}) in "
                f"{time.perf_counter() - t0:.1f}s")
        if args.warmup > 0:
            pb.run(args.warmup, mode)
        explicit = any(v is not None for v in (args.k1, args.max_blocks, args.xcd, args.pipeline))
        if args.autotune and not explicit and mode != pa.MODE_TASKS:
            # times every launch shape from a snapshot of the state and
            # restores it, so the problem does not advance
            tune_cycles = 16
            ms = pb.autotune(mode, cycles=tune_cycles)
            log(f"autotune: {ms / tune_cycles * 1e3:.1f} us/cycle at the best shape")
        tune = {k: pb.get_tuning(k) for k in ("k1_layout", "max_blocks", "xcd", "pipeline", "graph", "rect_maps")}
        log(f"tuning {tune}")
        sc0 = pb.scalars()
        cl0 = pb.closure_stats() if mode == pa.MODE_LEARNED else None
        rd0 = pb.lookahead_redos()
        # the timed cycles run as pennant_launch_cycles runs them (16-cycle
        # graphs); the per-kernel durations come from further cycles with
        # events between the kernels
        total_ms, _ = pb.time_cycles(args.steps, mode, phases=False)
        sc1 = pb.scalars()
        assert sc1["cycle"] == sc0["cycle"] + args.steps and not sc1["done"], sc1
        cl1 = pb.closure_stats() if mode == pa.MODE_LEARNED else None
        rd1 = pb.lookahead_redos()
        phase_ms = None
        if mode != pa.MODE_TASKS and args.phase_steps > 0:
            _, phase_ms = pb.time_cycles(args.phase_steps, mode, phases=True)
        res = {"total_ms": total_ms, "phase_ms": phase_ms, "tune": tune, "e0": e0,
               "t_window": (sc0["time"], sc1["time"]), "cycle_window": (int(sc0["cycle"]), int(sc1["cycle"])),
               "closure_window": None if cl0 is None else (cl1[0] - cl0[0], cl1[1] - cl0[1]),
               "redos_window": rd1 - rd0,
               "dims": (pb.numz, pb.nums, pb.nump)}
        if args.to_end:
            t0 = time.perf_counter()
            while not pb.scalars()["done"]:
                pb.run(8192, mode)
            sc = pb.scalars()
            res["end"] = {"time": sc["time"], "cycles": int(sc["cycle"]), "e": sum(pb.energy()),
                          "state": {f: pb.get(f) for f in ("zr", "ze", "zp")}, "px": pb.get("px"),
                          "closure": pb.closure_stats() if mode == pa.MODE_LEARNED else None,
                          "redos": pb.lookahead_redos()}
            log(f"  ran on to t={sc['time']!r} (cycle {int(sc['cycle'])}) in {time.perf_counter() - t0:.1f}s")
        pb.close()
        return res

    def rate(r):
        return r["dims"][0] * args.steps / (r["total_ms"] * 1e-3)

    def phase_roofline(r, mixed):
        if r["phase_ms"] is None:
            return None, None
        Z, S, P = r["dims"]
        pipelined = mixed and r["tune"]["pipeline"] == 1
        kb = kernel_bytes(S, Z, P, mixed, pipelined, r["tune"].get("rect_maps", 0) == 1)
        names = (("zone", "point", "unused", "finalize", "learned_pass") if pipelined
                 else ("zone_pred", "point", "zone_corr", "finalize", "learned_pass"))
        per = {n: r["phase_ms"][i] / args.phase_steps for i, n in enumerate(names)}
        for n in names:
            if n == "unused":
                continue
            gbs = kb[n] / (per[n] * 1e-3) / 1e9 if n in kb else 0.0
            log(f"  {n:12s} {per[n] * 1e3:8.1f} us/launch" +
                (f"  {kb[n] / 1e6:8.1f} MB  {gbs:8.1f} GB/s" if n in kb else ""))
        return kernel_roofline(kb, per, names), (kb, per, names, pipelined)

    mode = modes[args.mode]
    main_r = measure(mode)
    Z, S, P = main_r["dims"]
    value = rate(main_r)
    ms_per_step = main_r["total_ms"] / args.steps
    log(f"{args.mode}: {args.steps} cycles in {main_r['total_ms']:.3f} ms -> {ms_per_step * 1e3:.1f} us/cycle, "
        f"{value:.4e} zone-cycles/s; t window {main_r['t_window'][0]:.6f}..{main_r['t_window'][1]:.6f}")

    mixed = mode in (pa.MODE_MIXED, pa.MODE_LEARNED)
    roofline, rdetail = phase_roofline(main_r, mixed)
    pform = {pa.MODE_LEARNED: "learned", pa.MODE_MIXED: "mixed"}.get(mode, "fp64")
    if roofline is not None:
        kb, per, names, pipelined = rdetail
        roofline["traffic"] = (profiled_traffic(args.workload, roofline["kernel"], pform)
                               if main_r["tune"]["k1_layout"] == 0 else None)
        roofline["phase_us"] = {n: round(per[n] * 1e3, 2) for n in names
                                if n != "unused" and (n != "learned_pass" or mode == pa.MODE_LEARNED)}
        roofline["cycle_alg_GBs"] = round(sum(kb.values()) / (ms_per_step * 1e-3) / 1e9, 1)
        # the reference's task decomposition moves 688 S + 536 Z + 512 P bytes
        # per cycle; at the fused cycle's rate that is this effective bandwidth
        roofline["task_decomposition_effective_GBs"] = round(
            (688 * S + 536 * Z + 512 * P) / (ms_per_step * 1e-3) / 1e9, 1)
        if args.copy_mb > 0:
            import ctypes
            ms = ctypes.c_double()
            nbytes = args.copy_mb << 20
            if pa.lib().pennant_bench_copy(nbytes, 5, ctypes.byref(ms)) == 0:
                copy_gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
                log(f"  streaming copy of {args.copy_mb} MiB: {copy_gbs:.1f} GB/s (read + write)")
                roofline["copy_GBs"] = round(copy_gbs, 1)

    forms = {args.mode: main_r}
    roofline_fp64 = None
    accuracy = None
    if mixed:
        # the plain FP64 form on the same workload: its mid-run rate is the
        # denominator of the speedup and its state at tstop the accuracy
        # reference; the mixed form with the exact QCS beside them
        forms["fused"] = measure(pa.MODE_FUSED)
        if mode == pa.MODE_LEARNED:
            forms["mixed"] = measure(pa.MODE_MIXED)
        plain = rate(forms["fused"])
        roofline_fp64, _ = phase_roofline(forms["fused"], False)
        if roofline_fp64 is not None:
            roofline_fp64["traffic"] = (profiled_traffic(args.workload, roofline_fp64["kernel"], "fp64")
                                        if forms["fused"]["tune"]["k1_layout"] == 0 else None)
        log(f"plain fp64 fused: {forms['fused']['total_ms'] / args.steps * 1e3:.1f} us/cycle, "
            f"{plain:.4e} zone-cycles/s; speedup {value / plain:.3f}")
        accuracy = {"reference": "plain FP64 fused form (bit-identical to the CPU oracle)",
                    "plain_fp64_value": round(plain, 1),
                    "plain_fp64_tuning": forms["fused"]["tune"],
                    "speedup_vs_plain_fp64": round(value / plain, 4),
                    "timed_window": {k: {"t": [round(v, 6) for v in r["t_window"]],
                                         "cycles": list(r["cycle_window"]),
                                         "lookahead_redos": r["redos_window"]} for k, r in forms.items()}}
        if "mixed" in forms and args.mode != "mixed":
            mixed_v = rate(forms["mixed"])
            log(f"mixed exact QCS: {forms['mixed']['total_ms'] / args.steps * 1e3:.1f} us/cycle, "
                f"{mixed_v:.4e} zone-cycles/s")
            accuracy["mixed_exact_qcs_value"] = round(mixed_v, 1)
            accuracy["learned_vs_mixed_exact_speed"] = round(value / mixed_v, 4)
        if mode == pa.MODE_LEARNED:
            accuracy["closure_zone_evaluations_in_window"] = dict(
                zip(("learned", "exact_fallback"), main_r["closure_window"]))
        if args.to_end:
            ref = forms["fused"]["end"]
            cen = zone_centroids(znump, mapsp1, ref["px"])
            end = {"t": ref["time"], "fp64": {"cycles": ref["cycles"],
                                              "shock_radius": round(shock_radius(ref["state"]["zr"], cen), 6)}}
            for k, r in forms.items():
                if k == "fused":
                    continue
                e = r["end"]
                err = field_errors(e["state"], ref["state"])
                end[k] = {"cycles": e["cycles"], "lookahead_redos": e["redos"],
                          "rel_l1": {f: float(f"{v[0]:.3e}") for f, v in err.items()},
                          "rel_linf": {f: float(f"{v[1]:.3e}") for f, v in err.items()},
                          "shock_radius": round(shock_radius(e["state"]["zr"],
                                                             zone_centroids(znump, mapsp1, e["px"])), 6)}
                if e["closure"] is not None:
                    end[k]["closure_zone_evaluations"] = dict(zip(("learned", "exact_fallback"), e["closure"]))
                log(f"  {k} at t={e['time']!r}: " +
                    " ".join(f"{f} L1 {v[0]:.2e} Linf {v[1]:.2e}" for f, v in err.items()))
            accuracy["at_tstop"] = end
            gate = end[args.mode]["rel_l1"]["zr"] if args.mode in end else None
            accuracy["density_l1_at_tstop"] = gate
            accuracy["density_l1_gate_1pct_ok"] = (gate is not None and gate <= 0.01)
    # total energy E = internal + kinetic (per radian, as zm and zvol), at
    # the start and at tstop, per form
    if args.to_end:
        energy = {}
        for k, r in forms.items():
            e0, e1 = r["e0"], r["end"]["e"]
            energy[k] = {"E0": e0, "E_end": e1, "t_end": r["end"]["time"], "drift": float(f"{e1 / e0 - 1:.4e}")}
            log(f"  energy {k}: E0 {e0!r} E(t={r['end']['time']}) {e1!r} drift {e1 / e0 - 1:+.4e}")
    else:
        energy = None

    # the drop-in path: one launch per reference task ID, from t = 0 (its
    # kernels move the same bytes at any time of the run)
    tasks = None
    if args.tasks_steps > 0 and mode != pa.MODE_TASKS:
        with pa.Problem(text) as pt:
            pt.run(2, pa.MODE_TASKS)
            sct0 = pt.scalars()
            tms, _ = pt.time_cycles(args.tasks_steps, pa.MODE_TASKS, phases=False)
            sct1 = pt.scalars()
        us = tms / args.tasks_steps * 1e3
        tb = 688 * S + 536 * Z + 512 * P
        gbs = tb / (us * 1e-6) / 1e9
        log(f"drop-in tasks fp64: {us:.1f} us/cycle, {tb / 1e9:.3f} GB of task bytes per cycle, "
            f"{gbs:.1f} GB/s ({gbs / HBM_PEAK_GBS:.3f} of peak)")
        tasks = {"form": "plain-fp64-tasks", "us_per_cycle": round(us, 2),
                 "zone_cycles_per_s": round(Z / (us * 1e-6), 1),
                 "task_bytes_per_cycle": tb, "effective_GBs": round(gbs, 1),
                 "peak_GBs": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4),
                 "t_window": [round(sct0["time"], 9), round(sct1["time"], 9)]}

    cpu = None
    if args.cpu_budget > 0:
        res, threads = cpu_baseline(text, Z, args.cpu_budget)
        v_omp, n_omp, t_omp = res["omp"]
        v_ser, n_ser, t_ser = res["serial"]
        log(f"cpu oracle serial {v_ser:.3e} zone-cycles/s ({n_ser} cycles, {t_ser:.1f}s); "
            f"omp x{threads} {v_omp:.3e} ({n_omp} cycles, {t_omp:.1f}s)")
        cpu = {"value": round(v_omp, 1), "unit": "zone-cycles/s", "cores": threads, "kind": "port",
               "sample": f"{n_omp} cycles of {wname} after setup (OpenMP variant, {threads} threads); "
                         f"serial variant {v_ser:.4g} zone-cycles/s on 1 core over {n_ser} cycles",
               "serial_value": round(v_ser, 1), "host": host_cpu()}

    form = {"learned": "mixed-fp32-learned-qcs-fused", "mixed": "mixed-fp32-qcs-fused",
            "fused": "plain-fp64-fused", "tasks": "plain-fp64-tasks"}
    line = {
        "metric": HEADLINE_METRIC,
        "value": round(value, 1), "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {pa.MODE_LEARNED: "bf16+f32+f64", pa.MODE_MIXED: "f32+f64"}.get(mode, "f64"),
        "data": "synthetic",
        "config": {"workload": wname, "form": form[args.mode], "zones": Z, "sides": S, "points": P,
                   "timed_from_t": round(main_r["t_window"][0], 6), "tuning": main_r["tune"]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if roofline_fp64 is not None:
        line["roofline_fp64"] = roofline_fp64
    if tasks is not None:
        line["drop_in_tasks"] = tasks
    if accuracy is not None:
        line["accuracy"] = accuracy
    if energy is not None:
        line["energy"] = energy
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
