"""PENNANT hydro cycle benchmark on one MI355X.

Workload (default): Sedov blast on a 1440 x 1440 rect mesh (2,073,600 zones,
8,294,400 sides), fused cycle in mixed precision with the learned QCS
closure (trained weights loaded from the package; training is not timed),
reported against the plain FP64 fused cycle (bit-identical to the CPU
oracle) and the mixed form with the exact QCS on the same workload.  A step
is one hydro cycle over the whole mesh; the state is resident in HBM before
the timed region starts.

Prints ONE JSON line on stdout:
  value       zone-cycles/s = zones * cycles / wall (HIP events on the
              library's stream around exactly --steps cycles)
  roofline    dominant kernel: algorithmic bytes per launch / its average
              duration (HIP events between the launches in the same timed
              region) against MI355X's 8 TB/s
  cpu_baseline  the CPU oracle (a serial C++ restatement of the reference's
              CPU task bodies, and its OpenMP variant) on a bounded sample of
              the same workload, on this host
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


def kernel_bytes(S: int, Z: int, P: int, mixed: bool = False, pipelined: bool = False):
    """Algorithmic (compulsory) HBM bytes per launch of each fused kernel:
    every distinct input element read once, every output written once
    (int32 maps, FP64 values; FP32 side and corner arrays in the mixed form)."""
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


def physics_check(pa, decks, mlp, n):
    """Sedov n x n to t = 1 in the plain FP64, mixed and learned forms: density
    L1 error, shock radius (densest zone) and total energy against FP64."""
    from pennant_amd.accuracy import rel_l1
    from pennant_amd.diagnostics import shock_radius, zone_centroids
    text = decks.sedov_rect(n)
    hm = pa.HostMesh(text)
    znump, mapsp1 = hm.get("znump"), hm.get("mapsp1")
    res = {}
    for name, mode in (("fp64", pa.MODE_FUSED), ("mixed", pa.MODE_MIXED), ("learned", pa.MODE_LEARNED)):
        with pa.Problem(text) as pb:
            if mode == pa.MODE_LEARNED:
                pb.set_closure(mlp)
            cyc = 0
            while not pb.scalars()["done"]:
                cyc += pb.run(4096, mode)
            zr, px = pb.get("zr"), pb.get("px")
            ei, ek = pb.energy()
            res[name] = (zr, shock_radius(zr, zone_centroids(znump, mapsp1, px)), ei + ek, cyc,
                         pb.closure_stats() if mode == pa.MODE_LEARNED else None)
    e0 = res["fp64"][2]
    out = {"workload": f"sedov-rect-{n}x{n} to t=1", "cycles": res["fp64"][3],
           "shock_radius_fp64": round(res["fp64"][1], 6), "energy_fp64": e0}
    for name in ("mixed", "learned"):
        zr, rs, et, cyc, st = res[name]
        out[name] = {"rel_l1_zr": float(f"{rel_l1(zr, res['fp64'][0]):.3e}"), "shock_radius": round(rs, 6),
                     "rel_energy_diff": float(f"{(et - e0) / e0:.3e}"), "cycles": cyc}
        if st is not None:
            out[name]["closure_zone_evaluations"] = {"learned": st[0], "exact_fallback": st[1]}
    log(f"physics check {out}")
    return out


PROFILE_KERNELS = {
    "fp64": {"zone_pred": "k_pred_zone<4>", "point": "k_point<double>", "zone_corr": "k_corr<4, double>"},
    "mixed": {"zone_pred": "k_pred_mixed<4, false, false>", "point": "k_point<float>",
              "zone_corr": "k_corr<4, float>"},
    "learned": {"zone_pred": "k_pred_mixed<4, true, false>", "point": "k_point<float>",
                "zone_corr": "k_corr<4, float>"},
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--phase-steps", type=int, default=20,
                    help="cycles after the timed ones, with events between the kernels (direct "
                         "launches), for the per-kernel durations (0 disables)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: pennant_autotune picks each form's launch shape (grid cap, XCD "
                         "mapping, K1 layout, pipelined cycle) after the warm-up; 0: defaults or "
                         "the explicit options")
    ap.add_argument("--pipeline", type=int, default=None,
                    help="mixed forms: 1 pipelined cycle, 0 separate predictor and corrector "
                         "passes (the default)")
    ap.add_argument("--workload", default="sedov1440")
    ap.add_argument("--mode", default="learned", choices=["learned", "mixed", "fused", "tasks"],
                    help="learned: mixed precision with the learned QCS closure; mixed: mixed "
                         "precision with the exact QCS (both reported against the plain FP64 fused "
                         "cycle on the same workload); fused: plain FP64 fused cycle; tasks: one "
                         "launch per task")
    ap.add_argument("--closure", default=None,
                    help="trained closure (.npz from pennant_amd.train_closure); default: the "
                         "package's closure_ensemble.npz")
    ap.add_argument("--physics-n", type=int, default=360,
                    help="Sedov n x n run to t=1 comparing shock radius and energy of the forms "
                         "(0 disables)")
    ap.add_argument("--copy-mb", type=int, default=512,
                    help="size of the streaming-copy calibration (0 disables)")
    ap.add_argument("--k1", type=int, default=None, help="fused K1 layout: 0 zone/thread, 1 corner/lane")
    ap.add_argument("--max-blocks", type=int, default=None, help="grid cap of the fused zone kernels")
    ap.add_argument("--xcd", type=int, default=None, help="XCD band mapping of the zone kernels (0/1)")
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

    if args.workload == "ensemble256":
        bench_ensemble(args, pa, decks)
        return
    text, wname = workload_deck(args.workload, decks)
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

    def measure(mode):
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
        log(f"setup {time.perf_counter() - t0:.2f}s: zones {pb.numz} sides {pb.nums} points {pb.nump}")
        if args.warmup > 0:
            pb.run(args.warmup, mode)
        explicit = any(v is not None for v in (args.k1, args.max_blocks, args.xcd, args.pipeline))
        if args.autotune and not explicit and mode != pa.MODE_TASKS:
            # times every launch shape from a snapshot of the state and
            # restores it, so the problem does not advance
            tune_cycles = 16
            ms = pb.autotune(mode, cycles=tune_cycles)
            log(f"autotune: {ms / tune_cycles * 1e3:.1f} us/cycle at the best shape")
        tune = {k: pb.get_tuning(k) for k in ("k1_layout", "max_blocks", "xcd", "pipeline", "graph")}
        log(f"tuning {tune}")
        # the timed cycles run as pennant_launch_cycles runs them (16-cycle
        # graphs); the per-kernel durations come from further cycles with
        # events between the kernels
        total_ms, _ = pb.time_cycles(args.steps, mode, phases=False)
        phase_ms = None
        if mode != pa.MODE_TASKS and args.phase_steps > 0:
            _, phase_ms = pb.time_cycles(args.phase_steps, mode, phases=True)
        sc = pb.scalars()
        assert sc["cycle"] == args.warmup + args.steps + (args.phase_steps if phase_ms else 0), sc
        state = {f: pb.get(f) for f in ("zr", "ze", "zp")}
        if mode == pa.MODE_LEARNED:
            state["closure_stats"] = pb.closure_stats()
        dims = (pb.numz, pb.nums, pb.nump)
        pb.close()
        return total_ms, phase_ms, sc, state, dims, tune

    mode = modes[args.mode]
    total_ms, phase_ms, sc, state, (Z, S, P), tune = measure(mode)
    ms_per_step = total_ms / args.steps
    value = Z * args.steps / (total_ms * 1e-3)
    log(f"{args.mode}: {args.steps} cycles in {total_ms:.3f} ms -> {ms_per_step * 1e3:.1f} us/cycle, "
        f"{value:.4e} zone-cycles/s; t={sc['time']:.6e}")

    accuracy = None
    if mode in (pa.MODE_MIXED, pa.MODE_LEARNED):
        # the plain FP64 form on the same workload, same cycle count: its rate
        # is the denominator of the speedup and its state the accuracy reference
        p_ms, _, p_sc, p_state, _, p_tune = measure(pa.MODE_FUSED)
        plain = Z * args.steps / (p_ms * 1e-3)
        err = field_errors(state, p_state)
        log(f"plain fp64 fused: {p_ms / args.steps * 1e3:.1f} us/cycle, {plain:.4e} zone-cycles/s; "
            f"speedup {value / plain:.3f}; " +
            " ".join(f"{k} L1 {v[0]:.2e} Linf {v[1]:.2e}" for k, v in err.items()))
        accuracy = {"reference": "plain FP64 fused form (bit-identical to the CPU oracle)",
                    "cycles": args.warmup + args.steps,
                    "time": sc["time"], "time_plain": p_sc["time"],
                    "rel_l1": {k: float(f"{v[0]:.3e}") for k, v in err.items()},
                    "rel_linf": {k: float(f"{v[1]:.3e}") for k, v in err.items()},
                    "plain_fp64_value": round(plain, 1),
                    "plain_fp64_tuning": p_tune,
                    "speedup_vs_plain_fp64": round(value / plain, 4)}
        if mode == pa.MODE_LEARNED:
            m_ms, _, _, m_state, _, _ = measure(pa.MODE_MIXED)
            mixed_v = Z * args.steps / (m_ms * 1e-3)
            nl, nf = state["closure_stats"]
            log(f"mixed exact QCS: {m_ms / args.steps * 1e3:.1f} us/cycle, {mixed_v:.4e} zone-cycles/s; "
                f"closure zones learned {nl} fallback {nf}")
            accuracy["mixed_exact_qcs_value"] = round(mixed_v, 1)
            accuracy["closure_zone_evaluations"] = {"learned": nl, "exact_fallback": nf}
            accuracy["rel_l1_learned_vs_mixed_exact"] = float(f"{field_errors(state, m_state)['zr'][0]:.3e}")
            if args.physics_n > 0:
                accuracy["physics"] = physics_check(pa, decks, mlp, args.physics_n)

    roofline = None
    if phase_ms is not None:
        mixed = (mode in (pa.MODE_MIXED, pa.MODE_LEARNED))
        pipelined = mixed and tune["pipeline"] == 1
        kb = kernel_bytes(S, Z, P, mixed, pipelined)
        names = (("zone", "point", "unused", "finalize", "learned_pass") if pipelined
                 else ("zone_pred", "point", "zone_corr", "finalize", "learned_pass"))
        per = {n: phase_ms[i] / args.phase_steps for i, n in enumerate(names)}
        for n in names:
            if n == "unused" or (n == "learned_pass" and mode != pa.MODE_LEARNED):
                continue
            gbs = kb[n] / (per[n] * 1e-3) / 1e9 if n in kb else 0.0
            log(f"  {n:12s} {per[n] * 1e3:8.1f} us/launch" +
                (f"  {kb[n] / 1e6:8.1f} MB  {gbs:8.1f} GB/s" if n in kb else ""))
        dom = max(kb, key=lambda n: per[n])
        achieved = kb[dom] / (per[dom] * 1e-3) / 1e9
        # the reference's task decomposition moves 688 S + 536 Z + 512 P bytes
        # per cycle; at the fused cycle's rate that is this effective bandwidth
        b_tasks = 688 * S + 536 * Z + 512 * P
        pform = {pa.MODE_LEARNED: "learned", pa.MODE_MIXED: "mixed"}.get(mode, "fp64")
        traffic = (profiled_traffic(args.workload, dom, pform)
                   if tune["k1_layout"] == 0 and not pipelined else None)
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic,
                    "alg_bytes_per_launch": kb[dom], "avg_launch_us": round(per[dom] * 1e3, 2),
                    "phase_us": {n: round(per[n] * 1e3, 2) for n in names
                                 if n != "unused" and (n != "learned_pass" or mode == pa.MODE_LEARNED)},
                    "cycle_alg_GBs": round(sum(kb.values()) / (ms_per_step * 1e-3) / 1e9, 1),
                    "task_decomposition_effective_GBs": round(b_tasks / (ms_per_step * 1e-3) / 1e9, 1)}
        if args.copy_mb > 0:
            import ctypes
            ms = ctypes.c_double()
            nbytes = args.copy_mb << 20
            if pa.lib().pennant_bench_copy(nbytes, 5, ctypes.byref(ms)) == 0:
                copy_gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
                log(f"  streaming copy of {args.copy_mb} MiB: {copy_gbs:.1f} GB/s (read + write)")
                roofline["copy_GBs"] = round(copy_gbs, 1)

    cpu = None
    if args.cpu_budget > 0:
        res, threads = cpu_baseline(text, Z, args.cpu_budget)
        v_omp, n_omp, t_omp = res["omp"]
        v_ser, n_ser, t_ser = res["serial"]
        log(f"cpu oracle serial {v_ser:.3e} zone-cycles/s ({n_ser} cycles, {t_ser:.1f}s); "
            f"omp x{threads} {v_omp:.3e} ({n_omp} cycles, {t_omp:.1f}s)")
        cpu = {"value": round(v_omp, 1), "unit": "zone-cycles/s", "cores": threads, "kind": "port",
               "sample": f"{n_omp} cycles of {wname} after setup (OpenMP variant); serial "
                         f"variant {v_ser:.4g} zone-cycles/s on 1 core over {n_ser} cycles",
               "serial_value": round(v_ser, 1)}

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
                   "tuning": tune},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if accuracy is not None:
        line["accuracy"] = accuracy
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
