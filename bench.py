#!/usr/bin/env python3
"""Throughput benchmark of the MI355X 3DThesis path (one GPU).

Metric: point-node evaluations per second (one evaluation = one Gaussian node
term at one point: 1 exp + 14 FLOP). A "step" is one complete run of the
workload's mode driver (every timestep of the Solidification run, or every
snapshot), inputs parsed and device state allocated beforehand, no files
written. `value` is the FP32-compensated throughput; the FP64 run of the same
workload, the accuracy of FP32c against it, the node-sum kernel's roofline and
the reference's CPU throughput on a bounded sample are reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2|cfg1]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

FP32_PEAK_TFLOPS = 157.3  # MI355X vector/matrix FP32, published
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_EVAL = 14        # 3 sub, 3 mul, 3 fma (6), 1 mul by dtau, 1 add; plus one exp


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def workload_deck(name: str):
    import decks
    if name == "cfg1":
        return decks.cfg1(), "Snapshots/None, 2 mm line, 200x50x20 grid, snapshot at scan end"
    if name == "cfg2":
        return decks.cfg2(), ("Solidification/Surface, 20-hatch 2 mm serpentine at 1 m/s, dt 1e-5 s, "
                              "400x400x40 grid over 4x4x0.4 mm")
    if name == "cfg3":
        return decks.cfg3(), ("Solidification/Volume, 4 beams on seeded rasters, 500x500x50 grid over "
                              "5x5x0.5 mm")
    raise SystemExit(f"unknown workload {name}")


def cpu_sample_deck(name: str):
    """A bounded slice of the workload for the CPU baseline (same grid and physics)."""
    import decks
    if name == "cfg1":
        return decks.cfg1(), "the full workload"
    if name != "cfg2":
        raise SystemExit(f"the CPU baseline is sampled on cfg1 or cfg2; run {name} with --no-cpu")
    return decks.cfg2(), "the full workload (20 hatches)"


def bench_cfg1(pkg, torch, reps=20) -> dict:
    """Snapshot node sums over the 200x50x20 cfg1 grid, device-resident nodes."""
    import decks
    with tempfile.TemporaryDirectory() as d:
        th = pkg.Thesis(decks.cfg1().write(d))
        nodes = th.build_nodes(th.info.scan_end)
        th.upload_nodes(nodes)
        pnum = th.info.pnum
        T = torch.empty(pnum, dtype=torch.float64, device="cuda")
        out = {}
        s = torch.cuda.Stream()  # a real stream handle: 0 would mean "the context's own stream"
        for name, prec in (("fp32c", pkg.FP32C), ("fp64", pkg.FP64)):
            th.eval_T_dev(T.data_ptr(), prec, 0, pnum, s.cuda_stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                th.eval_T_dev(T.data_ptr(), prec, 0, pnum, s.cuda_stream)
            e1.record(s)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            out[name] = {"evals_per_s": pnum * nodes.shape[0] / (ms * 1e-3), "ms": ms}
        out["evals"] = pnum * nodes.shape[0]
        th.close()
    return out


def bench_cfg4(pkg, torch, reps=5) -> dict:
    """256 seeded parameter sets, 128x128x32 each, one persistent launch."""
    import decks
    sets = decks.cfg4_sets()
    beams = np.array([[q["beam"]["Power"], q["beam"]["Efficiency"], q["beam"]["Width_X"], q["beam"]["Width_Y"],
                       q["beam"]["Depth_Z"]] for q in sets])
    paths = [np.array(q["path"], dtype=np.float64) for q in sets]
    with tempfile.TemporaryDirectory() as d:
        th = pkg.Thesis(decks.cfg4_deck(0).write(d))
        t0 = time.perf_counter()
        th.batch_set(beams, paths)
        build_s = time.perf_counter() - t0
        nodes, evals = th.batch_info()
        pnum = th.info.pnum
        T = torch.empty(len(sets) * pnum, dtype=torch.float64, device="cuda")
        ppt, blocks = th.batch_autotune(T.data_ptr(), pkg.FP32C)
        out = {"sets": len(sets), "points_per_set": pnum, "nodes": nodes, "evals": evals,
               "node_build_s": build_s, "tuned": {"ppt": ppt, "blocks": blocks}}
        s = torch.cuda.Stream()
        for name, prec, n in (("fp32c", pkg.FP32C, reps), ("fp64", pkg.FP64, 1)):
            th.batch_eval_dev(T.data_ptr(), prec, s.cuda_stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(n):
                th.batch_eval_dev(T.data_ptr(), prec, s.cuda_stream)
            e1.record(s)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / n
            out[name] = {"evals_per_s": evals / (ms * 1e-3), "ms": ms}
        out["roofline_fp32c"] = {"achieved_tflops": out["fp32c"]["evals_per_s"] * FLOP_PER_EVAL / 1e12,
                                 "frac_of_fp32_peak": out["fp32c"]["evals_per_s"] * FLOP_PER_EVAL / 1e12
                                 / FP32_PEAK_TFLOPS}
        th.close()
    return out


def bench_cfg3(pkg, torch) -> dict:
    """Multi-beam Volume solidification, 4 beams over a 500x500x50 grid: one
    untimed run, then one timed run (outputs off)."""
    import decks
    with tempfile.TemporaryDirectory() as d:
        th = pkg.Thesis(decks.cfg3().write(d), precision=pkg.FP32C)
        th.set_option("outputs", 0)
        run_once(th)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = run_once(th)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        th.close()
    return {"s_per_run": secs, "evals": r["evals"], "evals_per_s": r["evals"] / secs, "timesteps": r["steps"],
            "switched_to_per_step": bool(r["window"]["per_step_switch"])}


def bench_cfg5(pkg, torch, reps=3) -> dict:
    """Adjoint: 256x256x64 field, 200 perturbed vertices, MSE loss vs a target
    made from the true parameters; one gradient = FP32c forward + adjoint."""
    import decks
    autograd = importlib.import_module("3dthesis_amd.autograd")
    with tempfile.TemporaryDirectory() as d:
        th = pkg.Thesis(decks.cfg5().write(d))
        beam_t, rows_t = th.design()
        pnum = th.info.pnum
        Tt = torch.empty(pnum, dtype=torch.float64, device="cuda")
        th.forward_dev(Tt.data_ptr(), pkg.FP64)
        torch.cuda.synchronize()
        rng = np.random.default_rng(5001)
        beam0 = beam_t * (1 + rng.uniform(-0.1, 0.1, 5))
        rows0 = rows_t.copy()
        rows0[:, 1:3] += rng.normal(0.0, 0.010, size=(rows0.shape[0], 2))
        field = autograd.SnapshotField(th, precision=pkg.FP32C, rows_template=rows0)

        def grad_once():
            b = torch.tensor(beam0, dtype=torch.float64, requires_grad=True)
            xy = torch.tensor(rows0[:, 1:3], dtype=torch.float64, requires_grad=True)
            loss = ((field(b, xy) - Tt) ** 2).mean()
            loss.backward()
            return loss, b.grad, xy.grad

        grad_once()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            grad_once()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / reps
        ncomp = 5 + 2 * rows0.shape[0]
        nodes = th.design_nodes().shape[0]
        th.close()
    return {"grads_per_s": 1.0 / sec, "components": ncomp, "ms_per_grad": sec * 1e3,
            "forward_evals": pnum * nodes, "points": pnum, "nodes": nodes}


def run_once(th) -> dict:
    st = th.run()
    return dict(evals=st.evals, wall=st.wall_s, steps=st.steps, cat_evals=list(st.cat_evals),
                cat_ms=list(st.cat_ms), cat_launches=list(st.cat_launches), evals_window=st.evals_window,
                window=th.window_stats())


def load_traffic(kernels, workload: str):
    """HBM bytes per launch of a kernel group (the kernels' names without
    template arguments) from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE in KiB per dispatch, scaled by the copy-kernel
    calibration factors recorded beside them), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    total = 0.0
    for kernel in kernels:
        k = next((v for name, v in d.get("kernels", {}).items() if name.split("<")[0] == kernel), None)
        if not k:
            return None
        total += (k["fetch_kib"] * d.get("fetch_factor", 1.0) + k["write_kib"] * d.get("write_factor", 1.0)) * 1024.0
    return total


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Host threads given to the CPU baseline: the CPU share of one GPU on the
# one-GPU boxes this bench runs on is 16 cores, whatever os.cpu_count() reports
# for the whole machine.
CPU_SHARE = 16


def cpu_baseline(name: str, pkg) -> dict:
    """The reference's OpenMP build (oracle/_ref) timed on the host on the
    workload's deck, with the evaluation count of the same deck from the GPU
    FP64 run (equal to the reference's own count: a GPU test holds the two
    equal on the golden decks); beside it the FP64 C restatement (oracle/,
    single-threaded) on a 4-hatch slice of the deck with its own count."""
    import decks
    import oracle_ffi
    deck, what = cpu_sample_deck(name)
    cores = min(os.cpu_count() or 1, CPU_SHARE)
    out = {"unit": "point-node evals/s", "cores": cores, "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    with tempfile.TemporaryDirectory() as d:
        deck.settings = {"Compute": {"MaxThreads": cores}}
        pi = deck.write(d)
        th = pkg.Thesis(pi, precision=pkg.FP64)
        th.set_option("outputs", 0)
        evals = th.run().evals
        th.close()
        ref = os.path.join(ROOT, "oracle", "_ref", "thesis_ref")
        if os.path.exists(ref):
            r = subprocess.run([ref, "run", pi], capture_output=True, text=True, check=True,
                               env=dict(os.environ, OMP_PROC_BIND="close"))
            secs = float(re.search(r"Execution time \(s\): ([0-9.eE+-]+)", r.stdout).group(1))
            out.update(value=evals / secs, kind="reference", seconds=secs,
                       sample=f"{name}: {what}; {evals:.4g} evals in {secs:.2f} s on {cores} threads")
    # the restatement on a slice (4 of the 20 hatches), one thread
    sl = decks.cfg2(nh=4) if name == "cfg2" else deck
    with tempfile.TemporaryDirectory() as d:
        sl.settings = {"Compute": {"MaxThreads": 1}}
        sim = oracle_ffi.Sim(sl.write(d))
        t0 = time.perf_counter()
        ev = sim.run(None)
        secs = time.perf_counter() - t0
    rest = {"value": ev / secs, "cores": 1, "seconds": secs, "evals": ev,
            "sample": f"{name} with 4 hatches: {ev:.4g} evals in {secs:.2f} s"}
    if "value" not in out:
        out.update(value=rest["value"], kind="port", cores=1, seconds=secs, sample=rest["sample"])
    else:
        out["restatement"] = rest
    return out


def read_full_dump(path: str):
    """A "<p> v1 v2 ..." full-precision output dump -> (p, values)."""
    a = np.loadtxt(path, ndmin=2)
    if a.size == 0:
        return np.zeros(0, np.int64), np.zeros((0, 0))
    return a[:, 0].astype(np.int64), a[:, 1:]


def accuracy(pkg, pi: str, deck) -> dict:
    """FP32c against FP64 on the workload itself, outside the timed region:
    (1) the whole run's outputs of both precisions (full-precision dumps):
    melted row sets, tSol in units of dt, G and V relative; (2) T of the FP32c
    node sums against the FP64 ones over the whole grid at eight times of the
    scan, with the node sets the run uses."""
    out = {}
    dumps = {}
    with tempfile.TemporaryDirectory() as d:
        for name, prec in (("fp32c", pkg.FP32C), ("fp64", pkg.FP64)):
            fd = os.path.join(d, name)
            os.makedirs(fd)
            th = pkg.Thesis(pi, precision=prec)
            th.set_full_dump(fd)
            th.run()
            th.close()
            fulls = sorted(f for f in os.listdir(fd) if f.endswith(".full.txt"))
            dumps[name] = read_full_dump(os.path.join(fd, fulls[-1]))
    data = os.path.join(os.path.dirname(pi), "Data")
    csvs = sorted(f for f in os.listdir(data) if f.endswith(".csv"))
    cols = open(os.path.join(data, csvs[-1])).readline().strip().split(",")
    (p32, v32), (p64, v64) = dumps["fp32c"], dumps["fp64"]
    common, i32, i64 = np.intersect1d(p32, p64, return_indices=True)
    out["rows_fp64"] = int(len(p64))
    out["rows_fp32c"] = int(len(p32))
    out["rows_only_in_one"] = int(len(p32) + len(p64) - 2 * len(common))
    dt = float(deck.mode.get("Timestep", 1e-5))
    a, b = v32[i32], v64[i64]
    for c in ("tSol", "G", "V", "T"):
        if c in cols:
            k = cols.index(c)
            if c == "tSol":
                e = np.abs(a[:, k] - b[:, k]) / dt
                out["tSol_max_abs_err_dt"] = float(e.max()) if e.size else 0.0
                out["tSol_frac_within_1dt"] = float(np.mean(e <= 1.0)) if e.size else 1.0
            else:
                e = np.abs(a[:, k] - b[:, k]) / np.maximum(np.abs(b[:, k]), 1e-300)
                out[f"{c}_rel_err_max"] = float(e.max()) if e.size else 0.0
                out[f"{c}_rel_err_p99"] = float(np.quantile(e, 0.99)) if e.size else 0.0
    th = pkg.Thesis(pi)
    T_end = th.info.scan_end
    worst = 0.0
    worst_hot = 0.0
    for f in np.linspace(0.05, 1.0, 8):
        nodes = th.build_nodes(f * T_end)
        T64 = th.eval_T(nodes, pkg.FP64)
        T32 = th.eval_T(nodes, pkg.FP32C)
        rel = np.abs(T32 - T64) / T64
        worst = max(worst, float(rel.max()))
        hot = T64 >= th.info.T_liq
        if hot.any():
            worst_hot = max(worst_hot, float(rel[hot].max()))
    th.close()
    out["T_rel_err_max_grid"] = worst
    out["T_rel_err_max_liquid"] = worst_hot
    out["T_points_x_times"] = f"{th.info.pnum} points x 8 times"
    out["bar"] = "T rel err <= 1e-5; tSol within 1 dt"
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the cfg1/cfg3/cfg4/cfg5 side measurements")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison runs")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP32c-vs-FP64 output comparison")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is "
              "a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    import torch
    pkg = importlib.import_module("3dthesis_amd")
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_mib:  # an elementwise kernel that reads and writes copy_mib MiB once each
        src = torch.empty(args.copy_mib << 20, dtype=torch.uint8, device="cuda")
        src.fill_(1)
        dst = src.mul(3)
        torch.cuda.synchronize()
        del src, dst
    deck, desc = workload_deck(args.workload)
    with tempfile.TemporaryDirectory() as d:
        pi = deck.write(d)
        th = pkg.Thesis(pi, precision=pkg.FP32C)
        th.set_option("outputs", 0)
        surface = th.info.tracking == 2
        log(f"bench: {args.workload} FP32c warmup x{args.warmup}")
        for _ in range(args.warmup):
            run_once(th)
        torch.cuda.synchronize()
        log(f"bench: {args.workload} FP32c timed x{args.steps}")
        t0 = time.perf_counter()
        res = [run_once(th) for _ in range(args.steps)]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        evals = res[-1]["evals"]
        value = evals * args.steps / elapsed
        # instrumented pass: per-category kernel time from HIP events
        th.set_option("timing", 1)
        log("bench: FP32c instrumented pass")
        inst = run_once(th)
        th.set_option("timing", 0)
        # the tracker's phase clocks in a pass of their own (each reading waits)
        th.set_option("phase_times", 1)
        log("bench: FP32c tracker phase pass")
        inst["window"].update({k: v for k, v in run_once(th)["window"].items() if k.startswith("t_")})
        th.set_option("phase_times", 0)
        th.close()

        if args.no_fp64:
            r64, secs64, inst64 = res[-1], elapsed / args.steps, inst
        else:
            log("bench: FP64 pass")
            th64 = pkg.Thesis(pi, precision=pkg.FP64)
            th64.set_option("outputs", 0)
            run_once(th64)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r64 = run_once(th64)
            torch.cuda.synchronize()
            secs64 = time.perf_counter() - t1
            th64.set_option("timing", 1)
            inst64 = run_once(th64)
            th64.close()

    names = ["box sums", "solidification (k_solidify)", "tracking (k_track_window, k_window_apply)"]
    cat = int(np.argmax(inst["cat_ms"]))
    # Per-kernel HIP-event times of the instrumented pass, each on the stream
    # its kernel is launched on. A window's box sums are one launch of
    # k_window_eval for Volume tracking and three for Surface (the surface
    # layer, the column list, the deep layers of the listed columns); their
    # events bracket the whole group.
    box_members = ["k_surf_eval_f32c", "k_list_columns", "k_deep_eval_sep_f32c"] if surface else ["k_window_eval_f32c"]
    # the window tracker (Surface: the surface-layer tracker; its column walks
    # run beside the next window's tracker, before the replay)
    track_members = ["k_track_surface"] if surface else ["k_track_window_lds"]
    members = {"box_sums": box_members, "tracker": track_members, "k_window_apply": ["k_window_apply"],
               "k_solidify": ["k_solidify"], "columns": ["k_surface_columns"]}
    kernels = [("box_sums", "box_sums"), ("tracker", "tracker"), ("k_window_apply", "replay"),
               ("k_solidify", "solidify")] + ([("columns", "columns")] if surface else [])

    def kern_table(r, fp64=False):
        w = r["window"]
        t = {k: {"ms": w[f"ms_{f}"], "launches": w[f"launches_{f}"],
                 "avg_launch_us": w[f"ms_{f}"] / max(w[f"launches_{f}"], 1.0) * 1e3} for k, f in kernels}
        for k in t:
            names = ", ".join(members[k])
            t[k]["kernels"] = names.replace("_sep_f32c", "_f64").replace("_f32c", "_f64") if fp64 else names
        return t

    ktab = kern_table(inst)
    dom = max(ktab, key=lambda k: ktab[k]["ms"])
    # Algorithmic work: the evaluations the windowed steps stand for, counted
    # as the reference counts them (each step's evaluated points times its node
    # count, 14 FLOP and one exp each), spread over the launches that decide or
    # produce them. The box sums evaluate a superset of those points.
    evals_alg = inst["window"]["evals_windowed"]

    def roof(kname, r, peak):
        t = kern_table(r)[kname]
        if not t["ms"]:
            return None
        ach = evals_alg * FLOP_PER_EVAL / (t["ms"] * 1e-3) / 1e12
        return {"kernel": kname, "kernels": t["kernels"], "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                "frac": ach / peak,
                "launches": t["launches"], "avg_launch_us": t["avg_launch_us"],
                "evals_per_launch": evals_alg / max(t["launches"], 1.0)}

    rd = roof(dom, inst, FP32_PEAK_TFLOPS)
    rbox = roof("box_sums", inst, FP32_PEAK_TFLOPS)
    traffic = load_traffic(members[dom], args.workload)
    ach64 = (inst64["window"]["evals_windowed"] * FLOP_PER_EVAL / (inst64["window"]["ms_box_sums"] * 1e-3) / 1e12
             if inst64["window"]["ms_box_sums"] else 0.0)
    whole = evals * FLOP_PER_EVAL / (elapsed / args.steps) / 1e12
    out = {
        "metric": "point-node evals/s (FP32-compensated)",
        "value": value,
        "unit": "point-node evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (exp, Kahan) with f64 offsets and times",
        "data": "synthetic input decks generated from the workload parameters",
        "config": {"workload": args.workload, "description": desc, "evals_per_step": evals,
                   "timesteps_per_step": res[-1]["steps"]},
        "fp64": {"value": r64["evals"] / secs64, "ms_per_step": secs64 * 1e3,
                 "speedup_fp32c_over_fp64": value / (r64["evals"] / secs64)},
        "roofline": {"bound": "valu", **rd,
                     "traffic": traffic,  # HBM bytes per launch (rocprofv3 PMC), None without the summary
                     "flop_per_eval": FLOP_PER_EVAL,
                     "basis": "evaluations of the windowed steps as the reference counts them x 14 FLOP / "
                              "the kernel's HIP-event time",
                     "box_sums": rbox,
                     "superset_ratio": inst["evals_window"] / evals_alg if evals_alg else None,
                     "whole_run_achieved": whole, "whole_run_frac": whole / FP32_PEAK_TFLOPS,
                     "fp64_kernel": "box sums, FP64", "fp64_achieved": ach64,
                     "fp64_frac_of_fp64_peak": ach64 / FP64_PEAK_TFLOPS},
        "kernels": ktab,
        "kernels_fp64": kern_table(inst64, fp64=True),
        "kernel_ms_by_category": dict(zip(names, inst["cat_ms"])),
        "dominant_category": names[cat],
        "window_stats": inst["window"],
    }
    if not args.no_accuracy:
        log("bench: FP32c vs FP64 accuracy on the workload")
        with tempfile.TemporaryDirectory() as d:
            out["accuracy"] = accuracy(pkg, deck.write(d), deck)
    if not args.no_extra:
        log("bench: cfg1 snapshot sums")
        out["cfg1"] = bench_cfg1(pkg, torch)
        log("bench: cfg4 batched snapshots")
        out["cfg4"] = bench_cfg4(pkg, torch)
        log("bench: cfg5 adjoint")
        out["cfg5"] = bench_cfg5(pkg, torch)
        log("bench: cfg3 multi-beam Volume solidification")
        out["cfg3"] = bench_cfg3(pkg, torch)
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.workload, pkg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
