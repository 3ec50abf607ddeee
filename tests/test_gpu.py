"""GPU parity: the HIP path through the C-ABI against the FP64 CPU oracle and the
reference's golden outputs.

Tolerances: FP64 kernels reproduce the reference's arithmetic except for the
device exp/log (<= 1 ulp), so T agrees to 1e-12 relative; FP32-compensated
kernels are held to the 1e-5 relative-T bar. Solidification times are held to
one timestep and, in FP64, to 1e-9 relative.
"""
from __future__ import annotations

import gzip
import importlib
import os
import sys

import numpy as np
import pytest

import decks
import oracle_ffi as of
from make_golden import CLANG_RUN_CASES, RUN_CASES, scan_end

pkg = importlib.import_module("3dthesis_amd")
pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg1_full(tmp_path_factory):
    d = tmp_path_factory.mktemp("cfg1")
    pi = decks.cfg1().write(str(d))
    sim = of.Sim(pi)
    t = float(sim.s.all_end)
    T_ref, _ = sim.snapshot_T(t, nthreads=8)
    return pi, t, T_ref


def test_fp64_grid_matches_oracle(cfg1_full):
    pi, t, T_ref = cfg1_full
    _flush("case: cfg1 full grid, FP64")
    th = pkg.Thesis(pi)
    nodes = th.build_nodes(t)
    T = th.eval_T(nodes, pkg.FP64)
    rel = np.abs(T - T_ref) / T_ref
    assert rel.max() <= 1e-12, rel.max()


@pytest.mark.parametrize("ppt", [1, 2, 4])
def test_fp32c_grid_within_1e5(cfg1_full, ppt):
    pi, t, T_ref = cfg1_full
    _flush(f"case: cfg1 full grid, FP32c ppt={ppt}")
    th = pkg.Thesis(pi)
    th.set_option("ppt", ppt)
    nodes = th.build_nodes(t)
    T = th.eval_T(nodes, pkg.FP32C)
    rel = np.abs(T - T_ref) / T_ref
    assert rel.max() <= 1e-5, rel.max()


def test_grid_subrange_and_golden(golden, tmp_path):
    for case in ("cfg1_small", "snapv"):
        m = golden["meta"]["tgrid"][case]
        _flush(f"case: golden T grid {case}")
        th = pkg.Thesis(decks.WORKLOADS[m["workload"]]().write(str(tmp_path / case)))
        nodes = th.build_nodes(m["t"])
        ref = golden["tgrid"][case]
        T = th.eval_T(nodes, pkg.FP64)
        assert np.max(np.abs(T - ref) / ref) <= 1e-12
        # ragged sub-range
        p0, cnt = 7, min(1000, ref.size - 7)
        assert np.max(np.abs(th.eval_T(nodes, pkg.FP64, p0, cnt) - ref[p0:p0 + cnt]) / ref[p0:p0 + cnt]) <= 1e-12
        assert np.max(np.abs(th.eval_T(nodes, pkg.FP32C, p0, cnt) - ref[p0:p0 + cnt]) / ref[p0:p0 + cnt]) <= 1e-5


def test_point_lists(cfg1_full):
    pi, t, T_ref = cfg1_full
    _flush("case: cfg1 point lists")
    th = pkg.Thesis(pi)
    nodes = th.build_nodes(t)
    rng = np.random.default_rng(7)
    pts = rng.integers(0, T_ref.size, size=4097).astype(np.int32)
    for prec, tol in ((pkg.FP64, 1e-12), (pkg.FP32C, 1e-5)):
        T = th.eval_T_points(nodes, pts, prec)
        assert np.max(np.abs(T - T_ref[pts]) / T_ref[pts]) <= tol
    # empty list and empty node set
    assert th.eval_T_points(nodes, np.zeros(0, np.int32)).size == 0
    T0 = th.eval_T(nodes[:0], pkg.FP64, 0, 10)
    assert np.all(T0 == decks.MATERIAL["T_0"])


def _compare_run(wl, tag, fname, golden, d, prec, deck, group="runs"):
    runs = golden[group]
    ref_csv = gzip.decompress(runs[f"{wl}/{tag}/csv"].tobytes()).decode().splitlines()
    got_csv = open(os.path.join(d, "Data", fname)).read().splitlines()
    assert got_csv[0] == ref_csv[0]  # header / column order
    return ref_csv, got_csv


def _run_fp64_vs_reference(wl, golden, tmp_path, window=None, tracker_lds=True, opts=(), group="runs", abs_mode=0):
    """Run a deck in FP64 and compare its outputs with the reference's (group
    "runs": the g++ build; "runs_clang" with abs_mode 1: the clang++ build);
    returns the run's window diagnostics."""
    deck = decks.WORKLOADS[wl]()
    d = str(tmp_path / wl)
    pi = deck.write(d)
    th = pkg.Thesis(pi, precision=pkg.FP64, abs_mode=abs_mode)
    if window is not None:
        th.set_option("window", window)
    if not tracker_lds:
        th.set_option("tracker_lds", 0)
    for k, v in opts:
        th.set_option(k, v)
    os.makedirs(os.path.join(d, "full"), exist_ok=True)
    th.set_full_dump(os.path.join(d, "full"))
    st = th.run()
    assert st.evals > 0
    ws = th.window_stats()
    meta = golden["meta"][group][wl]
    for tag, fname in meta["files"].items():
        ref_csv, got_csv = _compare_run(wl, tag, fname, golden, d, pkg.FP64, deck, group)
        assert len(got_csv) == len(ref_csv), fname
        a = np.array([[float(v) for v in l.split(",")] for l in ref_csv[1:]]) if len(ref_csv) > 1 else np.zeros((0, 1))
        b = np.array([[float(v) for v in l.split(",")] for l in got_csv[1:]]) if len(got_csv) > 1 else np.zeros((0, 1))
        # 6 printed digits; values below 1e-280 are the stack residue that the
        # reference's uninitialised zb of merged compressed nodes (Calc.cpp:452)
        # leaves in Gz at surface points, where this library has 0
        assert np.allclose(a, b, rtol=2e-6, atol=1e-280), fname
    if "rrdf_sha256" in meta:
        # Stork writes records, no grid table; their full-precision text is
        # pinned by digest for the oracle, which is run here as the checker
        od = tmp_path / "oracle"
        sim = of.Sim(deck.write(str(od)))
        os.makedirs(od / "full", exist_ok=True)
        sim.run(str(od / "full"))
        ref = np.loadtxt(od / "full" / f"{deck.name}.RRDF.full.txt", ndmin=2)
        got = np.loadtxt(os.path.join(d, "full", f"{deck.name}.RRDF.full.txt"), ndmin=2)
        assert got.shape == ref.shape and ref.shape[0] == meta["rrdf_records"]
        assert np.array_equal(got[:, :5], ref[:, :5])
        assert np.max(np.abs(got[:, 5:] - ref[:, 5:]) / ref[:, 5:]) <= 1e-12
        return ws
    runs = golden[group]
    p_ref, v_ref = runs[f"{wl}/final/p"], runs[f"{wl}/final/v"]
    fulls = [f for f in sorted(os.listdir(os.path.join(d, "full")))
             if not any(k in f for k in (".T_hist.", ".RDF.", ".RRDF."))]
    p, v = of.read_full(os.path.join(d, "full", fulls[-1]))
    assert np.array_equal(p, p_ref)
    rel = np.abs(v - v_ref) / np.maximum(np.abs(v_ref), 1e-280)
    assert np.nanmax(rel) <= 1e-9, np.nanmax(rel, axis=0)
    if f"{wl}/rdf/p" in runs:  # every RDF melt event: (tm, tl, cr)
        rp, rv = of.read_full(os.path.join(d, "full", f"{deck.name}.RDF.full.txt"))
        assert np.array_equal(rp, runs[f"{wl}/rdf/p"])
        ref = runs[f"{wl}/rdf/v"]
        if len(rp):  # Volume tracking records no melt times: no events then
            assert np.array_equal(rv[:, 0], ref[:, 0])  # melt times are step times
            assert np.max(np.abs(rv[:, 1:] - ref[:, 1:]) / np.maximum(np.abs(ref[:, 1:]), 1e-280)) <= 1e-9
    if f"{wl}/thist/p" in runs:  # every recorded (t, T) of the T_hist output
        hp, hv = of.read_full(os.path.join(d, "full", f"{deck.name}.T_hist.full.txt"))
        assert np.array_equal(hp, runs[f"{wl}/thist/p"])
        ref = runs[f"{wl}/thist/v"]
        assert np.array_equal(hv[:, 0], ref[:, 0])
        assert np.max(np.abs(hv[:, 1] - ref[:, 1]) / ref[:, 1]) <= 1e-12
    return ws


@pytest.mark.parametrize("wl", RUN_CASES)
def test_mode_run_fp64_matches_reference(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64")
    _run_fp64_vs_reference(wl, golden, tmp_path)


# Tracked Solidification runs in windows of steps by default; one step per
# window, a short odd window and the per-step launch sequence (window 0) must
# give the same outputs.
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg2_small_volume", "cfg3_small", "perim_small", "perim_small_volume",
                                "thist_surface_small", "thist_volume_small"])
@pytest.mark.parametrize("window", [0, 1, 7])
def test_tracked_run_window_sizes(wl, window, golden, tmp_path):
    _flush(f"case: run {wl} FP64 window={window}")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, window)
    assert ws["per_step_switch"] == 0  # these melt pools fit the LDS tracker


# The clang++ build of the reference (its abs() at Grid.cpp:381/:535 is the
# floating one: the regula falsi iterates to Sol_Tol and dTdt is not
# truncated) against abs_mode 1, windowed and per step.
@pytest.mark.parametrize("wl", CLANG_RUN_CASES)
@pytest.mark.parametrize("window", [None, 0])
def test_fabs_mode_run_matches_clang_reference(wl, window, golden, tmp_path):
    _flush(f"case: run {wl} FP64 abs_mode=1 window={window}")
    _run_fp64_vs_reference(wl, golden, tmp_path, window, group="runs_clang", abs_mode=1)


# Windows that stop at a point outside their boxes resume from that step with
# wider boxes: a zero base margin makes that happen, with the same outputs.
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg2_small_volume", "perim_small"])
def test_window_misses_resume_with_the_same_outputs(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64 box_margin=0")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, opts=[("box_margin", 0)])
    assert ws["misses"] > 0 and ws["windows"] > ws["misses"]


# A column walk that leaves the boxes ends the window at the step before it;
# no margin below the deepest melt makes that happen, with the same outputs.
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg2_small_volume", "perim_small", "thist_surface_small"])
def test_window_depth_misses_resume_with_the_same_outputs(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64 depth_margin=0")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, opts=[("depth_margin", 0)])
    assert ws["misses"] > 0 and ws["windows"] > ws["misses"]


# Fewer windows queued on the device at once, with windows that stop early
# (a zero base margin): the host plans each window from the last it has
# processed, and a window that stops early makes the ones behind it skip.
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg2_small_volume", "perim_small"])
@pytest.mark.parametrize("pipeline", [1, 2])
def test_window_pipeline_depths(wl, pipeline, golden, tmp_path):
    _flush(f"case: run {wl} FP64 pipeline={pipeline} box_margin=0")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, opts=[("pipeline", pipeline), ("box_margin", 0)])
    assert ws["misses"] > 0


# The Surface tracker at each of its workgroup sizes (by default the number
# of surface words of the window picks one).
@pytest.mark.parametrize("wl", ["cfg2_small", "perim_small", "thist_surface_small", "rdf_small"])
@pytest.mark.parametrize("threads", [64, 256, 1024])
def test_surface_tracker_workgroup_sizes(wl, threads, golden, tmp_path):
    _flush(f"case: run {wl} FP64 surface_threads={threads}")
    _run_fp64_vs_reference(wl, golden, tmp_path, opts=[("surface_threads", threads)])


# The Surface tracker without the window's seed masks in LDS (seeds claimed
# with atomics step by step), and without its surface rows there either.
@pytest.mark.parametrize("wl", ["cfg2_small", "perim_small", "rdf_small"])
@pytest.mark.parametrize("rows", [0, 1])
def test_surface_tracker_lds_modes(wl, rows, golden, tmp_path):
    _flush(f"case: run {wl} FP64 surface_rows={rows}")
    _run_fp64_vs_reference(wl, golden, tmp_path, opts=[("surface_rows", rows)])


# Path compression and wall images through the per-step launch sequence too.
@pytest.mark.parametrize("wl", ["compress_small_surface", "walls_small_volume"])
def test_compressed_and_walled_runs_per_step(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64 window=0")
    _run_fp64_vs_reference(wl, golden, tmp_path, 0)


# The tracker's device-memory form (used when a window's state outgrows LDS).
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg3_small"])
def test_tracked_run_tracker_in_device_memory(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64 tracker state in device memory")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, 5, tracker_lds=False, opts=[("volume_switch", 0)])
    assert ws["per_step_switch"] == 0 and ws["lds_windows"] == 0 and ws["steps"] > 0


# A Volume run whose windows keep outgrowing the LDS tracker starts again with
# the per-step launch sequence; the restart must leave no trace of the windows
# it abandoned (point state, queue, event logs, step tables).
@pytest.mark.parametrize("wl", ["cfg3_small", "walls_small_volume"])
def test_volume_switch_to_per_step(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP64 switched to per-step after 2 device-memory windows")
    ws = _run_fp64_vs_reference(wl, golden, tmp_path, 5, tracker_lds=False, opts=[("volume_switch", 2)])
    assert ws["per_step_switch"] == 1 and ws["windows"] >= 2


@pytest.mark.parametrize("wl", ["cfg2_small", "cfg3_small", "none_small", "snapvol_small", "compress_small_surface",
                                "walls_small_volume", "points_small_sol"])
def test_mode_run_fp32c_within_tolerance(wl, golden, tmp_path):
    _flush(f"case: run {wl} FP32c")
    deck = decks.WORKLOADS[wl]()
    d = str(tmp_path / wl)
    pi = deck.write(d)
    th = pkg.Thesis(pi, precision=pkg.FP32C)
    os.makedirs(os.path.join(d, "full"), exist_ok=True)
    th.set_full_dump(os.path.join(d, "full"))
    th.run()
    p_ref, v_ref = golden["runs"][f"{wl}/final/p"], golden["runs"][f"{wl}/final/v"]
    fulls = [f for f in sorted(os.listdir(os.path.join(d, "full")))
             if not any(k in f for k in (".T_hist.", ".RDF.", ".RRDF."))]
    p, v = of.read_full(os.path.join(d, "full", fulls[-1]))
    common, ia, ib = np.intersect1d(p_ref, p, return_indices=True)
    # melted sets may differ only by points within the FP32 tolerance of the liquidus
    assert len(common) >= 0.995 * max(len(p_ref), len(p))
    cols = [c.strip() for c in gzip.decompress(golden["runs"][f"{wl}/{list(golden['meta']['runs'][wl]['files'])[-1]}/csv"]
                                                .tobytes()).decode().splitlines()[0].split(",")]
    a, b = v_ref[ia], v[ib]
    if "T" in cols:
        c = cols.index("T")
        assert np.max(np.abs(a[:, c] - b[:, c]) / a[:, c]) <= 1e-5
    if "tSol" in cols:
        c = cols.index("tSol")
        dt = deck.mode.get("Timestep", 1e-5)
        assert np.max(np.abs(a[:, c] - b[:, c])) <= dt


# The metric's numerator: the point-node evaluations of a run, counted as the
# reference counts them (every Calc_T over the node list it is given, including
# the solidification searches and refined builds, Grid.cpp:275-404), equal to
# the oracle's own count on the same deck.
@pytest.mark.parametrize("wl,window", [("none_small", None), ("snapvol_small", None), ("cfg2_small", None),
                                       ("cfg2_small", 0), ("cfg2_small_volume", None), ("cfg2_small_volume", 0),
                                       ("cfg3_small", None), ("compress_small_surface", None),
                                       ("walls_small_volume", 0), ("rdf_small", None)])
def test_run_eval_count_matches_oracle(wl, window, tmp_path):
    _flush(f"case: eval count {wl} window={window}")
    deck = decks.WORKLOADS[wl]()
    th = pkg.Thesis(deck.write(str(tmp_path / "gpu")), precision=pkg.FP64)
    th.set_option("outputs", 0)
    if window is not None:
        th.set_option("window", window)
    st = th.run()
    th.close()
    ev = of.Sim(deck.write(str(tmp_path / "oracle"))).run(None)
    assert st.evals == ev, (st.evals, ev)


def test_eval_sol_matches_oracle(tmp_path):
    """Grid::Solidify's seam (Grid.h:291) at points crossing the liquidus
    between t - dt and t: the g++ build's one regula-falsi step
    (Grid.cpp:340-404), the refined build at t_sol (Grid.cpp:325) and the
    primary quantities (Grid.cpp:406-453, :522-548), from the oracle's node
    builder and gradient sums."""
    deck = decks.cfg2_small()
    pi = deck.write(str(tmp_path / "d"))
    sim = of.Sim(pi)
    dt = float(deck.mode["Timestep"])
    T_L = decks.MATERIAL["T_L"]
    alpha = decks.MATERIAL["k"] / (decks.MATERIAL["p"] * decks.MATERIAL["c"])
    t = 163 * dt
    T1, _ = sim.snapshot_T(t - dt)
    T2, _ = sim.snapshot_T(t)
    pts = np.nonzero((T1 >= T_L) & (T2 < T_L))[0][:48].astype(np.int32)
    _flush(f"case: eval_sol at {len(pts)} points of cfg2_small, t={t}")
    assert len(pts) >= 8
    th = pkg.Thesis(pi)
    out = th.eval_sol(t, pts, T2[pts], T1[pts])
    th.close()
    for q, p in enumerate(pts):
        t1, t2 = t - dt, t
        m = (T2[p] - T1[p]) / (t2 - t1)
        t0 = t1 + (T_L - T1[p]) / m
        x, y, z = sim.xyz(int(p))
        gx, gy, gz, lap, S = of.sol_primary(sim.nodes(t0, True), x, y, z)
        G = np.sqrt(gx * gx + gy * gy + gz * gz)
        dTdt = float(abs(int(alpha * lap + S)))  # int abs of the g++ build (Grid.cpp:535)
        want = np.array([t0, G, gx / G, gy / G, gz / G, dTdt, dTdt / G])
        got = out[q]
        assert abs(got[0] - t0) <= 1e-12 * t0, (p, got[0], t0)
        assert np.all(np.abs(got[[1, 2, 3, 4]] - want[[1, 2, 3, 4]]) <= 1e-9 * np.maximum(np.abs(want[[1, 2, 3, 4]]), G * 1e-6))
        assert abs(got[5] - dTdt) <= 1.0  # integer truncation of a value known to ~1e-9
        assert abs(got[6] - want[6]) <= 1e-9 * want[6] + 1.0 / G


def test_full_cfg2_matches_reference(golden, tmp_path):
    """The benchmarked workload at full size (400x400x40, 4244 steps) against
    the reference's own run of it (MaxThreads 1): FP64 gives the reference's
    melted row set (count and digest) and, on 1000 sampled rows, its values to
    1e-9; FP32c keeps all but 0.1% of the rows and every common row's tSol
    within one timestep."""
    from conftest import check_example_outputs
    deck = decks.cfg2()
    d = str(tmp_path / "cfg2")
    pi = deck.write(d)
    outs = {}
    for name, prec in (("fp64", pkg.FP64), ("fp32c", pkg.FP32C)):
        _flush(f"case: full cfg2 run, {name}")
        fd = os.path.join(d, "full_" + name)
        os.makedirs(fd, exist_ok=True)
        th = pkg.Thesis(pi, precision=prec)
        th.set_full_dump(fd)
        th.run()
        th.close()
        if prec == pkg.FP64:
            check_example_outputs("cfg2", golden, os.path.join(d, "Data"), fd, exact=False, group="full")
        outs[name] = of.read_full(os.path.join(fd, "cfg2.Solidification.Final.full.txt"))
    (p64, v64), (p32, v32) = outs["fp64"], outs["fp32c"]
    common, i64, i32 = np.intersect1d(p64, p32, return_indices=True)
    assert len(common) >= 0.999 * max(len(p64), len(p32)), (len(common), len(p64), len(p32))
    cols = open(os.path.join(d, "Data", "cfg2.Solidification.Final.csv")).readline().strip().split(",")
    c = cols.index("tSol")
    assert np.max(np.abs(v64[i64, c] - v32[i32, c])) <= float(deck.mode["Timestep"])


@pytest.mark.parametrize("opts", [(("pipeline", 4), ("window", 75)),
                                  (("pipeline", 4), ("depth_margin", 0), ("box_margin", 0)),
                                  (("pipeline", 2), ("window", 30), ("sol_stream", 0)),
                                  (("pipeline", 6), ("window", 30))])
def test_full_cfg2_window_pipelines_match_reference(opts, golden, tmp_path):
    """The full cfg2 workload in FP64 with other window lengths, pipeline
    depths, margins and stream placements: the reference's row set and values,
    and the same evaluation count as the default settings. Windows that end
    early while others are queued behind them send the driver back to the
    last processed window's liquid list, which no queued window may
    overwrite."""
    from conftest import check_example_outputs
    deck = decks.cfg2()
    d = str(tmp_path / "cfg2")
    pi = deck.write(d)
    evals = []
    for o in ((), opts):
        _flush(f"case: full cfg2 run, FP64, options {o}")
        fd = os.path.join(d, "full")
        os.makedirs(fd, exist_ok=True)
        th = pkg.Thesis(pi, precision=pkg.FP64)
        for k, v in o:
            th.set_option(k, v)
        th.set_full_dump(fd)
        evals.append(th.run().evals)
        th.close()
        check_example_outputs("cfg2", golden, os.path.join(d, "Data"), fd, exact=False, group="full")
    assert evals[0] == evals[1], evals


def _cfg4_batch(nsets):
    sets = decks.cfg4_sets()[:nsets]
    beams = np.array([[s["beam"]["Power"], s["beam"]["Efficiency"], s["beam"]["Width_X"], s["beam"]["Width_Y"],
                       s["beam"]["Depth_Z"]] for s in sets])
    paths = [np.array(s["path"], dtype=np.float64) for s in sets]
    return beams, paths


def test_batched_nodes_and_sums_match_per_deck(tmp_path):
    """Device-built node sets and the persistent batched sum agree with the
    per-deck host builder / oracle for each parameter set."""
    import torch
    nsets = 12
    _flush(f"case: batched {nsets} sets on a 32x32x8 grid")
    beams, paths = _cfg4_batch(nsets)
    base = decks.cfg4_deck(0, 32, 32, 8)
    th = pkg.Thesis(base.write(str(tmp_path / "base")))
    th.batch_set(beams, paths)
    total, evals = th.batch_info()
    pnum = th.info.pnum
    T64 = torch.empty(nsets * pnum, dtype=torch.float64, device="cuda")
    T32 = torch.empty_like(T64)
    th.batch_eval_dev(T64.data_ptr(), pkg.FP64)
    th.batch_eval_dev(T32.data_ptr(), pkg.FP32C)
    torch.cuda.synchronize()
    T64 = T64.cpu().numpy().reshape(nsets, pnum)
    T32 = T32.cpu().numpy().reshape(nsets, pnum)
    count = 0
    for s in range(nsets):
        deck = decks.cfg4_deck(s, 32, 32, 8)
        pi = deck.write(str(tmp_path / f"s{s}"))
        ref_th = pkg.Thesis(pi)
        t = ref_th.info.scan_end
        host = ref_th.build_nodes(t)
        dev = th.batch_nodes(s)
        assert dev.shape == host.shape
        assert np.array_equal(dev[:, :7], host[:, :7])
        assert np.max(np.abs(dev[:, 7] - host[:, 7]) / np.abs(host[:, 7])) <= 1e-14
        count += host.shape[0]
        sim = of.Sim(pi)
        T_ref, _ = sim.snapshot_T(t)
        assert np.max(np.abs(T64[s] - T_ref) / T_ref) <= 1e-12
        assert np.max(np.abs(T32[s] - T_ref) / T_ref) <= 1e-5
    assert count == total and evals == total * pnum
    # autotuned configuration gives the same answer
    T2 = torch.empty(nsets * pnum, dtype=torch.float64, device="cuda")
    ppt, blocks = th.batch_autotune(T2.data_ptr(), pkg.FP32C)
    assert ppt in (1, 2, 4) and blocks > 0
    th.batch_eval_dev(T2.data_ptr(), pkg.FP32C)
    torch.cuda.synchronize()
    # another tile size rounds differently: it meets the same FP32c bar against the
    # FP64 sums (themselves within 1e-12 of the oracle above)
    assert np.max(np.abs(T2.cpu().numpy().reshape(nsets, pnum) - T64) / T64) <= 1e-5


def _adjoint_deck(nx, ny, nz, nh, pieces, seed):
    rng = np.random.default_rng(seed)
    rows = decks.cfg5_path(rng, nh=nh, pieces=pieces, length=1.0, x0=0.5, y0=0.8)
    return decks.Deck(
        name="adj", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[dict(decks.BEAM)], paths=[rows],
        domain={"X": decks.axis(0.0, 2e-3, nx), "Y": decks.axis(0.0, 2e-3, ny), "Z": decks.axis(-5e-4, 0.0, nz)},
        output=decks.out_cols("T"))


def test_adjoint_gradient_matches_fd_of_fp64(tmp_path):
    """FP32 forward + adjoint gradient within 1e-4 of central differences of the
    FP64 GPU forward, for every beam parameter and every path vertex coordinate."""
    import torch
    autograd = importlib.import_module("3dthesis_amd.autograd")
    _flush("case: adjoint 48x48x12, 2 hatches x 4 pieces")
    truth = _adjoint_deck(48, 48, 12, 2, 4, 5000)
    th = pkg.Thesis(truth.write(str(tmp_path / "truth")))
    beam_t, rows_t = th.design()
    pnum = th.info.pnum
    T_target = torch.empty(pnum, dtype=torch.float64, device="cuda")
    th.forward_dev(T_target.data_ptr(), pkg.FP64)
    torch.cuda.synchronize()
    rng = np.random.default_rng(5001)
    beam0 = beam_t * (1 + rng.uniform(-0.1, 0.1, 5))
    rows0 = rows_t.copy()
    rows0[:, 1:3] += rng.normal(0.0, 0.010, size=(rows0.shape[0], 2))

    field = autograd.SnapshotField(th, precision=pkg.FP32C, rows_template=rows0)
    b = torch.tensor(beam0, dtype=torch.float64, requires_grad=True)
    xy = torch.tensor(rows0[:, 1:3], dtype=torch.float64, requires_grad=True)
    T = field(b, xy)
    loss = ((T - T_target) ** 2).mean()
    loss.backward()
    g = np.concatenate([b.grad.numpy(), xy.grad.numpy().reshape(-1)])

    def L(beam, rows):
        th.set_design(beam, rows)
        T = torch.empty(pnum, dtype=torch.float64, device="cuda")
        th.forward_dev(T.data_ptr(), pkg.FP64)
        torch.cuda.synchronize()
        return float(((T - T_target) ** 2).mean())

    assert abs(beam0[0] * g[0] - beam0[1] * g[1]) <= 1e-5 * abs(beam0[0] * g[0])
    errs = []
    for k in range(g.size):
        bp, bm, rp, rm = beam0.copy(), beam0.copy(), rows0.copy(), rows0.copy()
        if k < 5:
            h = 1e-6 * abs(beam0[k])
            bp[k] += h
            bm[k] -= h
        else:
            r, c = divmod(k - 5, 2)
            h = 1e-6
            rp[r, 1 + c] += h
            rm[r, 1 + c] -= h
        fd = (L(bp, rp) - L(bm, rm)) / (2 * h)
        errs.append(abs(g[k] - fd) / max(abs(fd), 1e-30))
    gscale = np.abs(g[5:]).max()
    # vertex components much smaller than the largest are compared on that scale
    for k, e in enumerate(errs):
        if k >= 5 and abs(g[k]) < 1e-3 * gscale:
            continue
        assert e <= 1e-4, (k, e, g[k])


def test_cfg4_all_sets_full_size(tmp_path):
    """The batched workload at its full size -- 256 seeded parameter sets,
    128x128x32 each, one persistent launch -- FP32c within 1e-5 of FP64 for
    every set, and FP64 within 1e-12 of the oracle on the set with the most
    nodes and on set 0."""
    import torch
    _flush("case: batched 256 sets on 128x128x32, FP32c vs FP64 per set")
    beams, paths = _cfg4_batch(256)
    base = decks.cfg4_deck(0)
    th = pkg.Thesis(base.write(str(tmp_path / "base")))
    th.batch_set(beams, paths)
    pnum = th.info.pnum
    T64 = torch.empty(256 * pnum, dtype=torch.float64, device="cuda")
    T32 = torch.empty_like(T64)
    th.batch_eval_dev(T64.data_ptr(), pkg.FP64)
    th.batch_eval_dev(T32.data_ptr(), pkg.FP32C)
    torch.cuda.synchronize()
    rel = ((T32 - T64).abs() / T64).reshape(256, pnum).amax(dim=1).cpu().numpy()
    assert rel.max() <= 1e-5, (int(rel.argmax()), rel.max())
    sizes = [th.batch_nodes(k).shape[0] for k in range(256)]
    for k in sorted({0, int(np.argmax(sizes))}):
        _flush(f"case: batched set {k} ({sizes[k]} nodes) against the oracle")
        pi = decks.cfg4_deck(k).write(str(tmp_path / f"s{k}"))
        sim = of.Sim(pi)
        T_ref, _ = sim.snapshot_T(float(sim.s.all_end), nthreads=16)
        got = T64[k * pnum:(k + 1) * pnum].cpu().numpy()
        assert np.max(np.abs(got - T_ref) / T_ref) <= 1e-12
    th.close()


def test_cfg5_adjoint_full_size_fd(tmp_path):
    """The adjoint workload at full size (256x256x64, 200 perturbed vertices,
    405 gradient components): the FP32 forward + adjoint gradient within 1e-4
    of central differences of the FP64 forward on the 5 beam parameters and 30
    seeded vertex coordinates (those much smaller than the largest are compared
    on that scale)."""
    import torch
    autograd = importlib.import_module("3dthesis_amd.autograd")
    _flush("case: adjoint at full size, 256x256x64, 200 vertices")
    th = pkg.Thesis(decks.cfg5().write(str(tmp_path / "truth")))
    beam_t, rows_t = th.design()
    pnum = th.info.pnum
    T_target = torch.empty(pnum, dtype=torch.float64, device="cuda")
    th.forward_dev(T_target.data_ptr(), pkg.FP64)
    torch.cuda.synchronize()
    rng = np.random.default_rng(5001)
    beam0 = beam_t * (1 + rng.uniform(-0.1, 0.1, 5))
    rows0 = rows_t.copy()
    rows0[:, 1:3] += rng.normal(0.0, 0.010, size=(rows0.shape[0], 2))
    field = autograd.SnapshotField(th, precision=pkg.FP32C, rows_template=rows0)
    b = torch.tensor(beam0, dtype=torch.float64, requires_grad=True)
    xy = torch.tensor(rows0[:, 1:3], dtype=torch.float64, requires_grad=True)
    loss = ((field(b, xy) - T_target) ** 2).mean()
    loss.backward()
    g = np.concatenate([b.grad.numpy(), xy.grad.numpy().reshape(-1)])
    assert g.size == 405

    def L(beam, rows):
        th.set_design(beam, rows)
        T = torch.empty(pnum, dtype=torch.float64, device="cuda")
        th.forward_dev(T.data_ptr(), pkg.FP64)
        torch.cuda.synchronize()
        return float(((T - T_target) ** 2).mean())

    picks = list(range(5)) + sorted(np.random.default_rng(5002).choice(np.arange(5, 405), 30, replace=False).tolist())
    gscale = np.abs(g[5:]).max()
    for k in picks:
        bp, bm, rp, rm = beam0.copy(), beam0.copy(), rows0.copy(), rows0.copy()
        if k < 5:
            h = 1e-6 * abs(beam0[k])
            bp[k] += h
            bm[k] -= h
        else:
            r, c = divmod(k - 5, 2)
            h = 1e-6
            rp[r, 1 + c] += h
            rm[r, 1 + c] -= h
        fd = (L(bp, rp) - L(bm, rm)) / (2 * h)
        if k >= 5 and abs(g[k]) < 1e-3 * gscale:
            continue
        assert abs(g[k] - fd) <= 1e-4 * abs(fd), (k, g[k], fd)
    th.close()


# Intermediate outputs (OutputFrequency) end windows and stop the pipelining of
# windows; every written file must match the one-step-at-a-time run's, with
# its queued solidifications solved at the outputs ("sync") or in ranges of a
# few entries beside the steps on 8 CUs ("async").
@pytest.mark.parametrize("wl", ["cfg2_small", "cfg3_small"])
def test_windowed_intermediate_outputs_match_per_step(wl, tmp_path):
    _flush(f"case: run {wl} FP64 with OutputFrequency, windowed vs per-step")
    runs = {"sync": (0, [("async_sol", 0)]), "async": (0, [("step_sol_batch", 8), ("step_sol_cus", 8)]),
            "windowed": (16, [])}
    files = {}
    for name, (window, opts) in runs.items():
        deck = decks.WORKLOADS[wl]()
        deck.mode = dict(deck.mode, OutputFrequency=23)
        d = str(tmp_path / name)
        pi = deck.write(d)
        th = pkg.Thesis(pi, precision=pkg.FP64)
        th.set_option("window", window)
        for k, v in opts:
            th.set_option(k, v)
        th.run()
        data = os.path.join(d, "Data")
        files[name] = {f: open(os.path.join(data, f)).read().splitlines() for f in sorted(os.listdir(data))}
        th.close()
    assert len(files["sync"]) > 3
    for name in ("async", "windowed"):
        assert files["sync"].keys() == files[name].keys(), name
        for f in files["sync"]:
            a, b = files["sync"][f], files[name][f]
            assert len(a) == len(b) and a[0] == b[0], (name, f)
            if len(a) > 1:
                A = np.array([[float(v) for v in l.split(",")] for l in a[1:]])
                B = np.array([[float(v) for v in l.split(",")] for l in b[1:]])
                assert np.allclose(A, B, rtol=2e-6, atol=0), (name, f)


# Queued solidifications of the one-step-at-a-time driver solved beside the
# steps, in ranges of a few entries, on a few CUs or on the whole device.
@pytest.mark.parametrize("wl,cus", [("cfg2_small", 8), ("cfg3_small", 8), ("cfg3_small", 0),
                                    ("compress_small_surface", 4)])
def test_per_step_solidify_ranges_beside_steps(wl, cus, golden, tmp_path):
    _flush(f"case: run {wl} FP64 window=0 step_sol_batch=16 step_sol_cus={cus}")
    _run_fp64_vs_reference(wl, golden, tmp_path, 0, opts=[("step_sol_batch", 16), ("step_sol_cus", cus)])
