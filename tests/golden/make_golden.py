"""Write the golden fixtures from the reference itself.

Runs oracle/_ref/thesis_ref (the reference's library sources compiled in place
by oracle/Makefile and linked with oracle/ref_harness.cpp) on the workload decks
of decks.py, and stores what it printed as small fixtures:

  tests/golden/ref_nodes.npz   node lists (N x 8: xb,yb,zb,1/phix,1/phiy,1/phiz,dtau,expmod)
  tests/golden/ref_tgrid.npz   T at every grid point of small snapshot decks
  tests/golden/ref_runs.npz    full mode runs: for every output file, the point
                               indices, the full-precision columns and the gzip of
                               the reference's own 6-digit CSV; with T_hist on, every
                               recorded (p, t, T) in full precision; with RDF every
                               (p, tm, tl, cr); with Stork tracking the SHA-256 of the
                               full-precision RRDF record dump (in ref_meta.json)

  tests/golden/ref_runs_clang.npz  the same run records for a few Solidification
                               decks from oracle/_ref_clang/thesis_ref, the reference
                               built with clang++, whose abs() at Grid.cpp:381/:535
                               is the floating one (the product's abs_mode 1)
  tests/golden/ref_examples.npz  the reference's own example decks
                               (tests/golden/examples/, copied from its examples/
                               directory): their outputs are megabytes, so each is
                               pinned by its row count, the SHA-256 of its row set
                               (point indices), about a thousand evenly spaced rows
                               in full precision and as printed, and for Stork the
                               SHA-256 of the full-precision RRDF records

  tests/golden/ref_full.npz    the full-size cfg2 workload pinned the same way

Usage:  python tests/golden/make_golden.py [nodes] [tgrid] [runs] [clang] [examples] [full]
(no argument: all). Needs the reference checkout; run `make -C oracle ref
ref_clang` first. All decks pin MaxThreads 1.
"""
from __future__ import annotations

import glob
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import decks  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "thesis_ref")
REF_CLANG = os.path.join(ROOT, "oracle", "_ref_clang", "thesis_ref")
EXAMPLES_DIR = os.path.join(HERE, "examples")
EXAMPLE_SAMPLE = 1000

# (case name, workload, time as fraction of scan end or absolute seconds, is_sol)
NODE_CASES = [
    ("cfg1_end", "cfg1", ("frac", 1.0), 0),
    ("cfg1_mid", "cfg1", ("frac", 0.37), 0),
    ("cfg1_after", "cfg1", ("abs", 3.1e-3), 0),
    ("cfg1_sol", "cfg1", ("frac", 0.5), 1),
    ("cfg2_mid", "cfg2", ("abs", 0.0213456), 0),
    ("cfg2_end", "cfg2", ("frac", 1.0), 0),
    ("cfg2_sol", "cfg2", ("abs", 0.0123456789), 1),
    ("cfg2_cool", "cfg2", ("abs", 0.05), 0),
    ("cfg3_mid", "cfg3", ("abs", 0.011), 0),
    ("cfg3_sol", "cfg3", ("abs", 0.0071), 1),
    ("cfg5_end", "cfg5", ("frac", 1.0), 0),
    ("snapv_t0", "snapvol_small", ("abs", 0.0), 0),
    ("comp_mid", "compress_small", ("abs", 0.001), 0),
    ("comp_merge", "compress_small", ("abs", 0.002), 0),
    ("comp_sol", "compress_small", ("abs", 0.0015), 1),
    ("comp_cool", "compress_small", ("abs", 0.004), 0),
    ("walls_mid", "walls_small", ("abs", 0.0008), 0),
    ("walls_sol", "walls_small", ("abs", 0.0008), 1),
    ("walls_cool", "walls_small", ("abs", 0.002), 0),
]
TGRID_CASES = [("cfg1_small", "cfg1_small", 1.0), ("snapv", "snapvol_small", 0.55)]
RUN_CASES = ["none_small", "snapvol_small", "snapsurf_small", "cfg2_small", "cfg2_small_volume",
             "cfg3_small", "cfg1_small", "compress_small", "compress_small_surface", "walls_small",
             "walls_small_volume", "points_small", "points_small_sol", "thist_none_small", "rdf_small",
             "rdf_small_volume", "stork_small", "rdf_small_windowed", "perim_small", "perim_small_volume",
             "thist_surface_small", "thist_volume_small", "thist_snapvol_small"]
# (no compressed deck: a merged compressed node's zb is left uninitialised by
# the reference, Calc.cpp:452; the g++ build's stack residue is subnormal, the
# clang build's moves T at the iterated liquidus times)
CLANG_RUN_CASES = ["cfg2_small", "cfg2_small_volume", "rdf_small", "walls_small_volume", "perim_small"]
# full-size workloads pinned the way the examples are (their outputs are megabytes)
FULL_CASES = ["cfg2"]
EXAMPLE_CASES = ["snapshot", "snapshot_T_hist", "solidification", "solidification_multibeam",
                 "solidification_mpstats", "solidification_stork"]


def scan_end(deck: decks.Deck) -> float:
    """End time of the scan from the path rows (Init.cpp:607-622 timing rule)."""
    end = 0.0
    for path in deck.paths:
        t, px, py, pz = 0.0, 0.0, 0.0, 0.0
        for m, x, y, z, q, prm in path:
            x, y, z = x * 1e-3, y * 1e-3, z * 1e-3
            if m == 1:
                t += prm
            else:
                t += float(np.sqrt((x - px) ** 2 + (y - py) ** 2 + (z - pz) ** 2)) / prm
            px, py, pz = x, y, z
        end = max(end, t)
    return end


def read_full(path):
    rows = [l.split() for l in open(path) if l.strip()]
    if not rows:
        return np.zeros(0, np.int32), np.zeros((0, 0))
    a = np.array([[float(v) for v in r] for r in rows])
    return a[:, 0].astype(np.int32), a[:, 1:]


def main() -> None:
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    groups = set(sys.argv[1:]) or {"nodes", "tgrid", "runs", "clang", "examples", "full"}
    work = tempfile.mkdtemp(prefix="thesis_golden_")
    meta_path = os.path.join(HERE, "ref_meta.json")
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    nodes_out, tgrid_out = {}, {}
    if "nodes" in groups:
        meta["nodes"] = {}
    if "tgrid" in groups:
        meta["tgrid"] = {}
    for case, wl, (kind, val), is_sol in (NODE_CASES if "nodes" in groups else []):
        deck = decks.WORKLOADS[wl]()
        pi = deck.write(os.path.join(work, "n_" + case))
        t = val * scan_end(deck) if kind == "frac" else val
        out = os.path.join(work, case + ".nodes")
        subprocess.run([REF, "nodes", pi, repr(t), str(is_sol), out], check=True, capture_output=True)
        lines = open(out).read().split("\n")
        n = int(lines[0])
        arr = np.array([[float(v) for v in l.split()] for l in lines[1:1 + n]]).reshape(n, 8)
        nodes_out[case] = arr
        meta["nodes"][case] = {"workload": wl, "t": t, "is_sol": is_sol, "n": n}
    for case, wl, frac in (TGRID_CASES if "tgrid" in groups else []):
        deck = decks.WORKLOADS[wl]()
        pi = deck.write(os.path.join(work, "t_" + case))
        t = frac * scan_end(deck)
        out = os.path.join(work, case + ".T")
        subprocess.run([REF, "tgrid", pi, repr(t), out], check=True, capture_output=True)
        lines = open(out).read().split()
        pnum, nn = int(lines[0]), int(lines[1])
        tgrid_out[case] = np.array([float(v) for v in lines[2:2 + pnum]])
        meta["tgrid"][case] = {"workload": wl, "t": t, "nodes": nn}
    if "runs" in groups:
        meta["runs"] = {}
        runs_out = {}
        for wl in RUN_CASES:
            run_case(REF, wl, os.path.join(work, "r_" + wl), runs_out, meta["runs"])
        np.savez_compressed(os.path.join(HERE, "ref_runs.npz"), **runs_out)
    if "clang" in groups:
        if not os.path.exists(REF_CLANG):
            sys.exit(f"{REF_CLANG} missing: run `make -C oracle ref_clang` first")
        meta["runs_clang"] = {}
        clang_out = {}
        for wl in CLANG_RUN_CASES:
            run_case(REF_CLANG, wl, os.path.join(work, "c_" + wl), clang_out, meta["runs_clang"])
        np.savez_compressed(os.path.join(HERE, "ref_runs_clang.npz"), **clang_out)
    if "examples" in groups:
        meta["examples"] = {}
        ex_out = {}
        for name in EXAMPLE_CASES:
            example_case(name, os.path.join(work, "e_" + name), ex_out, meta["examples"])
        np.savez_compressed(os.path.join(HERE, "ref_examples.npz"), **ex_out)
    if "full" in groups:
        meta["full"] = {}
        full_out = {}
        for wl in FULL_CASES:
            d = os.path.join(work, "f_" + wl)
            decks.WORKLOADS[wl]().write(d)
            sampled_case(wl, d, full_out, meta["full"])
        np.savez_compressed(os.path.join(HERE, "ref_full.npz"), **full_out)
    if "nodes" in groups:
        np.savez_compressed(os.path.join(HERE, "ref_nodes.npz"), **nodes_out)
    if "tgrid" in groups:
        np.savez_compressed(os.path.join(HERE, "ref_tgrid.npz"), **tgrid_out)
    with open(os.path.join(HERE, "ref_meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    shutil.rmtree(work, ignore_errors=True)
    print("wrote fixtures:", {k: len(v) for k, v in meta.items()})


def run_case(ref: str, wl: str, d: str, out: dict, meta: dict) -> None:
    """One whole run of workload wl by a reference build: its CSVs (gzip), the
    final output in full precision, and the T_hist, RDF and RRDF records."""
    deck = decks.WORKLOADS[wl]()
    pi = deck.write(d)
    r = subprocess.run([ref, "run", pi], check=True, capture_output=True, text=True)
    files = {}
    for csv in sorted(glob.glob(os.path.join(d, "Data", "*.csv"))):
        base = os.path.basename(csv)
        # "<Name>.<tag>.csv", or "<Name>T.hist.csv" (Grid.cpp:110)
        tag = base[len(deck.name) + 1:-4] if base.startswith(deck.name + ".") else base[len(deck.name):-4]
        out[f"{wl}/{tag}/csv"] = np.frombuffer(gzip.compress(open(csv, "rb").read(), mtime=0), np.uint8)
        files[tag] = os.path.basename(csv)
    p, v = read_full(glob.glob(os.path.join(d, "Data", "*.final.full.txt"))[0])
    out[f"{wl}/final/p"] = p
    out[f"{wl}/final/v"] = v
    hist = glob.glob(os.path.join(d, "Data", "*.T_hist.full.txt"))
    if hist:
        hp, hv = read_full(hist[0])
        out[f"{wl}/thist/p"] = hp
        out[f"{wl}/thist/v"] = hv
    rdf = glob.glob(os.path.join(d, "Data", "*.RDF.full.txt"))
    if rdf:
        rp, rv = read_full(rdf[0])
        out[f"{wl}/rdf/p"] = rp
        out[f"{wl}/rdf/v"] = rv
    meta[wl] = {"files": files, "exec": r.stdout.strip().splitlines()[-1]}
    rrdf = glob.glob(os.path.join(d, "Data", "*.RRDF.full.txt"))
    if rrdf:  # large: pinned by the digest of the full-precision text
        meta[wl]["rrdf_sha256"] = hashlib.sha256(open(rrdf[0], "rb").read()).hexdigest()
        meta[wl]["rrdf_records"] = sum(1 for _ in open(rrdf[0]))


def sample_rows(n: int) -> np.ndarray:
    return np.unique(np.linspace(0, n - 1, min(n, EXAMPLE_SAMPLE)).astype(np.int64)) if n else np.zeros(0, np.int64)


def example_case(name: str, d: str, out: dict, meta: dict) -> None:
    """One of the reference's example decks, run by the reference with
    MaxThreads 1 (the deck's own thread count changes node times by ~1e-19 s,
    Calc.cpp:34): row counts, row-set digests and sampled rows of its outputs."""
    shutil.copytree(os.path.join(EXAMPLES_DIR, name), d)
    sampled_case(name, d, out, meta)


def sampled_case(name: str, d: str, out: dict, meta: dict) -> None:
    """Run the deck in d (ParamInput.txt) by the reference with MaxThreads 1 and
    keep row counts, row-set digests and sampled rows of its outputs."""
    with open(os.path.join(d, "Settings.txt"), "w") as f:
        f.write("Compute\n{\n\tMaxThreads\t1\n}\n")
    os.makedirs(os.path.join(d, "Data"), exist_ok=True)
    r = subprocess.run([REF, "run", "ParamInput.txt"], check=True, capture_output=True, text=True, cwd=d)
    m = {"exec": r.stdout.strip().splitlines()[-1], "files": {}}
    data = os.path.join(d, "Data")
    for csv in sorted(glob.glob(os.path.join(data, "*.csv"))):
        base = os.path.basename(csv)
        if ".T_hist." in base and not base.endswith("T.hist.csv"):
            continue
        lines = open(csv).read().splitlines()
        rows = sample_rows(len(lines) - 1)
        m["files"][base] = {"rows": len(lines) - 1, "header": lines[0]}
        out[f"{name}/{base}/idx"] = rows
        out[f"{name}/{base}/lines"] = np.frombuffer(
            gzip.compress("\n".join(lines[1 + i] for i in rows).encode(), mtime=0), np.uint8)
    for kind in ("final", "T_hist", "RDF"):
        f = glob.glob(os.path.join(data, f"*.{kind}.full.txt"))
        if not f:
            continue
        p, v = read_full(f[0])
        rows = sample_rows(len(p))
        m[kind] = {"rows": int(len(p)), "p_sha256": hashlib.sha256(p.astype(np.int64).tobytes()).hexdigest()}
        out[f"{name}/{kind}/idx"] = rows
        out[f"{name}/{kind}/p"] = p[rows]
        out[f"{name}/{kind}/v"] = v[rows]
    rrdf = glob.glob(os.path.join(data, "*.RRDF.full.txt"))
    if rrdf:
        m["rrdf_sha256"] = hashlib.sha256(open(rrdf[0], "rb").read()).hexdigest()
        m["rrdf_records"] = sum(1 for _ in open(rrdf[0]))
    meta[name] = m


if __name__ == "__main__":
    main()
