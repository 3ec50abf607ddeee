"""The `thesis` command line as a drop-in for the reference's `3DThesis
ParamInput.txt` (bin/Main.cpp:44-124): the same input files in, the same
Data/<Name>.<tag>.csv files out, checked against the reference's own outputs
for those decks (tests/golden/ref_runs.npz)."""
from __future__ import annotations

import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

import decks
from conftest import check_example_outputs, example_deck
from make_golden import EXAMPLE_CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "3dthesis_amd", "thesis")


def _table(lines):
    if len(lines) <= 1:
        return np.zeros((0, 1))
    return np.array([[float(v) for v in l.split(",")] for l in lines[1:]])


def test_cli_without_input_names_its_usage():
    if not os.path.exists(CLI):
        pytest.fail("the thesis command line is not built (run __graft_entry__.build())")
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "ParamInput.txt" in r.stderr


@pytest.mark.parametrize("args,msg", [(["deck.txt", "--fulll"], "unknown option --fulll"),
                                      (["deck.txt", "--full"], "--full needs a directory"),
                                      (["--full", "--fp32c", "deck.txt"], "--full needs a directory"),
                                      (["a.txt", "b.txt"], "more than one input file"),
                                      (["--fp32c"], "Input file argument required")])
def test_cli_rejects_bad_arguments(args, msg):
    if not os.path.exists(CLI):
        pytest.fail("the thesis command line is not built (run __graft_entry__.build())")
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert msg in r.stderr and "usage: thesis ParamInput.txt" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("wl", ["none_small", "snapvol_small", "cfg2_small", "cfg3_small", "compress_small",
                                "points_small", "rdf_small"])
def test_cli_run_writes_the_reference_outputs(wl, golden, tmp_path):
    print(f"case: thesis command line, {wl}", flush=True)
    sys.stdout.flush()
    deck = decks.WORKLOADS[wl]()
    d = str(tmp_path / wl)
    pi = deck.write(d)
    r = subprocess.run([CLI, pi], capture_output=True, text=True, timeout=600, cwd=d)
    assert r.returncode == 0, r.stderr
    assert "Execution time (s):" in r.stdout
    meta = golden["meta"]["runs"][wl]
    for tag, fname in meta["files"].items():
        ref = gzip.decompress(golden["runs"][f"{wl}/{tag}/csv"].tobytes()).decode().splitlines()
        got = open(os.path.join(d, "Data", fname)).read().splitlines()
        assert got[0] == ref[0], fname  # header / column order
        assert len(got) == len(ref), fname
        # 6 printed digits, as in the library-level comparison
        assert np.allclose(_table(ref), _table(got), rtol=2e-6, atol=1e-280), fname


@pytest.mark.gpu
def test_cli_reports_a_bad_input_file(tmp_path):
    print("case: thesis command line, missing input file", flush=True)
    r = subprocess.run([CLI, str(tmp_path / "missing.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Fatal Error" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", EXAMPLE_CASES)
def test_cli_runs_the_reference_examples(name, golden, tmp_path):
    """The reference's own example decks, as shipped (their MaxThreads is read
    and ignored), through the command line from their own directory, against
    the reference's outputs for them: row counts, row-set digests and sampled
    rows (6 printed digits; 1e-9 in full precision)."""
    print(f"case: thesis command line, example {name}", flush=True)
    sys.stdout.flush()
    pi = example_deck(name, str(tmp_path / name))
    d = os.path.dirname(pi)
    os.makedirs(os.path.join(d, "full"), exist_ok=True)
    r = subprocess.run([CLI, "ParamInput.txt", "--full", "full"], capture_output=True, text=True, timeout=600, cwd=d)
    assert r.returncode == 0, r.stderr
    check_example_outputs(name, golden, os.path.join(d, "Data"), os.path.join(d, "full"), exact=False)
    meta = golden["meta"]["examples"][name]
    if "rrdf_sha256" in meta:  # Stork records: counted here; their values against the oracle in test_gpu
        n = sum(1 for _ in open(os.path.join(d, "full", "stork.RRDF.full.txt")))
        assert n == meta["rrdf_records"]
