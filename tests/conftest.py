import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, os.path.join(HERE, "golden"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    import json
    g = os.path.join(HERE, "golden")
    return {
        "nodes": np.load(os.path.join(g, "ref_nodes.npz")),
        "tgrid": np.load(os.path.join(g, "ref_tgrid.npz")),
        "runs": np.load(os.path.join(g, "ref_runs.npz")),
        "runs_clang": np.load(os.path.join(g, "ref_runs_clang.npz")),
        "examples": np.load(os.path.join(g, "ref_examples.npz")),
        "full": np.load(os.path.join(g, "ref_full.npz")),
        "meta": json.load(open(os.path.join(g, "ref_meta.json"))),
    }


def example_deck(name, dest):
    """Copy one of the reference's example decks (tests/golden/examples/<name>)
    into dest; its ParamInput names its files relative to that directory."""
    import shutil
    src = os.path.join(HERE, "golden", "examples", name)
    os.makedirs(dest)  # file contents only: the checkout may be read-only, the copy must not be
    for f in os.listdir(src):
        shutil.copyfile(os.path.join(src, f), os.path.join(dest, f))
    os.makedirs(os.path.join(dest, "Data"), exist_ok=True)
    return os.path.join(dest, "ParamInput.txt")


def check_example_outputs(name, golden, data_dir, full_dir, exact, group="examples"):
    """Compare a run of the reference's example deck `name` with its fixtures:
    every output CSV's header and row count, the full-precision final table's
    (and T_hist / RDF records') row count and row-set digest, and the sampled
    rows -- exactly (exact=True) or to 1e-9 relative in full precision and 6
    printed digits (2e-6) in the CSVs."""
    import gzip
    import hashlib
    import numpy as np
    import oracle_ffi as of
    meta = golden["meta"][group][name]
    ex = golden[group]
    for base, fm in meta["files"].items():
        lines = open(os.path.join(data_dir, base)).read().splitlines()
        assert lines[0] == fm["header"], base
        assert len(lines) - 1 == fm["rows"], (base, len(lines) - 1, fm["rows"])
        idx = ex[f"{name}/{base}/idx"]
        if not len(idx):
            continue
        ref = gzip.decompress(ex[f"{name}/{base}/lines"].tobytes()).decode().split("\n")
        got = [lines[1 + i] for i in idx]
        if exact:
            assert got == ref, base
        else:
            a = np.array([[float(v) for v in l.split(",")] for l in ref])
            b = np.array([[float(v) for v in l.split(",")] for l in got])
            assert np.allclose(a, b, rtol=2e-6, atol=1e-280), base
    for kind, fname in (("final", None), ("T_hist", ".T_hist.full.txt"), ("RDF", ".RDF.full.txt")):
        if kind not in meta or (kind == "final" and "rrdf_sha256" in meta):
            continue  # Stork writes its records instead of the grid table (Main.cpp:104-112)
        files = sorted(os.listdir(full_dir))
        if kind == "final":
            path = [f for f in files if f.endswith(".full.txt") and not any(k in f for k in (".T_hist.", ".RDF.", ".RRDF."))][-1]
        else:
            path = [f for f in files if f.endswith(fname)][0]
        p, v = of.read_full(os.path.join(full_dir, path))
        assert len(p) == meta[kind]["rows"], (kind, len(p))
        assert hashlib.sha256(p.astype(np.int64).tobytes()).hexdigest() == meta[kind]["p_sha256"], kind
        idx = ex[f"{name}/{kind}/idx"]
        if not len(idx):
            continue
        assert np.array_equal(p[idx], ex[f"{name}/{kind}/p"])
        rv = ex[f"{name}/{kind}/v"]
        if exact:
            assert np.array_equal(v[idx], rv), kind
        else:
            rel = np.abs(v[idx] - rv) / np.maximum(np.abs(rv), 1e-280)
            assert np.nanmax(rel) <= 1e-9, (kind, np.nanmax(rel))


@pytest.fixture
def deckdir(tmp_path):
    """Write a named workload deck into a fresh directory; returns ParamInput path."""
    import decks

    def make(name, **kw):
        d = decks.WORKLOADS[name]() if not kw else decks.WORKLOADS[name](**kw)
        return d.write(str(tmp_path / name)), d
    return make
