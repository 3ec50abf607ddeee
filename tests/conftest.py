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
        "meta": json.load(open(os.path.join(g, "ref_meta.json"))),
    }


@pytest.fixture
def deckdir(tmp_path):
    """Write a named workload deck into a fresh directory; returns ParamInput path."""
    import decks

    def make(name, **kw):
        d = decks.WORKLOADS[name]() if not kw else decks.WORKLOADS[name](**kw)
        return d.write(str(tmp_path / name)), d
    return make
