"""CPU-side checks of the product: the C-ABI library loads and exports what the
header declares, the input surface parses decks like the reference, and the
host node builder reproduces the reference's node lists bit for bit.
No GPU is touched here (node builds and parsing are host code)."""
from __future__ import annotations

import ctypes as C
import importlib
import os
import re
import shutil

import numpy as np
import pytest

import decks
import oracle_ffi as of

pkg = importlib.import_module("3dthesis_amd")


def test_library_exports_every_declared_symbol():
    L = C.CDLL(pkg.LIB_PATH)
    decl = re.findall(r"^\s*(?:int|const char\*)\s+(thesis_\w+)\s*\(", open(pkg.HEADER_PATH).read(), re.M)
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(pkg.exported_symbols())


@pytest.mark.parametrize("wl", ["cfg1", "cfg2", "cfg3", "cfg5", "cfg2_small", "cfg3_small", "none_small",
                                "snapvol_small", "compress_small", "walls_small_volume", "points_small",
                                "points_small_sol", "rdf_small", "stork_small"])
def test_parse_matches_oracle(wl, tmp_path):
    pi = decks.WORKLOADS[wl]().write(str(tmp_path / wl))
    th = pkg.Thesis(pi)
    sim = of.Sim(pi)
    s = sim.s
    i = th.info
    assert (i.xnum, i.ynum, i.znum, i.pnum) == (s.xnum, s.ynum, s.znum, s.pnum)
    for a in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        assert getattr(i, a) == getattr(s, a)
    assert i.scan_end == s.all_end and i.nbeams == s.nbeams and i.T_init == s.T_init and i.T_liq == s.T_liq
    assert i.alpha == s.a
    assert i.mode == s.mode and i.tracking == s.tracking
    if i.mode == 0:
        assert np.array_equal(th.snapshot_times(), np.ctypeslib.as_array(s.snap_times, shape=(s.nsnap,)))
    else:
        assert i.dt == s.dt


@pytest.mark.parametrize("case", [c for c in ("cfg1_end", "cfg1_mid", "cfg1_after", "cfg1_sol", "cfg2_mid",
                                              "cfg2_end", "cfg2_sol", "cfg2_cool", "cfg3_mid", "cfg3_sol",
                                              "cfg5_end", "snapv_t0", "comp_mid", "comp_merge", "comp_sol",
                                              "comp_cool", "walls_mid", "walls_sol", "walls_cool")])
def test_host_nodes_bit_identical_to_reference(case, golden, tmp_path):
    m = golden["meta"]["nodes"][case]
    pi = decks.WORKLOADS[m["workload"]]().write(str(tmp_path / case))
    th = pkg.Thesis(pi)
    got = th.build_nodes(m["t"], bool(m["is_sol"]))
    ref = golden["nodes"][case].copy()
    assert got.shape == ref.shape
    if case.startswith("comp"):
        # merged compressed nodes: the reference leaves zb uninitialised
        # (Calc.cpp:452, subnormal residue in its build); this library writes 0
        ref[(np.abs(ref[:, 2]) < np.finfo(np.float64).tiny), 2] = 0.0
    assert np.array_equal(got, ref)


def test_host_nodes_follow_the_oracle_over_time(tmp_path):
    """Compressed and wall-image builds at many times, main and per-point
    kinds, against the oracle (itself pinned to the reference's fixtures)."""
    for wl in ("compress_small", "walls_small"):
        pi = decks.WORKLOADS[wl]().write(str(tmp_path / wl))
        th = pkg.Thesis(pi)
        sim = of.Sim(pi)
        end = th.info.scan_end
        for t in np.linspace(0.0, 1.3 * end, 41):
            for is_sol in (False, True):
                assert np.array_equal(th.build_nodes(t, is_sol), sim.nodes(t, is_sol)), (wl, t, is_sol)


def test_wall_images_make_walls_adiabatic(tmp_path):
    """Method of images: with one reflection of an X_min wall, the node set is
    symmetric about the wall, so T is mirror-symmetric and dT/dx vanishes on it."""
    d = decks.walls_small(reflections=1)
    d.domain["BoundaryConditions"] = {"X_min": 0.0, "Reflections": 1}
    pi = d.write(str(tmp_path / "w1"))
    sim = of.Sim(pi)
    th = pkg.Thesis(pi)
    n = th.build_nodes(3e-4)
    plain = n.shape[0] // 2
    assert n.shape[0] == 2 * plain
    assert np.array_equal(n[plain:, 0], -n[:plain, 0])
    assert np.array_equal(n[plain:, 1:], n[:plain, 1:])
    for x in (1e-5, 4e-5, 1.2e-4):
        a = of.calc_T_xyz(sim, n, x, 1e-4, -2e-5)
        b = of.calc_T_xyz(sim, n, -x, 1e-4, -2e-5)
        assert abs(a - b) <= 1e-12 * a


def test_compression_refuses_vertices_the_reference_loops_on(tmp_path):
    """GaussCompressIntegrate's R_max skip (Calc.cpp:290-302) never advances
    past a vertex outside the R_max box; both the oracle and the library refuse
    such a deck rather than loop."""
    d = decks.compress_small()
    d.paths[0].append((0, 200.0, 0.3, 0.0, 1.0, 50.0))  # 0.2 m away: beyond R_max (about 21 mm here)
    pi = d.write(str(tmp_path / "far"))
    _expect_error(pi, 1, "R_max box")
    assert "R_max box" in of.Sim(pi).error
    d.settings = {"Compute": {"MaxThreads": 1}}
    pkg.Thesis(d.write(str(tmp_path / "far_plain")))  # without compression the deck is fine


def test_reference_examples_parse_and_build(tmp_path):
    """The reference's own example decks (inputs only) parse; multi-beam wildcards expand."""
    here = os.path.dirname(os.path.abspath(__file__))
    ex = os.path.join(here, "golden", "examples")
    for name in sorted(os.listdir(ex)):
        d = tmp_path / name
        shutil.copytree(os.path.join(ex, name), d)
        cwd = os.getcwd()
        os.chdir(d)
        try:
            th = pkg.Thesis("ParamInput.txt")
            sim = of.Sim("ParamInput.txt")
            assert th.info.pnum == sim.s.pnum and th.info.nbeams == sim.s.nbeams
            t = 0.5 * th.info.scan_end
            assert np.array_equal(th.build_nodes(t), sim.nodes(t))
        finally:
            os.chdir(cwd)


def _expect_error(pi, code, fragment):
    with pytest.raises(pkg.ThesisError) as e:
        pkg.Thesis(pi)
    assert e.value.code == code
    assert fragment in str(e.value)


def test_input_errors_are_status_codes(tmp_path):
    d = decks.cfg1(10, 5, 3)
    pi = d.write(str(tmp_path / "bad"))
    with open(tmp_path / "bad" / "Mode.txt", "a") as f:
        f.write("Solidification\n{\n\tTimestep 1e-5\n}\n")
    _expect_error(pi, 1, "Too many modes")
    open(tmp_path / "bad" / "Mode.txt", "w").write("Nothing\n{\n}\n")
    _expect_error(pi, 1, "No modes")
    open(tmp_path / "bad" / "Mode.txt", "w").write("Snapshots\n{\n\tTimes 1e-3\n\tTracking Geometry\n}\n")
    _expect_error(pi, 1, "outside this library's scope")
    _expect_error(str(tmp_path / "missing.txt"), 1, "cannot open")
    s = open(pi).read().replace("Beam.txt", "Be*am*.txt")
    open(pi, "w").write(s)
    _expect_error(pi, 1, "Too many '*'")


def test_run_options_are_checked(tmp_path):
    """Driver options are validated on the host, before any GPU work; the
    window diagnostics name one field per value the C-ABI fills."""
    th = pkg.Thesis(decks.cfg2_small().write(str(tmp_path / "opt")))
    for key, value in (("volume_switch", 24), ("volume_switch", 0), ("window", 0), ("async_sol", 1)):
        th.set_option(key, value)
    for key, value in (("volume_switch", -1), ("window", 5000), ("async_sol", 2), ("no_such_option", 1)):
        with pytest.raises(pkg.ThesisError) as e:
            th.set_option(key, value)
        assert e.value.code == 2, (key, value)
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "thesis_amd.h")).read()
    m = re.search(r"out\[0\.\.(\d+)\] = windows", hdr)
    assert m and int(m.group(1)) + 1 == len(pkg.Thesis.WINDOW_FIELDS)
    th.close()


def test_defaults_when_files_missing(tmp_path):
    """Missing optional files fall back to the reference defaults (Init.cpp:627-814)."""
    d = decks.cfg1(10, 5, 3)
    pi = d.write(str(tmp_path / "dflt"))
    os.remove(tmp_path / "dflt" / "Settings.txt")
    th = pkg.Thesis(pi)
    sim = of.Sim(pi)
    assert th.info.pnum == sim.s.pnum
    assert np.array_equal(th.build_nodes(1e-3), sim.nodes(1e-3))


def test_design_interface_refuses_compression_and_walls(tmp_path):
    """The design gradients differentiate plain node lists; decks with path
    compression or wall images are refused there with a status code."""
    for wl in ("compress_small", "walls_small"):
        th = pkg.Thesis(decks.WORKLOADS[wl]().write(str(tmp_path / wl)))
        beam, rows = th.design()
        with pytest.raises(pkg.ThesisError) as e:
            th.set_design(beam, rows)
        assert "compression and wall images are off" in str(e.value)


def test_custom_points_and_history_gates(tmp_path):
    """Custom points read like Init::FileRead_Points (header line skipped, mm to
    m by division); T_hist with Stork tracking on a grid and an empty point
    file are refused (the reference's own T_hist output indexes point 0 of the
    list); T_hist with Volume or Surface tracking on a grid is accepted."""
    d = decks.points_small()
    pi = d.write(str(tmp_path / "pts"))
    th = pkg.Thesis(pi)
    assert th.info.pnum == 400
    sim = of.Sim(pi)
    xyz = np.array([sim.xyz(p) for p in range(400)])
    assert np.array_equal(xyz, d.points / 1000.0)
    open(tmp_path / "pts" / "Points.txt", "w").write("x y z\n")
    _expect_error(pi, 1, "No points found")
    g = decks.snapvol_small()
    g.output["Temperature"]["T_hist"] = 1
    pkg.Thesis(g.write(str(tmp_path / "grid"))).close()
    k = decks.stork_small()
    k.output["Temperature"]["T_hist"] = 1
    ki = k.write(str(tmp_path / "stork"))
    _expect_error(ki, 1, "T_hist with Stork tracking on a grid")
    assert "T_hist with Stork tracking on a grid" in of.Sim(ki).error
