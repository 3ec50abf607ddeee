"""Pins for the FP64 CPU oracle (oracle/oracle.c).

Each function of the oracle is checked against something it did not produce:
  * the reference's own output on the same decks (golden fixtures written by
    oracle/_ref, i.e. the reference library compiled in place): node lists,
    T fields and whole mode runs must be bit-identical;
  * closed forms of the underlying heat-conduction integral (stationary
    isotropic Gaussian spot) and the energy balance of the half-space solution;
  * finite differences for the gradient / Laplacian / cooling-rate accumulators;
  * the closure property of the melt-pool flood fill.
"""
from __future__ import annotations

import gzip
import hashlib
import math
import os

import numpy as np
import pytest

import decks
import oracle_ffi as of
from conftest import check_example_outputs, example_deck
from make_golden import CLANG_RUN_CASES, EXAMPLE_CASES, NODE_CASES, RUN_CASES, TGRID_CASES, scan_end


@pytest.mark.parametrize("case", [c[0] for c in NODE_CASES])
def test_nodes_bit_identical_to_reference(case, golden, tmp_path):
    m = golden["meta"]["nodes"][case]
    deck = decks.WORKLOADS[m["workload"]]()
    sim = of.Sim(deck.write(str(tmp_path / case)))
    got = sim.nodes(m["t"], bool(m["is_sol"]))
    ref = golden["nodes"][case].copy()
    assert got.shape == ref.shape
    # A merged node of the compressed quadrature takes zb from a field the
    # reference never initialises (Calc.cpp:452); its build printed 0 or
    # subnormal stack residue there. The oracle writes 0; a subnormal z offset
    # leaves every temperature unchanged (dz*dz underflows, zp - zb == zp).
    junk = (ref[:, 2] != 0) & (np.abs(ref[:, 2]) < np.finfo(np.float64).tiny)
    if case.startswith("comp"):
        ref[junk, 2] = 0.0
    else:
        assert not junk.any()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", [c[0] for c in TGRID_CASES])
def test_snapshot_T_bit_identical_to_reference(case, golden, tmp_path):
    m = golden["meta"]["tgrid"][case]
    deck = decks.WORKLOADS[m["workload"]]()
    sim = of.Sim(deck.write(str(tmp_path / case)))
    T, nn = sim.snapshot_T(m["t"])
    assert nn == m["nodes"]
    assert np.array_equal(T, golden["tgrid"][case])


@pytest.mark.parametrize("wl", RUN_CASES)
def test_mode_runs_identical_to_reference(wl, golden, tmp_path):
    meta = golden["meta"]["runs"][wl]
    deck = decks.WORKLOADS[wl]()
    d = tmp_path / wl
    sim = of.Sim(deck.write(str(d)))
    os.makedirs(d / "full", exist_ok=True)
    sim.run(str(d / "full"))
    runs = golden["runs"]
    residue = wl.startswith("compress")
    for tag, fname in meta["files"].items():
        ref_csv = gzip.decompress(runs[f"{wl}/{tag}/csv"].tobytes()).decode()
        got_csv = open(d / "Data" / fname).read()
        if residue:
            _assert_csv_equal_but_residue(got_csv, ref_csv, fname)
        else:
            assert got_csv == ref_csv, f"{fname} differs"
    if "rrdf_sha256" in meta:  # Stork writes its records instead of the grid table
        text = open(d / "full" / f"{deck.name}.RRDF.full.txt", "rb").read()
        assert text.count(b"\n") == meta["rrdf_records"]
        assert hashlib.sha256(text).hexdigest() == meta["rrdf_sha256"]
        return
    # full-precision final state
    fulls = [f for f in sorted(os.listdir(d / "full")) if not any(k in f for k in (".T_hist.", ".RDF.", ".RRDF."))]
    p, v = of.read_full(str(d / "full" / fulls[-1]))
    assert np.array_equal(p, runs[f"{wl}/final/p"])
    ref_v = runs[f"{wl}/final/v"].copy()
    if residue:
        ref_v[_residue(ref_v) & (v == 0)] = 0.0
    assert np.array_equal(v, ref_v)
    if f"{wl}/rdf/p" in runs:  # every RDF melt event
        rp, rv = of.read_full(str(d / "full" / f"{deck.name}.RDF.full.txt"))
        assert np.array_equal(rp, runs[f"{wl}/rdf/p"])
        assert np.array_equal(rv, runs[f"{wl}/rdf/v"])
    if f"{wl}/thist/p" in runs:  # every recorded (t, T) of the T_hist output
        hp, hv = of.read_full(str(d / "full" / f"{deck.name}.T_hist.full.txt"))
        assert np.array_equal(hp, runs[f"{wl}/thist/p"])
        assert np.array_equal(hv, runs[f"{wl}/thist/v"])


@pytest.mark.parametrize("wl", CLANG_RUN_CASES)
def test_fabs_mode_runs_identical_to_clang_reference(wl, golden, tmp_path):
    """abs_mode 1 (the floating abs() that clang++ binds at Grid.cpp:381 and
    :535: the regula falsi iterates to Sol_Tol, dTdt is not truncated) against
    the reference built with clang++ (oracle/_ref_clang)."""
    meta = golden["meta"]["runs_clang"][wl]
    deck = decks.WORKLOADS[wl]()
    d = tmp_path / wl
    sim = of.Sim(deck.write(str(d)), abs_mode=1)
    os.makedirs(d / "full", exist_ok=True)
    sim.run(str(d / "full"))
    runs = golden["runs_clang"]
    residue = wl.startswith("compress")
    for tag, fname in meta["files"].items():
        ref_csv = gzip.decompress(runs[f"{wl}/{tag}/csv"].tobytes()).decode()
        got_csv = open(d / "Data" / fname).read()
        if residue:
            _assert_csv_equal_but_residue(got_csv, ref_csv, fname)
        else:
            assert got_csv == ref_csv, f"{fname} differs"
    fulls = [f for f in sorted(os.listdir(d / "full")) if not any(k in f for k in (".T_hist.", ".RDF.", ".RRDF."))]
    p, v = of.read_full(str(d / "full" / fulls[-1]))
    assert np.array_equal(p, runs[f"{wl}/final/p"])
    ref_v = runs[f"{wl}/final/v"].copy()
    if residue:
        ref_v[_residue(ref_v) & (v == 0)] = 0.0
    assert np.array_equal(v, ref_v)
    if f"{wl}/rdf/p" in runs:
        rp, rv = of.read_full(str(d / "full" / f"{deck.name}.RDF.full.txt"))
        assert np.array_equal(rp, runs[f"{wl}/rdf/p"])
        assert np.array_equal(rv, runs[f"{wl}/rdf/v"])


@pytest.mark.parametrize("name", EXAMPLE_CASES)
def test_reference_examples_identical(name, golden, tmp_path):
    """The reference's own example decks (examples/, inputs committed under
    tests/golden/examples) through the oracle, against the reference's outputs
    for them: row counts, row-set digests and sampled rows, bit for bit; Stork's
    RRDF records by the digest of their full-precision text."""
    pi = example_deck(name, str(tmp_path / name))
    d = os.path.dirname(pi)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        sim = of.Sim("ParamInput.txt")
        os.makedirs("full", exist_ok=True)
        sim.run(os.path.join(d, "full"))
    finally:
        os.chdir(cwd)
    check_example_outputs(name, golden, os.path.join(d, "Data"), os.path.join(d, "full"), exact=True)
    meta = golden["meta"]["examples"][name]
    if "rrdf_sha256" in meta:
        text = open(os.path.join(d, "full", [f for f in os.listdir(os.path.join(d, "full"))
                                              if f.endswith(".RRDF.full.txt")][0]), "rb").read()
        assert text.count(b"\n") == meta["rrdf_records"]
        assert hashlib.sha256(text).hexdigest() == meta["rrdf_sha256"]


def _residue(a):
    """Values below 1e-280: what the uninitialised zb of a merged compressed
    node (Calc.cpp:452, subnormal stack residue in the reference build) leaves
    in Gz at surface points, where dz = zp - zb = -zb. A physical Gz is 0 or
    many orders of magnitude larger."""
    return (a != 0) & (np.abs(a) < 1e-280)


def _assert_csv_equal_but_residue(got, ref, fname):
    g_rows, r_rows = got.split("\n"), ref.split("\n")
    assert len(g_rows) == len(r_rows), fname
    for g, r in zip(g_rows, r_rows):
        if g == r:
            continue
        gc, rc = g.split(","), r.split(",")
        assert len(gc) == len(rc), fname
        for a, b in zip(gc, rc):
            if a != b:
                assert float(a) == 0.0 and _residue(np.array([float(b)]))[0], (fname, a, b)


def _spot_deck(a0, T_spot, P=300.0, eff=0.35):
    return decks.Deck(
        name="spot", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[dict(Width_X=a0, Width_Y=a0, Depth_Z=a0, Power=P, Efficiency=eff)],
        paths=[[(1, 0.0, 0.0, 0.0, 1.0, T_spot)]],
        domain={"X": decks.axis(-2e-4, 2e-4, 9), "Y": decks.axis(0.0, 0.0, 1),
                "Z": decks.axis(0.0, 0.0, 1)},
        output=decks.out_cols("T"))


def test_closed_form_stationary_isotropic_spot(tmp_path):
    """dT(r,t) = beta/(12 a) sqrt(pi/(3 r^2)) [erf(sqrt(3) r/a0) - erf(sqrt(3 r^2/(a0^2+12 a t)))],
    with beta = (3/pi)^1.5 * 2 P eff / (rho c); at r=0 beta/(6a) (1/a0 - 1/sqrt(a0^2+12 a t))."""
    a0, Ts = 50e-6, 2e-3
    deck = _spot_deck(a0, Ts)
    sim = of.Sim(deck.write(str(tmp_path / "spot")))
    m = decks.MATERIAL
    alpha = m["k"] / (m["p"] * m["c"])
    beta = (3 / math.pi) ** 1.5 * (2 * 300.0 * 0.35) / (m["p"] * m["c"])
    T, _ = sim.snapshot_T(Ts)
    for p in range(sim.s.pnum):
        x, _, _ = sim.xyz(p)
        r = abs(x)
        if r == 0:
            dT = beta / (6 * alpha) * (1 / a0 - 1 / math.sqrt(a0 ** 2 + 12 * alpha * Ts))
        else:
            dT = beta / (12 * alpha) * math.sqrt(math.pi / (3 * r * r)) * (
                math.erf(math.sqrt(3) * r / a0) - math.erf(math.sqrt(3 * r * r / (a0 ** 2 + 12 * alpha * Ts))))
        assert abs((T[p] - m["T_0"]) - dT) <= 2e-6 * dT, (p, r, T[p] - m["T_0"], dT)


def test_energy_balance_half_space(tmp_path):
    """rho c * integral over z<=0 of dT = P eff * (powered time): pins beta, the
    half-space factor 2 and the 1/phi convention of the node sum."""
    deck = decks.Deck(
        name="energy", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[dict(decks.BEAM)], paths=[[(0, 0.4, 0.0, 0.0, 1.0, 1.0)]],
        domain={"X": decks.axis(-1.6e-3, 2.0e-3, 145), "Y": decks.axis(-1.6e-3, 1.6e-3, 129),
                "Z": decks.axis(-1.6e-3, 0.0, 65)},
        output=decks.out_cols("T"))
    sim = of.Sim(deck.write(str(tmp_path / "energy")))
    t = 2e-3  # 0.4 ms of heating, 1.6 ms of diffusion
    T, _ = sim.snapshot_T(t, nthreads=4)
    dT = (T - decks.MATERIAL["T_0"]).reshape(145, 129, 65)
    hx, hy, hz = 3.6e-3 / 144, 3.2e-3 / 128, 1.6e-3 / 64
    integ = np.trapezoid(np.trapezoid(np.trapezoid(dT, dx=hz, axis=2), dx=hy, axis=1), dx=hx)
    m = decks.MATERIAL
    energy = m["p"] * m["c"] * integ
    expect = 300.0 * 0.35 * 0.4e-3
    assert abs(energy - expect) < 2e-3 * expect, (energy, expect)


def test_quadrature_covers_powered_time(golden):
    """Sum of node weights equals the powered time [0, t] (8-digit weights)."""
    for case in ("cfg1_end", "cfg1_mid", "cfg1_sol", "cfg2_mid"):
        m = golden["meta"]["nodes"][case]
        arr = golden["nodes"][case]
        if case.startswith("cfg1"):
            powered = m["t"]
            assert abs(arr[:, 6].sum() - powered) < 1e-7 * powered


def test_gradient_and_laplacian_match_finite_differences(golden, tmp_path):
    deck = decks.cfg1()
    sim = of.Sim(deck.write(str(tmp_path / "fd")))
    arr = golden["nodes"]["cfg1_mid"]
    xb = arr[:, 0][arr[:, 6] == 0][0]
    for (x, y, z) in [(xb - 60e-6, 20e-6, -10e-6), (xb - 150e-6, -35e-6, -40e-6)]:
        g = of.sol_primary(arr, x, y, z)
        h = 1e-7
        f = lambda *q: of.calc_T_xyz(sim, arr, *q)
        c = f(x, y, z)
        fd = [(f(x + h, y, z) - f(x - h, y, z)) / (2 * h), (f(x, y + h, z) - f(x, y - h, z)) / (2 * h),
              (f(x, y, z + h) - f(x, y, z - h)) / (2 * h)]
        lap = sum((f(*(np.array([x, y, z]) + s * h * e)) - 2 * c + f(*(np.array([x, y, z]) - s * h * e))) / h ** 2
                  for e in np.eye(3) for s in [1])
        G = math.sqrt(sum(v * v for v in g[:3]))
        for k in range(3):
            assert abs(g[k] - fd[k]) < 1e-5 * G
        assert abs(g[3] - lap) < 2e-3 * abs(g[3])


def test_cooling_rate_is_heat_equation_time_derivative(tmp_path):
    """alpha*Laplacian + S (instantaneous source) equals dT/dt at a fixed point
    (points away from the 1 um-deep source, where the standard quadrature resolves it)."""
    deck = decks.cfg1()
    sim = of.Sim(deck.write(str(tmp_path / "dtdt")))
    alpha = decks.MATERIAL["k"] / (decks.MATERIAL["p"] * decks.MATERIAL["c"])
    cases = [(1.1e-3, (0.9e-3, 30e-6, -20e-6)), (1.1e-3, (0.8e-3, 20e-6, -40e-6)),
             (1.1e-3, (0.9e-3, 0.0, -60e-6)), (2.5e-3, (1.06e-3, 10e-6, 0.0)),
             (2.5e-3, (0.8e-3, 20e-6, -40e-6))]
    for t, (x, y, z) in cases:
        h = 2e-7
        Tp = of.calc_T_xyz(sim, sim.nodes(t + h, hint=[1]), x, y, z)
        Tm = of.calc_T_xyz(sim, sim.nodes(t - h, hint=[1]), x, y, z)
        g = of.sol_primary(sim.nodes(t, hint=[1]), x, y, z)
        rate = alpha * g[3] + g[4]
        fd = (Tp - Tm) / (2 * h)
        assert abs(rate - fd) < 5e-5 * abs(fd), (t, x, y, z, rate, fd)


def test_flood_fill_closure(tmp_path):
    """Snapshot Volume tracking: every output point is liquid and every liquid
    point in the BFS neighbourhood (3x3x{k-1,k}) of an output point is output."""
    deck = decks.snapvol_small()
    pi = deck.write(str(tmp_path / "closure"))
    sim = of.Sim(pi)
    os.makedirs(tmp_path / "closure" / "full", exist_ok=True)
    sim.run(str(tmp_path / "closure" / "full"))
    nx, ny, nz = sim.s.xnum, sim.s.ynum, sim.s.znum
    end = scan_end(deck)
    for s, frac in enumerate([0.4, 0.7, 1.0]):
        sim2 = of.Sim(pi)
        T, _ = sim2.snapshot_T(frac * end)
        liq = (T >= decks.MATERIAL["T_L"]).reshape(nx, ny, nz)
        rows = np.loadtxt(tmp_path / "closure" / "Data" / f"snapv.Snapshot.{s:02d}.csv", delimiter=",",
                          skiprows=1, ndmin=2)
        out = np.zeros_like(liq)
        for x, y, z, _T in rows:
            i = int(round(x / 2e-5)); j = int(round(y / 2e-5)); k = int(round((z + 1e-4) / 2e-5))
            out[i, j, k] = True
        assert out.any()
        assert not (out & ~liq).any()
        for i, j, k in zip(*np.nonzero(out)):
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    for dk in (-1, 0):
                        a, b, c = i + di, j + dj, k + dk
                        if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz and liq[a, b, c]:
                            assert out[a, b, c]


def test_fabs_mode_iterates_regula_falsi(tmp_path):
    """clang semantics (fabs) keep tSol inside [t-dt, t] and change only tSol/V/dTdt."""
    deck = decks.none_small()
    d = tmp_path / "fabs"
    pi = deck.write(str(d))
    a = of.Sim(pi, abs_mode=0)
    os.makedirs(d / "fa", exist_ok=True)
    a.run(str(d / "fa"))
    b = of.Sim(pi, abs_mode=1)
    os.makedirs(d / "fb", exist_ok=True)
    b.run(str(d / "fb"))
    pa, va = of.read_full(str(d / "fa" / "nones.Solidification.Final.full.txt"))
    pb, vb = of.read_full(str(d / "fb" / "nones.Solidification.Final.full.txt"))
    assert np.array_equal(pa, pb)
    assert np.allclose(va[:, 4], vb[:, 4], rtol=0, atol=2e-5)  # tSol within one step
    dtdt_a = va[:, 10]
    assert np.all(dtdt_a == np.trunc(dtdt_a))  # int-truncated under g++ semantics
