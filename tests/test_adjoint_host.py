"""CPU check of the design-gradient chain (design.cpp) against central finite
differences of the FP64 oracle.

The node sensitivities S (the GPU adjoint kernel's output) are formed here in
numpy from the FP64 node list; the library then chains them through the
quadrature schedule with branch decisions frozen. The finite differences
re-parse perturbed decks and evaluate them with the oracle, so they share
nothing with the chain but the physics.
"""
from __future__ import annotations

import importlib

import numpy as np
import pytest

import decks
import oracle_ffi as of

pkg = importlib.import_module("3dthesis_amd")


def _deck(beam, rows, n=(20, 20, 6)):
    b = dict(Width_X=beam[2], Width_Y=beam[3], Depth_Z=beam[4], Power=beam[0], Efficiency=beam[1])
    return decks.Deck(
        name="adj", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[b], paths=[[tuple(r) for r in rows]],
        domain={"X": decks.axis(0.0, 1e-3, n[0]), "Y": decks.axis(0.0, 1e-3, n[1]),
                "Z": decks.axis(-2e-4, 0.0, n[2])},
        output=decks.out_cols("T"))


def _rows(rng):
    base = decks.serpentine(0.3, 0.7, 0.4, 0.15, 2, 1.0, pieces=3)
    out = []
    for r in base:
        d = rng.normal(0.0, 0.01, size=2)
        out.append([r[0], r[1] + d[0], r[2] + d[1], r[3], r[4], r[5]])
    return np.array(out)


def _T_oracle(tmp_path, tag, beam, rows):
    pi = _deck(beam, rows).write(str(tmp_path / tag))
    sim = of.Sim(pi)
    T, _ = sim.snapshot_T(float(sim.s.all_end))
    return T, sim


def _sensitivities(nodes, xyz, w):
    """S (n x 7) of L for w = dL/dT at the points xyz, FP64."""
    xb, yb, zb, ix, iy, iz, dtau, em = (nodes[:, c][:, None] for c in range(8))
    dx = xyz[None, :, 0] - xb
    dy = xyz[None, :, 1] - yb
    dz = xyz[None, :, 2] - zb
    E = dtau * np.exp(-3.0 * (dx * dx * ix + dy * dy * iy + dz * dz * iz) + em)
    wE = w[None, :] * E
    return np.stack([(wE * 6 * dx * ix).sum(1), (wE * 6 * dy * iy).sum(1), (wE * 6 * dz * iz).sum(1),
                     (wE * (3 * dx * dx * ix * ix - 0.5 * ix)).sum(1), (wE * (3 * dy * dy * iy * iy - 0.5 * iy)).sum(1),
                     (wE * (3 * dz * dz * iz * iz - 0.5 * iz)).sum(1), wE.sum(1)], axis=1)


def test_design_gradient_matches_finite_differences(tmp_path):
    rng = np.random.default_rng(11)
    beam_true = np.array([300.0, 0.35, 50e-6, 50e-6, 1e-6])
    rows_true = _rows(rng)
    beam = beam_true * (1 + rng.uniform(-0.1, 0.1, 5))
    rows = _rows(rng)
    T_target, _ = _T_oracle(tmp_path, "target", beam_true, rows_true)

    pi = _deck(beam, rows).write(str(tmp_path / "design"))
    th = pkg.Thesis(pi)
    b0, r0 = th.design()
    assert np.allclose(b0, beam, rtol=1e-15) and np.allclose(r0, rows, rtol=0, atol=1e-15)
    T, sim = _T_oracle(tmp_path, "base", beam, rows)
    nodes = th.design_nodes()
    assert np.array_equal(nodes, sim.nodes(float(sim.s.all_end)))

    N = T.size
    w = 2.0 * (T - T_target) / N
    xyz = np.array([sim.xyz(p) for p in range(N)])
    S = _sensitivities(nodes, xyz, w)
    g = th.design_chain(S)
    assert g.shape == (5 + 2 * rows.shape[0],)

    def loss(bm, rw, tag):
        Tp, _ = _T_oracle(tmp_path, tag, bm, rw)
        return float(np.mean((Tp - T_target) ** 2))

    # invariant: P enters only through P*eff
    assert abs(beam[0] * g[0] - beam[1] * g[1]) <= 1e-9 * abs(beam[0] * g[0])
    checks = list(range(5)) + [5 + 2 * r + c for r in range(rows.shape[0]) for c in (0, 1)]
    for k in checks:
        bp, bm_, rp, rm = beam.copy(), beam.copy(), rows.copy(), rows.copy()
        if k < 5:
            h = 1e-6 * abs(beam[k])
            bp[k] += h
            bm_[k] -= h
        else:
            r, c = divmod(k - 5, 2)
            h = 1e-6  # mm (1 nm)
            rp[r, 1 + c] += h
            rm[r, 1 + c] -= h
        fd = (loss(bp, rp, f"p{k}") - loss(bm_, rm, f"m{k}")) / (2 * h)
        scale = max(abs(fd), 1e-3 * np.max(np.abs(g[:5] * beam)) / max(abs(beam[min(k, 4)]), 1e-30)
                    if k < 5 else abs(fd))
        assert abs(g[k] - fd) <= 2e-5 * max(abs(fd), 1e-12) + 1e-12 * scale, (k, g[k], fd)
