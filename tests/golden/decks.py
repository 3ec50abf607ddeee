"""Input-deck writer for the 3DThesis file format, and the named workloads.

A deck is the set of text files the reference reads (ParamInput, Mode, Material,
Beam[.N], Path[.N], Domain, Output, Settings; formats per Init.cpp:233-814 of the
reference). Every workload used by the tests, the bench and the fixtures is
generated here from its parameters and seeds, so the decks themselves never
need to be committed.

Workloads (sizes in metres unless noted):
  cfg1  Snapshots/None, one 2 mm line at 1 m/s, 200x50x20 grid
  cfg2  Solidification/Surface, 20-hatch 2 mm serpentine, dt 1e-5, 400x400x40
  cfg3  Solidification/Volume, 4 seeded beams/rasters, 500x500x50
  cfg4  256 seeded snapshot parameter sets, 128x128x32 each (batched API)
  cfg5  adjoint workload: 200-vertex perturbed raster, 256x256x64
The *_small variants keep the physics and shrink grid and path so the
reference finishes in seconds.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Common material of the workloads: T0=300 K, TL=1600 K, k=26.6, c=600, rho=7800.
MATERIAL = dict(T_0=300.0, T_L=1600.0, k=26.6, c=600.0, p=7800.0)
BEAM = dict(Width_X=50e-6, Width_Y=50e-6, Depth_Z=1e-6, Power=300.0, Efficiency=0.35)

Row = Tuple[int, float, float, float, float, float]  # mode, x mm, y mm, z mm, pmod, speed|time


@dataclass
class Deck:
    name: str
    mode: Dict[str, str]                      # e.g. {"Snapshots": {...}} flattened below
    mode_block: str                           # "Snapshots" or "Solidification"
    beams: List[Dict[str, float]]
    paths: List[List[Row]]
    domain: Dict[str, Dict[str, float]]
    output: Dict[str, Dict[str, int]]
    material: Dict[str, float] = field(default_factory=lambda: dict(MATERIAL))
    settings: Dict[str, Dict[str, float]] = field(default_factory=lambda: {"Compute": {"MaxThreads": 1}})
    cet: Optional[Dict[str, float]] = None
    points: Optional[np.ndarray] = None       # custom points (mm), n x 3: Domain Custom File

    def write(self, d: str) -> str:
        """Write the deck into directory d; returns the ParamInput path."""
        os.makedirs(d, exist_ok=True)
        multi = len(self.beams) > 1
        beam_file = "Beam.*.txt" if multi else "Beam.txt"
        path_file = "Path.*.txt" if multi else "Path.txt"
        with open(os.path.join(d, "ParamInput.txt"), "w") as f:
            f.write("Simulation\n{\n")
            f.write(f"\tName\t\t{self.name}\n")
            f.write(f"\tMode\t\t{d}/Mode.txt\n")
            f.write(f"\tMaterial\t{d}/Material.txt\n")
            f.write(f"\tBeam\t\t{d}/{beam_file}\n")
            f.write(f"\tPath\t\t{d}/{path_file}\n")
            f.write("}\nOptions\n{\n")
            f.write(f"\tDomain\t\t{d}/Domain.txt\n")
            f.write(f"\tOutput\t\t{d}/Output.txt\n")
            f.write(f"\tSettings\t{d}/Settings.txt\n")
            f.write("}\n")
        _block_file(os.path.join(d, "Mode.txt"), {self.mode_block: self.mode})
        mat = {"Constants": self.material}
        if self.cet:
            mat["CET"] = self.cet
        _block_file(os.path.join(d, "Material.txt"), mat)
        for b, beam in enumerate(self.beams):
            fn = f"Beam.{b + 1}.txt" if multi else "Beam.txt"
            _block_file(os.path.join(d, fn), {
                "Shape": {k: beam[k] for k in ("Width_X", "Width_Y", "Depth_Z")},
                "Intensity": {k: beam[k] for k in ("Power", "Efficiency")}})
        for b, path in enumerate(self.paths):
            fn = f"Path.{b + 1}.txt" if multi else "Path.txt"
            with open(os.path.join(d, fn), "w") as f:
                f.write("Mode\tX(mm)\tY(mm)\tZ(mm)\tPmod\tParam\n")
                for r in path:
                    f.write("%d\t%s\t%s\t%s\t%s\t%s\n" % (r[0], _num(r[1]), _num(r[2]), _num(r[3]),
                                                           _num(r[4]), _num(r[5])))
        domain = dict(self.domain)
        if self.points is not None:
            fn = os.path.join(d, "Points.txt")
            with open(fn, "w") as f:
                f.write("X(mm)\tY(mm)\tZ(mm)\n")
                for x, y, z in self.points:
                    f.write("%s\t%s\t%s\n" % (_num(float(x)), _num(float(y)), _num(float(z))))
            domain["Custom"] = {"File": fn}
        _block_file(os.path.join(d, "Domain.txt"), domain)
        _block_file(os.path.join(d, "Output.txt"), self.output)
        _block_file(os.path.join(d, "Settings.txt"), self.settings)
        return os.path.join(d, "ParamInput.txt")


def _num(v) -> str:
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def _block_file(path: str, blocks: Dict[str, Dict[str, object]]) -> None:
    with open(path, "w") as f:
        for name, kv in blocks.items():
            f.write(f"{name}\n{{\n")
            for k, v in kv.items():
                f.write(f"\t{k}\t{v if isinstance(v, str) else _num(v)}\n")
            f.write("}\n")


def out_cols(*cols: str, xyz: bool = True) -> Dict[str, Dict[str, int]]:
    """Output.txt blocks with exactly the named columns switched on."""
    sol = ["tSol", "G", "Gx", "Gy", "Gz", "V", "dTdt", "eqFrac", "depth", "numMelt", "RDF", "MP_Stats"]
    return {
        "Grid": {"x": int(xyz), "y": int(xyz), "z": int(xyz)},
        "Temperature": {"T": int("T" in cols), "T_hist": 0},
        "Solidification": {k: int(k in cols) for k in sol},
    }


def axis(lo: float, hi: float, num: int) -> Dict[str, float]:
    return {"Min": lo, "Max": hi, "Num": num}


def serpentine(x_a: float, x_b: float, y0: float, dy: float, nh: int, speed: float,
               jump: float = 1e-4, pieces: int = 1) -> List[Row]:
    """Raster of nh hatches (mm) alternating x_a->x_b and back, each preceded by a
    zero-power dwell of `jump` seconds at its start point."""
    rows: List[Row] = []
    for h in range(nh):
        ys = y0 + dy * h
        xs, xe = (x_a, x_b) if h % 2 == 0 else (x_b, x_a)
        rows.append((1, xs, ys, 0.0, 0.0, jump))
        for k in range(1, pieces + 1):
            rows.append((0, xs + (xe - xs) * k / pieces, ys, 0.0, 1.0, speed))
    return rows


def point_raster(x0: float, y0: float, nx: int, ny: int, step: float, dwell: float, pmod: float = 1.0) -> List[Row]:
    """Spots on an nx x ny raster (mm), row by row, each held `dwell` seconds."""
    rows: List[Row] = []
    for j in range(ny):
        for i in range(nx):
            ii = i if j % 2 == 0 else nx - 1 - i
            rows.append((1, x0 + step * ii, y0 + step * j, 0.0, pmod, dwell))
    return rows


def compress_small(mode: str = "Snapshots", tracking: str = "None") -> Deck:
    """Path compression (Settings Path Compression 1, Calc.cpp:213-501): a
    point raster of 12x6 spots 15 um apart, a powered line and a second raster
    at half power, over a 0.6x0.4x0.1 mm box (20 um spacing for snapshots,
    10 um for solidification)."""
    path = point_raster(0.20, 0.15, 12, 6, 0.015, 2e-5)
    path += [(0, 0.40, 0.15, 0.0, 1.0, 0.8), (0, 0.40, 0.25, 0.0, 1.0, 0.8)]
    path += point_raster(0.38, 0.28, 8, 4, 0.012, 1.5e-5, pmod=0.5)
    mode_d = {"Tracking": tracking}
    if mode == "Snapshots":
        mode_d["Times"] = "0.0005,0.001,0.002"
        n = (31, 21, 6)
    else:
        mode_d["Timestep"] = 1e-5
        n = (61, 41, 11)
    d = Deck(name="comps", mode_block=mode, mode=mode_d, beams=[dict(BEAM)], paths=[path],
             domain={"X": axis(0.1e-3, 0.7e-3, n[0]), "Y": axis(0.05e-3, 0.45e-3, n[1]),
                     "Z": axis(-1e-4, 0.0, n[2])},
             output=out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "numMelt") if mode != "Snapshots"
             else out_cols("T"))
    d.settings = {"Compute": {"MaxThreads": 1}, "Path": {"Compression": 1}}
    return d


def walls_small(mode: str = "Snapshots", tracking: str = "None", reflections: int = 2) -> Deck:
    """Wall images (Domain BoundaryConditions, Calc.cpp:503-608): a line scan
    beside insulated walls at X_min, Y_max and Z_min, reflected `reflections`
    times."""
    path = serpentine(0.05, 0.35, 0.17, 0.02, 2, 1.0)
    mode_d = {"Tracking": tracking}
    if mode == "Snapshots":
        mode_d["Times"] = "0.0002,0.0005,0.0008"
    else:
        mode_d["Timestep"] = 1e-5
    d = Deck(name="walls", mode_block=mode, mode=mode_d, beams=[dict(BEAM)], paths=[path],
             domain={"X": axis(0.0, 0.4e-3, 41), "Y": axis(0.0, 0.2e-3, 21), "Z": axis(-0.1e-3, 0.0, 11),
                     "BoundaryConditions": {"X_min": 0.0, "Y_max": 0.2e-3, "Z_min": -0.1e-3,
                                            "Reflections": reflections}},
             output=out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "numMelt") if mode != "Snapshots"
             else out_cols("T"))
    return d

def probe_points(n: int, seed: int, lo=(0.0, -0.1, -0.08), hi=(0.5, 0.1, 0.0)) -> np.ndarray:
    """n seeded probe points (mm) in a box, a quarter of them on the top surface."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(lo, hi, size=(n, 3))
    pts[: n // 4, 2] = hi[2]
    return np.round(pts, 6)


def points_small() -> Deck:
    """Snapshots at custom probe points (Domain Custom File, Init.cpp:816-857)
    with the temperature history output (T_hist, Grid.cpp:101-174)."""
    d = Deck(name="pts", mode_block="Snapshots", mode={"Times": "0.0001,0.0002,0.0004,0.0007", "Tracking": "None"},
             beams=[dict(BEAM)], paths=[[(0, 0.5, 0.0, 0.0, 1.0, 1.0)]],
             domain={}, output=out_cols("T"), points=probe_points(400, 7100))
    d.output["Temperature"]["T_hist"] = 1
    return d


def points_small_sol() -> Deck:
    """Solidification at custom probe points. The Surface keyword is kept:
    custom points always run the untracked driver (Run.cpp:391) but the
    keyword still selects the per-point T_hist files and the depth column."""
    d = Deck(name="ptss", mode_block="Solidification", mode={"Tracking": "Surface", "Timestep": 2e-5},
             beams=[dict(BEAM)], paths=[[(0, 0.4, 0.0, 0.0, 1.0, 1.0)]],
             domain={"X": {"Min": 0.0, "Max": 0.5e-3}, "Y": {"Min": -1e-4, "Max": 1e-4},
                     "Z": {"Min": -8e-5, "Max": 0.0}},
             output=out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "depth", "numMelt"),
             points=probe_points(300, 7200, lo=(0.0, -0.06, -0.05), hi=(0.45, 0.06, 0.0)))
    d.output["Temperature"]["T_hist"] = 1
    return d


def thist_none_small() -> Deck:
    """Untracked Solidification on a grid with the one-table T_hist output."""
    d = none_small()
    d.name = "thist"
    d.output["Temperature"]["T_hist"] = 1
    return d


def rdf_small(tracking: str = "Surface") -> Deck:
    """The ExaCA coupling outputs on a small Surface run: RDF melt/liquidus/
    cooling-rate events (Grid.cpp:176-201), melt-pool statistics (MP_Stats,
    Melt.cpp:359-538) and numMelt, which with RDF counts recorded melt times."""
    d = cfg2_small(tracking)
    d.name = "rdf" if tracking == "Surface" else "rdfv"
    d.paths = [serpentine(0.3, 0.7, 0.3, 0.06, 3, 1.0)]
    d.output = out_cols("T", "tSol", "G", "V", "dTdt", "depth", "numMelt", "RDF", "MP_Stats")
    return d


def rdf_small_windowed() -> Deck:
    """RDF events without the melt-pool statistics: a Surface run whose melt
    times the windowed replay records (MP_Stats takes the per-step driver)."""
    d = rdf_small("Surface")
    d.name = "rdfw"
    d.output = out_cols("T", "tSol", "G", "V", "dTdt", "depth", "numMelt", "RDF")
    return d


def stork_small() -> Deck:
    """Stork tracking (Run.cpp:711-1079): the reduced RRDF records of every
    grid cell whose liquid-corner count changes."""
    d = cfg2_small("Stork")
    d.name = "stork"
    d.paths = [serpentine(0.3, 0.5, 0.35, 0.06, 1, 1.0)]
    d.domain = {"X": axis(0.2e-3, 0.6e-3, 21), "Y": axis(0.25e-3, 0.45e-3, 11), "Z": axis(-1e-4, 0.0, 6)}
    d.output = out_cols("T")
    return d


# ----------------------------------------------------------------------------
# named workloads

def cfg1(nx=200, ny=50, nz=20) -> Deck:
    return Deck(
        name="cfg1", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[dict(BEAM)], paths=[[(0, 2.0, 0.0, 0.0, 1.0, 1.0)]],
        domain={"X": axis(0.0, 2e-3, nx), "Y": axis(-2.5e-4, 2.5e-4, ny), "Z": axis(-2e-4, 0.0, nz)},
        output=out_cols("T"))


def cfg2(nx=400, ny=400, nz=40, nh=20, x_a=1.0, x_b=3.0, y0=1.05, lo=(0.0, 0.0, -4e-4),
         hi=(4e-3, 4e-3, 0.0), dt=1e-5, tracking="Surface") -> Deck:
    return Deck(
        name="cfg2", mode_block="Solidification",
        mode={"Tracking": tracking, "Timestep": dt},
        beams=[dict(BEAM)], paths=[serpentine(x_a, x_b, y0, 0.1, nh, 1.0)],
        domain={"X": axis(lo[0], hi[0], nx), "Y": axis(lo[1], hi[1], ny), "Z": axis(lo[2], hi[2], nz)},
        output=out_cols("tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt"))


def cfg2_small(tracking="Surface") -> Deck:
    """3 hatches of 0.6 mm over a 1.2x0.8x0.2 mm box at 20 um spacing."""
    d = cfg2(nx=61, ny=41, nz=11, nh=3, x_a=0.3, x_b=0.9, y0=0.3, lo=(0.0, 0.0, -2e-4),
             hi=(1.2e-3, 0.8e-3, 0.0), tracking=tracking)
    d.name = "cfg2s"
    d.output = out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "numMelt", "depth")
    return d


def cfg3_rasters(nbeams=4, box_mm=5.0) -> Tuple[List[Dict[str, float]], List[List[Row]]]:
    beams, paths = [], []
    for b in range(1, nbeams + 1):
        rng = np.random.default_rng(3000 + b)
        nh = int(rng.integers(5, 31))
        L = float(rng.uniform(1.0, 3.0))
        P = float(rng.uniform(200.0, 400.0))
        v = float(rng.uniform(0.5, 2.0))
        x0 = float(rng.uniform(0.0, box_mm - L))
        y0 = float(rng.uniform(0.0, box_mm - 0.1 * (nh - 1)))
        beam = dict(BEAM)
        beam["Power"] = P
        beams.append(beam)
        paths.append(serpentine(x0, x0 + L, y0, 0.1, nh, v))
    return beams, paths


def cfg3(nx=500, ny=500, nz=50) -> Deck:
    beams, paths = cfg3_rasters()
    return Deck(
        name="cfg3", mode_block="Solidification", mode={"Tracking": "Volume", "Timestep": 1e-5},
        beams=beams, paths=paths,
        domain={"X": axis(0.0, 5e-3, nx), "Y": axis(0.0, 5e-3, ny), "Z": axis(-5e-4, 0.0, nz)},
        output=out_cols("tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt"))


def cfg3_small() -> Deck:
    """Two beams with short rasters in a 1x1x0.2 mm box at 20 um spacing."""
    beams = [dict(BEAM), dict(BEAM)]
    beams[1]["Power"] = 250.0
    beams[1]["Width_X"] = 40e-6
    paths = [serpentine(0.2, 0.6, 0.3, 0.1, 2, 1.0),
             serpentine(0.5, 0.8, 0.7, 0.1, 2, 0.8)]
    d = Deck(name="cfg3s", mode_block="Solidification", mode={"Tracking": "Volume", "Timestep": 1e-5},
             beams=beams, paths=paths,
             domain={"X": axis(0.0, 1e-3, 51), "Y": axis(0.0, 1e-3, 51), "Z": axis(-2e-4, 0.0, 11)},
             output=out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "eqFrac", "numMelt"))
    d.cet = {"N0": 3e13, "n": 3, "a": 1.25e6}
    return d


def none_small() -> Deck:
    """Solidification with tracking None (every point every step) on a tiny box."""
    return Deck(
        name="nones", mode_block="Solidification", mode={"Tracking": "None", "Timestep": 2e-5},
        beams=[dict(BEAM)], paths=[[(0, 0.3, 0.0, 0.0, 1.0, 1.0)]],
        domain={"X": axis(0.0, 3e-4, 16), "Y": axis(-8e-5, 8e-5, 9), "Z": axis(-6e-5, 0.0, 4)},
        output=out_cols("T", "tSol", "G", "V", "Gx", "Gy", "Gz", "dTdt", "numMelt"))


def snapvol_small(tracking="Volume") -> Deck:
    """Snapshots with melt-pool tracking at three scan fractions."""
    return Deck(
        name="snapv", mode_block="Snapshots", mode={"ScanFracs": "40,70,100", "Tracking": tracking},
        beams=[dict(BEAM)], paths=[serpentine(0.2, 0.8, 0.3, 0.1, 2, 1.0)],
        domain={"X": axis(0.0, 1e-3, 51), "Y": axis(0.0, 0.6e-3, 31), "Z": axis(-1e-4, 0.0, 6)},
        output=out_cols("T"))


def cfg4_sets(n_sets=256) -> List[Dict[str, object]]:
    """The 256 seeded snapshot parameter sets (default_rng(4000))."""
    rng = np.random.default_rng(4000)
    sets = []
    for _ in range(n_sets):
        P = float(rng.uniform(150.0, 400.0))
        v = float(rng.uniform(0.5, 2.0))
        wx = float(rng.uniform(40e-6, 100e-6))
        wy = float(rng.uniform(40e-6, 100e-6))
        dz = float(rng.uniform(1e-6, 20e-6))
        nh = int(rng.integers(1, 21))
        beam = dict(Width_X=wx, Width_Y=wy, Depth_Z=dz, Power=P, Efficiency=0.35)
        path = serpentine(0.0, 1.0, 0.025, 0.05, nh, v)
        sets.append(dict(beam=beam, path=path))
    return sets


def cfg4_deck(k: int, nx=128, ny=128, nz=32) -> Deck:
    s = cfg4_sets()[k]
    return Deck(
        name=f"cfg4_{k:03d}", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[s["beam"]], paths=[s["path"]],
        domain={"X": axis(0.0, 1e-3, nx), "Y": axis(0.0, 1e-3, ny), "Z": axis(-2.5e-4, 0.0, nz)},
        output=out_cols("T"))


def cfg5_path(rng: np.random.Generator, nh=10, pieces=19, length=1.6, dy=0.1,
              x0=0.2, y0=0.55, sigma_mm=0.010) -> List[Row]:
    rows = serpentine(x0, x0 + length, y0, dy, nh, 1.0, pieces=pieces)
    out: List[Row] = []
    for r in rows:
        dx, dyy = rng.normal(0.0, sigma_mm, size=2)
        out.append((r[0], r[1] + float(dx), r[2] + float(dyy), r[3], r[4], r[5]))
    return out


def cfg5(nx=256, ny=256, nz=64) -> Deck:
    rng = np.random.default_rng(5000)
    return Deck(
        name="cfg5", mode_block="Snapshots", mode={"ScanFracs": "100", "Tracking": "None"},
        beams=[dict(BEAM)], paths=[cfg5_path(rng)],
        domain={"X": axis(0.0, 2e-3, nx), "Y": axis(0.0, 2e-3, ny), "Z": axis(-5e-4, 0.0, nz)},
        output=out_cols("T"))


def perim_small(tracking: str = "Surface") -> Deck:
    """Hatches that start and end outside the grid, two of them beside it (30 um
    past the y = 0 and y = 0.4 mm edges), so that the beam spends part of the
    scan off the grid and its melt pool reaches in only through the edge rows:
    the perimeter seeding of beam_trace_perimeter (Melt.cpp:20-96)."""
    rows: List[Row] = []
    for h, y in enumerate((-0.03, 0.1, 0.25, 0.43)):
        xs, xe = (-0.1, 0.7) if h % 2 == 0 else (0.7, -0.1)
        rows.append((1, xs, y, 0.0, 0.0, 1e-4))
        rows.append((0, xe, y, 0.0, 1.0, 1.0))
    return Deck(
        name="perim", mode_block="Solidification", mode={"Tracking": tracking, "Timestep": 1e-5},
        beams=[dict(BEAM)], paths=[rows],
        domain={"X": axis(0.0, 6e-4, 31), "Y": axis(0.0, 4e-4, 21), "Z": axis(-1.4e-4, 0.0, 8)},
        output=out_cols("T", "tSol", "G", "V", "dTdt", "numMelt"))


def thist_tracked_small(mode: str = "Solidification", tracking: str = "Surface") -> Deck:
    """T_hist with melt-pool tracking on a grid (Grid.cpp:138-174, :300-304):
    every evaluation with set records (t, T) for its point; one short hatch."""
    m = {"Tracking": tracking, "Timestep": 1e-5} if mode == "Solidification" else \
        {"ScanFracs": "50,100", "Tracking": tracking}
    d = Deck(
        name="thistt", mode_block=mode, mode=m,
        beams=[dict(BEAM)], paths=[[(1, 0.08, 0.15, 0.0, 0.0, 1e-4), (0, 0.38, 0.15, 0.0, 1.0, 1.0)]],
        domain={"X": axis(0.0, 5e-4, 26), "Y": axis(0.0, 3e-4, 16), "Z": axis(-1e-4, 0.0, 6)},
        output=out_cols("T", "tSol", "G", "V") if mode == "Solidification" else out_cols("T"))
    d.output["Temperature"]["T_hist"] = 1
    return d


WORKLOADS = {
    "cfg1": cfg1, "cfg1_small": lambda: cfg1(40, 11, 6), "cfg2": cfg2, "cfg2_small": cfg2_small,
    "cfg2_small_volume": lambda: cfg2_small("Volume"), "cfg3": cfg3, "cfg3_small": cfg3_small,
    "none_small": none_small, "snapvol_small": snapvol_small,
    "snapsurf_small": lambda: snapvol_small("Surface"), "cfg5": cfg5,
    "compress_small": compress_small,
    "compress_small_surface": lambda: compress_small("Solidification", "Surface"),
    "walls_small": walls_small,
    "walls_small_volume": lambda: walls_small("Solidification", "Volume"),
    "points_small": points_small, "points_small_sol": points_small_sol, "thist_none_small": thist_none_small,
    "rdf_small": rdf_small, "rdf_small_volume": lambda: rdf_small("Volume"), "stork_small": stork_small,
    "rdf_small_windowed": rdf_small_windowed,
    "perim_small": perim_small, "perim_small_volume": lambda: perim_small("Volume"),
    "thist_surface_small": thist_tracked_small,
    "thist_volume_small": lambda: thist_tracked_small("Solidification", "Volume"),
    "thist_snapvol_small": lambda: thist_tracked_small("Snapshots", "Volume"),
}
