"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Tuple

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "thesis_ref")


class OrNodes(C.Structure):
    _fields_ = [("size", C.c_size_t), ("cap", C.c_size_t)] + [
        (n, C.POINTER(C.c_double)) for n in ("xb", "yb", "zb", "phix", "phiy", "phiz", "dtau", "expmod")]


class OrSim(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256), ("data_dir", C.c_char * 1024),
        ("mode", C.c_int), ("tracking", C.c_int), ("nsnap", C.c_int),
        ("snap_times", C.POINTER(C.c_double)), ("dt", C.c_double), ("radius_check", C.c_double),
        ("out_freq", C.c_int), ("secondary", C.c_int),
        ("T_init", C.c_double), ("T_liq", C.c_double), ("kon", C.c_double), ("cps", C.c_double),
        ("rho", C.c_double), ("a", C.c_double), ("cet_N0", C.c_double), ("cet_n", C.c_double),
        ("cet_a", C.c_double),
        ("nbeams", C.c_int), ("beams", C.c_void_p), ("path_off", C.POINTER(C.c_int)),
        ("segs", C.c_void_p), ("all_end", C.c_double),
        ("xnum", C.c_int), ("ynum", C.c_int), ("znum", C.c_int), ("pnum", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double), ("xres", C.c_double), ("yres", C.c_double),
        ("zres", C.c_double), ("use_bcs", C.c_int), ("custom_points", C.c_int),
    ] + [(n, C.c_int) for n in ("o_x", "o_y", "o_z", "o_T", "o_Thist", "o_tSol", "o_G", "o_Gx", "o_Gy",
                                "o_Gz", "o_V", "o_dTdt", "o_eqFrac", "o_depth", "o_numMelt", "o_RDF",
                                "o_mpstats")] + [
        ("sol_tol", C.c_double), ("t_hist", C.c_double), ("p_hist", C.c_double), ("r_max", C.c_double),
        ("sol_iter", C.c_int), ("compress", C.c_int), ("thnum", C.c_int),
        ("abs_mode", C.c_int), ("parse_error", C.c_int), ("parse_msg", C.c_char * 512),
        ("bc_xmin", C.c_double), ("bc_xmax", C.c_double), ("bc_ymin", C.c_double), ("bc_ymax", C.c_double),
        ("bc_zmin", C.c_double), ("bc_refl", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.or_load.restype = C.POINTER(OrSim)
        L.or_load.argtypes = [C.c_char_p]
        L.or_free.argtypes = [C.POINTER(OrSim)]
        L.or_build_nodes.argtypes = [C.POINTER(OrSim), C.POINTER(C.c_int), C.c_double, C.c_int,
                                     C.POINTER(OrNodes)]
        L.or_nodes_free.argtypes = [C.POINTER(OrNodes)]
        L.or_point_xyz.argtypes = [C.POINTER(OrSim), C.c_int] + [C.POINTER(C.c_double)] * 3
        L.or_calc_T_xyz.restype = C.c_double
        L.or_calc_T_xyz.argtypes = [C.POINTER(OrSim), C.POINTER(OrNodes)] + [C.c_double] * 3
        L.or_sol_primary_xyz.argtypes = [C.POINTER(OrNodes)] + [C.c_double] * 3 + [C.POINTER(C.c_double)]
        L.or_run.restype = C.c_int
        L.or_run.argtypes = [C.POINTER(OrSim), C.c_char_p, C.POINTER(C.c_double)]
        L.or_snapshot_T.restype = C.c_long
        L.or_snapshot_T.argtypes = [C.POINTER(OrSim), C.c_double, C.POINTER(C.c_double), C.c_int]
        _lib = L
    return _lib


class Sim:
    """An oracle simulation loaded from a ParamInput file."""

    def __init__(self, paraminput: str, abs_mode: int = 0):
        self.ptr = lib().or_load(paraminput.encode())
        if not self.ptr:
            raise MemoryError("or_load failed")
        self.s = self.ptr.contents
        self.s.abs_mode = abs_mode
        self.hint = (C.c_int * max(1, self.s.nbeams))(*([1] * max(1, self.s.nbeams)))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_free(self.ptr)
            self.ptr = None

    @property
    def error(self) -> Optional[str]:
        return self.s.parse_msg.decode() if self.s.parse_error else None

    def nodes(self, t: float, is_sol: bool = False, hint: Optional[list] = None) -> np.ndarray:
        """(N, 8) array of xb,yb,zb,1/phix,1/phiy,1/phiz,dtau,expmod."""
        if hint is not None:
            for i, h in enumerate(hint):
                self.hint[i] = h
        n = OrNodes()
        lib().or_build_nodes(self.ptr, self.hint, t, int(is_sol), C.byref(n))
        out = np.empty((n.size, 8))
        for c, name in enumerate(("xb", "yb", "zb", "phix", "phiy", "phiz", "dtau", "expmod")):
            if n.size:
                out[:, c] = np.ctypeslib.as_array(getattr(n, name), shape=(n.size,))
        lib().or_nodes_free(C.byref(n))
        return out

    def snapshot_T(self, t: float, nthreads: int = 1) -> Tuple[np.ndarray, int]:
        T = np.empty(self.s.pnum)
        nn = lib().or_snapshot_T(self.ptr, t, T.ctypes.data_as(C.POINTER(C.c_double)), nthreads)
        return T, int(nn)

    def xyz(self, p: int) -> Tuple[float, float, float]:
        x, y, z = C.c_double(), C.c_double(), C.c_double()
        lib().or_point_xyz(self.ptr, p, C.byref(x), C.byref(y), C.byref(z))
        return x.value, y.value, z.value

    def run(self, full_dir: Optional[str] = None) -> float:
        ev = C.c_double()
        rc = lib().or_run(self.ptr, full_dir.encode() if full_dir else None, C.byref(ev))
        if rc:
            raise RuntimeError(f"oracle run failed: {self.error}")
        return ev.value


def nodes_from_arrays(arr: np.ndarray) -> Tuple[OrNodes, list]:
    """Wrap an (N, 8) node array as an OrNodes view (keeps buffers alive)."""
    cols = [np.ascontiguousarray(arr[:, c]) for c in range(8)]
    n = OrNodes()
    n.size = arr.shape[0]
    n.cap = arr.shape[0]
    for c, name in enumerate(("xb", "yb", "zb", "phix", "phiy", "phiz", "dtau", "expmod")):
        setattr(n, name, cols[c].ctypes.data_as(C.POINTER(C.c_double)))
    return n, cols


def sol_primary(arr: np.ndarray, x: float, y: float, z: float) -> np.ndarray:
    n, keep = nodes_from_arrays(arr)
    out = (C.c_double * 5)()
    lib().or_sol_primary_xyz(C.byref(n), x, y, z, out)
    return np.array(out[:])


def calc_T_xyz(sim: Sim, arr: np.ndarray, x: float, y: float, z: float) -> float:
    n, keep = nodes_from_arrays(arr)
    return lib().or_calc_T_xyz(sim.ptr, C.byref(n), x, y, z)


def read_full(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """Read a "<p> v1 v2 ..." full-precision dump -> (p int64 array, values 2-D)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                rows.append([float(v) for v in parts])
    if not rows:
        return np.zeros(0, dtype=np.int64), np.zeros((0, 0))
    a = np.array(rows)
    return a[:, 0].astype(np.int64), a[:, 1:]
