"""Python host side of the MI355X 3DThesis path, over the C-ABI of lib3dthesis_amd.so.

The package mirrors the reference's operator surface for this path:

  Thesis(paraminput)            Init::GetFileNames + Init::ReadSimParams (Init.cpp:233-303)
  .build_nodes(t, is_sol)       Calc::Integrate_Serial (Calc.cpp:48-65)
  .eval_T(nodes, ...)           Grid::Calc_T over grid points (Grid.cpp:275-314)
  .eval_sol(t, points, ...)     Grid::Solidify (Grid.cpp:316-338)
  .run()                        Run::Simulate + final output (Run.cpp:27-57, Main.cpp:104-112)

The package name starts with a digit, so import it with
importlib.import_module("3dthesis_amd"). Every GPU entry point goes through the
HIP library; there is no CPU fallback, and a missing or unloadable library is an
ImportError at first use.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib3dthesis_amd.so")
CLI_PATH = os.path.join(_HERE, "thesis")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "thesis_amd.h")

FP64, FP32C = 0, 1
NODE_FIELDS = ("xb", "yb", "zb", "ix", "iy", "iz", "dtau", "expmod")


class ThesisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Info(C.Structure):
    _fields_ = [("mode", C.c_int), ("tracking", C.c_int), ("nbeams", C.c_int), ("xnum", C.c_int),
                ("ynum", C.c_int), ("znum", C.c_int), ("pnum", C.c_longlong), ("nsnap", C.c_int),
                ("scan_end", C.c_double), ("dt", C.c_double), ("T_init", C.c_double), ("T_liq", C.c_double),
                ("alpha", C.c_double), ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("evals", C.c_double), ("wall_s", C.c_double), ("steps", C.c_double), ("nodes_built", C.c_double),
                ("cat_evals", C.c_double * 3), ("cat_ms", C.c_double * 3), ("cat_launches", C.c_double * 3),
                ("evals_window", C.c_double)]


_lib = None

_SIGS = {
    "thesis_ctx_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "thesis_ctx_destroy": (C.c_int, [C.c_void_p]),
    "thesis_last_error": (C.c_char_p, []),
    "thesis_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "thesis_set_full_dump": (C.c_int, [C.c_void_p, C.c_char_p]),
    "thesis_get_window_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int]),
    "thesis_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "thesis_get_snapshot_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int]),
    "thesis_build_nodes": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double), C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "thesis_reset_hints": (C.c_int, [C.c_void_p]),
    "thesis_eval_T": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_size_t, C.c_int, C.c_longlong,
                                C.c_longlong, C.POINTER(C.c_double)]),
    "thesis_eval_T_points": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_size_t, C.c_int,
                                       C.POINTER(C.c_int), C.c_size_t, C.POINTER(C.c_double)]),
    "thesis_upload_nodes": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_size_t]),
    "thesis_eval_T_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_void_p, C.c_void_p]),
    "thesis_eval_T_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_longlong, C.c_longlong,
                                       C.c_void_p, C.c_void_p]),
    "thesis_eval_sol": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int), C.c_size_t, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "thesis_run": (C.c_int, [C.c_void_p]),
    "thesis_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "thesis_batch_set": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "thesis_batch_eval_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "thesis_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "thesis_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]),
    "thesis_batch_nodes": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "thesis_design_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                    C.POINTER(C.c_int)]),
    "thesis_design_set": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    "thesis_forward_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    "thesis_adjoint_dev": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_double)]),
    "thesis_design_chain": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_size_t,
                                      C.POINTER(C.c_double)]),
    "thesis_design_nodes": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_size_t,
                                      C.POINTER(C.c_size_t)]),
}


def lib() -> C.CDLL:
    """Load lib3dthesis_amd.so (raises ImportError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        # PyTorch-ROCm ships its own libamdhip64 (same soname). Loading torch
        # first makes this library bind to that already-loaded runtime, so the
        # process holds one HIP runtime and torch tensors can be passed in.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> Sequence[str]:
    return tuple(_SIGS)


def _check(rc: int) -> None:
    if rc != 0:
        raise ThesisError(rc, lib().thesis_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Thesis:
    """One parsed 3DThesis problem (a ParamInput deck) and its GPU engine."""

    def __init__(self, paraminput: str, precision: int = FP64, abs_mode: int = 0):
        h = C.c_void_p()
        _check(lib().thesis_ctx_create(paraminput.encode(), C.byref(h)))
        self._h = h
        if abs_mode:
            self.set_option("abs_mode", abs_mode)
        self.set_option("precision", precision)
        info = Info()
        _check(lib().thesis_get_info(self._h, C.byref(info)))
        self.info = info

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().thesis_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- options and metadata
    def set_option(self, key: str, value: float) -> None:
        _check(lib().thesis_set_option(self._h, key.encode(), float(value)))

    def set_full_dump(self, d: Optional[str]) -> None:
        _check(lib().thesis_set_full_dump(self._h, d.encode() if d else None))

    @property
    def shape(self) -> Tuple[int, int, int]:
        return self.info.xnum, self.info.ynum, self.info.znum

    def snapshot_times(self) -> np.ndarray:
        out = np.zeros(max(1, self.info.nsnap))
        _check(lib().thesis_get_snapshot_times(self._h, _dptr(out), self.info.nsnap))
        return out[: self.info.nsnap]

    # --- node build (host)
    def build_nodes(self, t: float, is_sol: bool = False) -> np.ndarray:
        """(N, 8) float64 nodes: xb, yb, zb, 1/phix, 1/phiy, 1/phiz, dtau, expmod."""
        n = C.c_size_t()
        cap = 4096
        while True:
            out = np.empty((cap, 8))
            rc = lib().thesis_build_nodes(self._h, float(t), int(is_sol), _dptr(out), cap, C.byref(n))
            if rc == 4 and n.value > cap:  # too small: retry; the hints already sit at seg(t)
                cap = n.value
                continue
            _check(rc)
            return out[: n.value].copy()

    def reset_hints(self) -> None:
        _check(lib().thesis_reset_hints(self._h))

    # --- per-point sums (GPU)
    def eval_T(self, nodes: np.ndarray, precision: int = FP64, p_begin: int = 0,
               count: Optional[int] = None) -> np.ndarray:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        if count is None:
            count = self.info.pnum - p_begin
        out = np.empty(count)
        _check(lib().thesis_eval_T(self._h, _dptr(nodes), nodes.shape[0], precision, p_begin, count, _dptr(out)))
        return out

    def eval_T_points(self, nodes: np.ndarray, points: np.ndarray, precision: int = FP64) -> np.ndarray:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        pts = np.ascontiguousarray(points, dtype=np.int32)
        out = np.empty(pts.shape[0])
        _check(lib().thesis_eval_T_points(self._h, _dptr(nodes), nodes.shape[0], precision,
                                          pts.ctypes.data_as(C.POINTER(C.c_int)), pts.shape[0], _dptr(out)))
        return out

    def upload_nodes(self, nodes: np.ndarray) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        _check(lib().thesis_upload_nodes(self._h, _dptr(nodes), nodes.shape[0]))

    def eval_T_dev(self, T_dev_ptr: int, precision: int, p_begin: int, count: int, stream: int = 0) -> None:
        """Evaluate the uploaded nodes into a device array (e.g. a torch tensor's data_ptr()).
        stream: a hipStream_t handle; 0 means the context's own stream (not the legacy
        null stream), so callers timing on torch's stream should use a non-default stream."""
        _check(lib().thesis_eval_T_dev(self._h, precision, p_begin, count, C.c_void_p(T_dev_ptr),
                                       C.c_void_p(stream) if stream else None))

    def eval_T_static(self, nodes_dev_ptr: int, n_nodes: int, T_dev_ptr: int, precision: int, p_begin: int,
                      count: int, stream: int = 0) -> None:
        """Allocation-free evaluation: nodes already in device memory (n_nodes x 8 doubles,
        build_nodes layout), T into a device array. stream 0 is the null stream here."""
        _check(lib().thesis_eval_T_static(self._h, C.c_void_p(nodes_dev_ptr), n_nodes, precision, p_begin, count,
                                          C.c_void_p(T_dev_ptr), C.c_void_p(stream) if stream else None))

    def eval_sol(self, t: float, points: np.ndarray, T: np.ndarray, T_last: np.ndarray) -> np.ndarray:
        """(n, 7): tSol, G, Gx/G, Gy/G, Gz/G, dTdt, V."""
        pts = np.ascontiguousarray(points, dtype=np.int32)
        T = np.ascontiguousarray(T, dtype=np.float64)
        Tl = np.ascontiguousarray(T_last, dtype=np.float64)
        out = np.empty((pts.shape[0], 7))
        _check(lib().thesis_eval_sol(self._h, float(t), pts.ctypes.data_as(C.POINTER(C.c_int)), pts.shape[0],
                                     _dptr(T), _dptr(Tl), _dptr(out)))
        return out

    # --- batched snapshots (GPU)
    def batch_set(self, beams: np.ndarray, paths: Sequence[np.ndarray], times: Optional[np.ndarray] = None) -> None:
        """beams: (S, 5) Power, Efficiency, Width_X, Width_Y, Depth_Z; paths: S arrays of Path-file rows (n, 6)."""
        beams = np.ascontiguousarray(beams, dtype=np.float64)
        off = np.zeros(len(paths) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(p) for p in paths])
        rows = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64) for p in paths]))
        tp = None
        if times is not None:
            times = np.ascontiguousarray(times, dtype=np.float64)
            tp = _dptr(times)
        _check(lib().thesis_batch_set(self._h, len(paths), _dptr(beams), off.ctypes.data_as(C.POINTER(C.c_int)),
                                      _dptr(rows), tp))

    def batch_eval_dev(self, T_dev_ptr: int, precision: int = FP32C, stream: int = 0) -> None:
        _check(lib().thesis_batch_eval_dev(self._h, precision, C.c_void_p(T_dev_ptr),
                                           C.c_void_p(stream) if stream else None))

    def batch_autotune(self, T_dev_ptr: int, precision: int = FP32C) -> Tuple[int, int]:
        ppt, blocks = C.c_int(), C.c_int()
        _check(lib().thesis_batch_autotune(self._h, precision, C.c_void_p(T_dev_ptr), C.byref(ppt), C.byref(blocks)))
        return ppt.value, blocks.value

    def batch_info(self) -> Tuple[int, float]:
        n, ev = C.c_longlong(), C.c_double()
        _check(lib().thesis_batch_info(self._h, C.byref(n), C.byref(ev)))
        return n.value, ev.value

    def batch_nodes(self, s: int) -> np.ndarray:
        n = C.c_size_t()
        cap = 8192
        out = np.empty((cap, 8))
        rc = lib().thesis_batch_nodes(self._h, s, _dptr(out), cap, C.byref(n))
        if rc == 4:
            out = np.empty((n.value, 8))
            rc = lib().thesis_batch_nodes(self._h, s, _dptr(out), n.value, C.byref(n))
        _check(rc)
        return out[: n.value].copy()

    # --- design gradients
    def design(self) -> Tuple[np.ndarray, np.ndarray]:
        """(beam5, rows (n, 6)): Power, Efficiency, Width_X, Width_Y, Depth_Z and the Path-file rows."""
        n = C.c_int()
        beam = np.empty(5)
        rows = np.empty((4096, 6))
        rc = lib().thesis_design_get(self._h, _dptr(beam), _dptr(rows), 4096, C.byref(n))
        if rc == 4:
            rows = np.empty((n.value, 6))
            rc = lib().thesis_design_get(self._h, _dptr(beam), _dptr(rows), n.value, C.byref(n))
        _check(rc)
        return beam, rows[: n.value].copy()

    def set_design(self, beam5: np.ndarray, rows: np.ndarray) -> None:
        beam5 = np.ascontiguousarray(beam5, dtype=np.float64)
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        _check(lib().thesis_design_set(self._h, _dptr(beam5), _dptr(rows), rows.shape[0]))

    def forward_dev(self, T_dev_ptr: int, precision: int = FP64, t: float = -1.0, stream: int = 0) -> None:
        _check(lib().thesis_forward_dev(self._h, precision, float(t), C.c_void_p(T_dev_ptr),
                                        C.c_void_p(stream) if stream else None))

    def adjoint_dev(self, dLdT_dev_ptr: int, t: float = -1.0) -> np.ndarray:
        nrows = self.design()[1].shape[0]
        g = np.empty(5 + 2 * nrows)
        _check(lib().thesis_adjoint_dev(self._h, float(t), C.c_void_p(dLdT_dev_ptr), _dptr(g)))
        return g

    def design_nodes(self, t: float = -1.0) -> np.ndarray:
        n = C.c_size_t()
        out = np.empty((8192, 8))
        rc = lib().thesis_design_nodes(self._h, float(t), _dptr(out), 8192, C.byref(n))
        if rc == 4:
            out = np.empty((n.value, 8))
            rc = lib().thesis_design_nodes(self._h, float(t), _dptr(out), n.value, C.byref(n))
        _check(rc)
        return out[: n.value].copy()

    def design_chain(self, S: np.ndarray, t: float = -1.0) -> np.ndarray:
        S = np.ascontiguousarray(S, dtype=np.float64)
        nrows = self.design()[1].shape[0]
        g = np.empty(5 + 2 * nrows)
        _check(lib().thesis_design_chain(self._h, float(t), _dptr(S), S.shape[0], _dptr(g)))
        return g

    # --- whole run (GPU)
    def run(self) -> Stats:
        _check(lib().thesis_run(self._h))
        return self.stats()

    WINDOW_FIELDS = ("windows", "misses", "slots", "steps", "lds_windows", "waves", "t_setup_ms", "t_stage_ms",
                     "t_recheck_ms", "t_commit_work_ms", "t_waves_ms", "t_depth_ms", "t_commit_ms", "t_handoff_ms",
                     "scan_waves", "max_frontier", "max_liquid", "boxes", "box_columns", "box_layers",
                     "margin", "z_margin", "host_plan_ms", "host_stage_ms", "host_queue_ms", "host_build_ms",
                     "host_wait_ms", "host_process_ms", "per_step_switch",
                     "ms_box_sums", "ms_tracker", "ms_replay", "ms_solidify", "ms_columns", "ms_other",
                     "launches_box_sums", "launches_tracker", "launches_replay", "launches_solidify",
                     "launches_columns", "launches_other", "evals_windowed", "t_col_load_ms", "t_col_walk_ms",
                     "t_col_handoff_ms")

    def window_stats(self) -> dict:
        """Diagnostics of the windowed tracking of the last run."""
        out = np.zeros(len(self.WINDOW_FIELDS))
        _check(lib().thesis_get_window_stats(self._h, _dptr(out), out.size))
        return dict(zip(self.WINDOW_FIELDS, out.tolist()))

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().thesis_get_stats(self._h, C.byref(s)))
        return s
