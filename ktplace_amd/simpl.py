"""Host-side mirror of the reference's SimPL solver interface over the C ABI.

Names and argument meaning follow simpl_lss.h:

  SparseSystem                 simpl_lss.h:38-58
  LinearSystems.buildB2B       simpl_lss.h:75-76  (simpl_lss.cc:11-114)
  solveBothAxes                simpl_lss.h:86-88  (simpl_lss.cc:203-290)

Everything runs in libktplace_amd.so on the GPU; this module only marshals
numpy arrays through ctypes.  There is no CPU fallback: if the library or a
GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _lib
from ._lib import P, U32, U64, F64, I32, lib
from .design import Design, Flat

KT_PREC_FP64 = 0
KT_SPMV_ROW_GROUPS, KT_SPMV_ENTRY_BLOCKS = 0, 1
KT_PREC_MIXED = 1
KT_PREC_FP64_EXACT = 2
KT_STOP_RECURRENCE = 0
KT_STOP_TRUE = 1
KT_WARM_PREVIOUS = 0
KT_WARM_LEARNED = 1
KT_MLP_FEATURES = 17
KT_MLP_PARAMS = 20994

KERNEL_SPMV, KERNEL_UPDATE, KERNEL_PUPDATE, KERNEL_SPMV32, KERNEL_B2B_FILL = 0, 1, 2, 3, 4
KERNEL_WARM, KERNEL_MLP = 5, 6


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_uint64), ("precision", C.c_int32),
                ("stop_rule", C.c_int32), ("check_every", C.c_int32), ("profile", C.c_int32),
                ("warm_start", C.c_int32), ("reserved", C.c_int32)]


class SolveResult(C.Structure):
    _fields_ = [("resid", C.c_double * 2), ("true_resid", C.c_double * 2),
                ("iterations", C.c_uint64), ("steps", C.c_uint64 * 2),
                ("refinements", C.c_uint64), ("warm_gamma", C.c_double * 2)]

    def as_dict(self) -> dict:
        return dict(resid=(self.resid[0], self.resid[1]),
                    true_resid=(self.true_resid[0], self.true_resid[1]),
                    iterations=int(self.iterations), steps=(int(self.steps[0]), int(self.steps[1])),
                    refinements=int(self.refinements),
                    warm_gamma=(self.warm_gamma[0], self.warm_gamma[1]))


class LoopState(C.Structure):
    _fields_ = [("lower_hpwl", C.c_double), ("upper_hpwl", C.c_double), ("gap", C.c_double),
                ("overflow_upper", C.c_double), ("overflow_lower", C.c_double), ("gap_reference", C.c_double),
                ("best_upper", C.c_double), ("iteration", C.c_uint64), ("stalled_for", C.c_uint64),
                ("converged", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class SystemView(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p),
                ("diag", C.c_void_p), ("rhs", C.c_void_p)]


class BatchInstanceResult(C.Structure):
    _fields_ = [("resid", C.c_double * 2), ("true_resid", C.c_double * 2), ("steps", C.c_uint64 * 2),
                ("done_ms", C.c_double)]


OBSERVER = C.CFUNCTYPE(None, C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_int)

_SIGS = {
    "kt_ctx_create": (I32, [C.POINTER(C.c_void_p), I32]),
    "kt_ctx_destroy": (I32, [P]),
    "kt_sync": (I32, [P]),
    "kt_last_error": (C.c_char_p, []),
    "kt_ctx_device_bytes": (I32, [P, C.POINTER(C.c_uint64)]),
    "kt_upload_netlist": (I32, [P, P]),
    "kt_set_positions": (I32, [P, P, P, U64]),
    "kt_get_positions": (I32, [P, P, P, U64]),
    "kt_set_anchors": (I32, [P, P, P]),
    "kt_save_positions": (I32, [P]),
    "kt_restore_positions": (I32, [P]),
    "kt_build_b2b": (I32, [P, P, P, F64, P, P, F64, F64]),
    "kt_build_b2b_resident": (I32, [P, F64, F64, F64]),
    "kt_build_star": (I32, [P, C.POINTER(C.c_uint64)]),
    "kt_system_size": (I32, [P, I32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kt_get_system": (I32, [P, I32, P, P, P, P, P]),
    "kt_set_system": (I32, [P, I32, U64, P, P, P, P, P]),
    "kt_multiply": (I32, [P, I32, P, P]),
    "kt_solve_both": (I32, [P, P, P, U64, C.POINTER(SolveOpts), OBSERVER, P,
                            C.POINTER(SolveResult)]),
    "kt_solve_both_resident": (I32, [P, C.POINTER(SolveOpts), C.POINTER(SolveResult)]),
    "kt_set_tuning": (I32, [P, I32, U32, I32, F64]),
    "kt_set_tuning2": (I32, [P, I32, U32, I32, F64, I32, I32]),
    "kt_set_spmv_halo": (I32, [P, U32]),
    "kt_set_spmv_layout": (I32, [P, I32, U32, I32, I32, I32]),
    "kt_set_spmv_columns": (I32, [P, I32]),
    "kt_set_spmv_wide_loads": (I32, [P, I32]),
    "kt_set_mixed_scaling": (I32, [P, I32]),
    "kt_set_mixed_fused_update": (I32, [P, I32]),
    "kt_kernel_times": (I32, [P, I32, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "kt_reset_kernel_times": (I32, [P]),
    "kt_probe_copy": (I32, [P, U64, I32, C.POINTER(C.c_double)]),
    "kt_set_rows": (I32, [P, U32, P, P, P, P, P, P]),
    "kt_build_grid": (I32, [P, U64, F64, C.POINTER(C.c_uint64)]),
    "kt_grid_avail": (I32, [P, P, C.POINTER(C.c_double)]),
    "kt_occupancy": (I32, [P, P]),
    "kt_overflow": (I32, [P, C.POINTER(C.c_double)]),
    "kt_hpwl": (I32, [P, C.POINTER(C.c_double)]),
    "kt_hpwl_noalloc": (I32, [P, C.POINTER(C.c_double)]),
    "kt_set_lal_options": (I32, [P, U64, U64, F64]),
    "kt_lal": (I32, [P, P, P]),
    "kt_lal_resident": (I32, [P]),
    "kt_lal_batches": (I32, [P, C.POINTER(C.c_uint64)]),
    "kt_lal_reruns": (I32, [P, C.POINTER(C.c_uint64)]),
    "kt_set_lal_fused": (I32, [P, U32]),
    "kt_lal_fused_batches": (I32, [P, C.POINTER(C.c_uint64)]),
    "kt_floating_cells": (I32, [P, C.POINTER(C.c_uint32), P]),
    "kt_loop_reset": (I32, [P, C.c_uint64, C.c_double, C.c_double, C.c_uint64]),
    "kt_anchors_from_positions": (I32, [P]),
    "kt_get_anchors": (I32, [P, P, P]),
    "kt_loop_update": (I32, [P, C.c_uint64, P]),
    "kt_set_warm_start": (I32, [P, P, U64]),
    "kt_warm_features": (I32, [P, P, P]),
    "kt_warm_predict": (I32, [P, P, P]),
    "kt_mlp_create": (I32, [I32, C.POINTER(C.c_void_p)]),
    "kt_mlp_destroy": (I32, [P]),
    "kt_mlp_set_params": (I32, [P, P, U64]),
    "kt_mlp_get_params": (I32, [P, P, U64]),
    "kt_mlp_set_data": (I32, [P, P, P, P, U64]),
    "kt_mlp_forward": (I32, [P, P, U64, P]),
    "kt_mlp_gradient": (I32, [P, P, U64, P, C.POINTER(C.c_double)]),
    "kt_mlp_train": (I32, [P, U64, U32, F64, U64, P]),
    "kt_batch_create": (I32, [I32, C.POINTER(C.c_void_p)]),
    "kt_batch_destroy": (I32, [P]),
    "kt_batch_clear": (I32, [P]),
    "kt_batch_add": (I32, [P, P, P, P, P, C.POINTER(C.c_uint32)]),
    "kt_batch_add_ctx": (I32, [P, P, C.POINTER(C.c_uint32)]),
    "kt_batch_size": (I32, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kt_batch_set_tuning": (I32, [P, I32, U32, I32, I32]),
    "kt_batch_set_layout": (I32, [P, I32, I32]),
    "kt_batch_set_columns": (I32, [P, I32]),
    "kt_batch_set_wide_loads": (I32, [P, I32]),
    "kt_batch_reset": (I32, [P]),
    "kt_batch_solve": (I32, [P, C.POINTER(SolveOpts), P, C.POINTER(SolveResult)]),
    "kt_batch_get_positions": (I32, [P, U32, P, P, U64]),
}

EXPORTS = tuple(_SIGS)

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class KtError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        raise KtError(f"status {rc}: {lib.kt_last_error().decode()}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def netlist_struct(flat: Flat) -> _lib.KtgFlat:
    """kt_netlist view of a Flat (arrays are kept alive by the returned struct)."""
    s = _lib.KtgFlat()
    s.num_vertices = flat.num_vertices
    s.num_movable = flat.num_movable
    s.num_nets = flat.num_nets
    s.num_pins = flat.num_pins
    s.num_cell_pins = len(flat.cell_pin_net)
    s.num_obstacles = len(flat.obstacles)
    keep = {}
    for name in ("net_start", "pin_cell", "pin_off_x", "pin_off_y", "net_weight", "var_of_vertex",
                 "mov_vertex", "fixed_x", "fixed_y", "pin_start", "cell_pin_net", "cell_pin_idx",
                 "width", "height", "area"):
        arr = np.ascontiguousarray(getattr(flat, name))
        keep[name] = arr
        setattr(s, name, _lib.ptr(arr))
    obs = np.ascontiguousarray(flat.obstacles.reshape(-1), np.float64)
    keep["obstacles"] = obs
    s.obstacles = _lib.ptr(obs) if len(obs) else None
    s._keep = keep
    return s


@dataclass
class SparseSystem:
    """One axis' system (simpl_lss.h:38-58): CSR off-diagonals + diag + rhs."""
    row_ptr: np.ndarray  # u64 [n+1]
    col: np.ndarray      # u32 [nnz]
    val: np.ndarray      # f64 [nnz]
    diag: np.ndarray     # f64 [n]
    rhs: np.ndarray      # f64 [n]

    def size(self) -> int:
        return len(self.diag)


class Context:
    """One design on one GPU (a kt_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib.kt_ctx_create(C.byref(h), device))
        self.h = h
        self.n = 0
        self.floors = (1.0, 1.0)

    def close(self):
        if getattr(self, "h", None):
            lib.kt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # netlist -----------------------------------------------------------------
    def upload(self, design_or_flat) -> "Context":
        flat = design_or_flat.flat if isinstance(design_or_flat, Design) else design_or_flat
        s = netlist_struct(flat)
        _check(lib.kt_upload_netlist(self.h, C.byref(s)))
        self.n = flat.num_movable
        self.floors = flat.floors
        if isinstance(design_or_flat, Design):
            self.set_rows(design_or_flat.raw)
        return self

    # rows, grid, bookkeeping, LAL -------------------------------------------
    def set_rows(self, raw):
        arrs = [np.ascontiguousarray(getattr(raw, k), np.float64)
                for k in ("row_y", "row_height", "row_x0", "row_x1", "row_pitch")]
        die = np.ascontiguousarray(raw.die, np.float64)
        _check(lib.kt_set_rows(self.h, len(arrs[0]), *[_lib.ptr(a) for a in arrs], _lib.ptr(die)))
        self._rows_keep = (arrs, die)

    def build_grid(self, bins_per_axis: int = 0, density: float = 1.0) -> int:
        b = C.c_uint64()
        _check(lib.kt_build_grid(self.h, bins_per_axis, density, C.byref(b)))
        self.bins = b.value
        return b.value

    def grid_avail(self):
        a = np.zeros(self.bins * self.bins)
        t = C.c_double()
        _check(lib.kt_grid_avail(self.h, _lib.ptr(a), C.byref(t)))
        return a, t.value

    def occupancy(self):
        occ = np.zeros(self.bins * self.bins)
        _check(lib.kt_occupancy(self.h, _lib.ptr(occ)))
        return occ

    def overflow(self) -> float:
        v = C.c_double()
        _check(lib.kt_overflow(self.h, C.byref(v)))
        return v.value

    def hpwl(self) -> float:
        v = C.c_double()
        _check(lib.kt_hpwl(self.h, C.byref(v)))
        return v.value

    def hpwl_noalloc(self) -> float:
        """kt_hpwl_noalloc: the same HPWL with no device allocation."""
        v = C.c_double()
        _check(lib.kt_hpwl_noalloc(self.h, C.byref(v)))
        return v.value

    def set_lal_options(self, max_level: int = 0, min_cells_to_split: int = 0,
                        stripe_area_fraction: float = 0.0):
        _check(lib.kt_set_lal_options(self.h, max_level, min_cells_to_split, stripe_area_fraction))

    def lal(self, x, y):
        """LookAheadLegalizer::legalize on copies of x, y (returned)."""
        x = np.array(x, np.float64)
        y = np.array(y, np.float64)
        _check(lib.kt_lal(self.h, _lib.ptr(x), _lib.ptr(y)))
        return x, y

    def lal_resident(self):
        _check(lib.kt_lal_resident(self.h))

    def lal_batches(self) -> int:
        b = C.c_uint64()
        _check(lib.kt_lal_batches(self.h, C.byref(b)))
        return b.value

    # the outer loop's bookkeeping on the device ---------------------------------
    def loop_reset(self, reference_iter: int = 10, tight_frac: float = 0.05, relaxed_frac: float = 0.25,
                   stall_iters: int = 6):
        _check(lib.kt_loop_reset(self.h, reference_iter, tight_frac, relaxed_frac, stall_iters))

    def anchors_from_positions(self):
        _check(lib.kt_anchors_from_positions(self.h))

    def get_anchors(self):
        x = np.zeros(self.n)
        y = np.zeros(self.n)
        _check(lib.kt_get_anchors(self.h, _lib.ptr(x), _lib.ptr(y)))
        return x, y

    def loop_update(self, iteration: int) -> dict:
        st = LoopState()
        _check(lib.kt_loop_update(self.h, iteration, C.byref(st)))
        return st.as_dict()

    def floating_cells(self):
        """(count, flags[n]) of the movable cells with no net path to a fixed vertex."""
        c = C.c_uint32()
        f = np.zeros(self.n, np.uint8)
        _check(lib.kt_floating_cells(self.h, C.byref(c), _lib.ptr(f)))
        return c.value, f

    def lal_reruns(self) -> int:
        b = C.c_uint64()
        _check(lib.kt_lal_reruns(self.h, C.byref(b)))
        return b.value

    def set_lal_fused(self, max_cells: int):
        """Batches of at most max_cells cells (<= 1024) run in one launch each; 0: all level by level."""
        _check(lib.kt_set_lal_fused(self.h, int(max_cells)))

    def lal_fused_batches(self) -> int:
        b = C.c_uint64()
        _check(lib.kt_lal_fused_batches(self.h, C.byref(b)))
        return b.value

    # positions ---------------------------------------------------------------
    def set_positions(self, x, y):
        x, y = _f64(x), _f64(y)
        _check(lib.kt_set_positions(self.h, _lib.ptr(x), _lib.ptr(y), len(x)))

    def get_positions(self, count: Optional[int] = None):
        count = self.n if count is None else count
        x = np.zeros(count)
        y = np.zeros(count)
        _check(lib.kt_get_positions(self.h, _lib.ptr(x), _lib.ptr(y), count))
        return x, y

    def set_anchors(self, ax, ay):
        ax, ay = _f64(ax), _f64(ay)
        _check(lib.kt_set_anchors(self.h, _lib.ptr(ax), _lib.ptr(ay)))

    def save_positions(self):
        _check(lib.kt_save_positions(self.h))

    def restore_positions(self):
        _check(lib.kt_restore_positions(self.h))

    def sync(self):
        _check(lib.kt_sync(self.h))

    def device_bytes(self) -> int:
        b = C.c_uint64()
        _check(lib.kt_ctx_device_bytes(self.h, C.byref(b)))
        return b.value

    # LinearSystems -----------------------------------------------------------
    def build_b2b(self, x, y, alpha: float = 0.0, anchor_x=None, anchor_y=None, floors=None):
        """LinearSystems::buildB2B (simpl_lss.cc:11-114) into the device systems."""
        fx, fy = floors if floors is not None else self.floors
        x, y = _f64(x), _f64(y)
        ax = None if anchor_x is None else _f64(anchor_x)
        ay = None if anchor_y is None else _f64(anchor_y)
        _check(lib.kt_build_b2b(self.h, _lib.ptr(x), _lib.ptr(y), float(alpha), _lib.ptr(ax),
                                _lib.ptr(ay), float(fx), float(fy)))

    def build_star(self) -> int:
        """LinearSystems::buildStar (simpl_lss.cc:116-201); returns n + stars."""
        size = C.c_uint64()
        _check(lib.kt_build_star(self.h, C.byref(size)))
        return size.value

    def build_b2b_resident(self, alpha: float = 0.0, floors=None):
        fx, fy = floors if floors is not None else self.floors
        _check(lib.kt_build_b2b_resident(self.h, float(alpha), float(fx), float(fy)))

    def system_size(self, axis: int):
        rows = C.c_uint64()
        nnz = C.c_uint64()
        _check(lib.kt_system_size(self.h, axis, C.byref(rows), C.byref(nnz)))
        return rows.value, nnz.value

    def get_system(self, axis: int) -> SparseSystem:
        rows, nnz = self.system_size(axis)
        s = SparseSystem(np.zeros(rows + 1, np.uint64), np.zeros(nnz, np.uint32), np.zeros(nnz),
                         np.zeros(rows), np.zeros(rows))
        _check(lib.kt_get_system(self.h, axis, _lib.ptr(s.row_ptr), _lib.ptr(s.col),
                                 _lib.ptr(s.val), _lib.ptr(s.diag), _lib.ptr(s.rhs)))
        return s

    def set_system(self, axis: int, s):
        rp = np.ascontiguousarray(s.row_ptr, np.uint64)
        col = np.ascontiguousarray(s.col, np.uint32)
        val = _f64(s.val)
        diag = _f64(s.diag)
        rhs = _f64(s.rhs)
        _check(lib.kt_set_system(self.h, axis, len(diag), _lib.ptr(rp), _lib.ptr(col),
                                 _lib.ptr(val), _lib.ptr(diag), _lib.ptr(rhs)))

    def multiply(self, axis: int, v) -> np.ndarray:
        v = _f64(v)
        out = np.zeros_like(v)
        _check(lib.kt_multiply(self.h, axis, _lib.ptr(v), _lib.ptr(out)))
        return out

    # solveBothAxes -----------------------------------------------------------
    def solve_both(self, x, y, tol: float, max_iter: int, precision: int = KT_PREC_FP64,
                   stop_rule: int = KT_STOP_RECURRENCE, observe: Optional[Callable] = None,
                   check_every: int = 0, profile: bool = False, warm_start: int = KT_WARM_PREVIOUS):
        """solveBothAxes: x, y are updated in place (numpy float64 arrays)."""
        if x.dtype != np.float64 or y.dtype != np.float64 or not x.flags["C_CONTIGUOUS"]:
            raise TypeError("x and y must be contiguous float64 arrays (updated in place)")
        opts = SolveOpts(tol=tol, max_iter=max_iter, precision=precision, stop_rule=stop_rule,
                         check_every=check_every, profile=int(profile), warm_start=warm_start)
        res = SolveResult()
        cb = OBSERVER(lambda _u, it, rx, ry, last: observe(int(it), (rx, ry), bool(last))) \
            if observe else OBSERVER()
        _check(lib.kt_solve_both(self.h, _lib.ptr(x), _lib.ptr(y), len(x), C.byref(opts), cb,
                                 None, C.byref(res)))
        return res.as_dict()

    def solve_resident(self, tol: float, max_iter: int, precision: int = KT_PREC_FP64,
                       stop_rule: int = KT_STOP_RECURRENCE, check_every: int = 0,
                       profile: bool = False, warm_start: int = KT_WARM_PREVIOUS) -> dict:
        opts = SolveOpts(tol=tol, max_iter=max_iter, precision=precision, stop_rule=stop_rule,
                         check_every=check_every, profile=int(profile), warm_start=warm_start)
        res = SolveResult()
        _check(lib.kt_solve_both_resident(self.h, C.byref(opts), C.byref(res)))
        return res.as_dict()

    def set_tuning(self, lanes: int = 0, tile: int = 0, use_graphs: bool = True,
                   refresh_factor: float = 0.0, xcd_split: int = -1, rows: int = 0, halo: int = 0,
                   layout: int = 0, block_entries: int = 0, unroll: int = 0, row_lanes: int = 0,
                   block_threads: int = 0, col_bytes: int = 0, wide: int = 0):
        _check(lib.kt_set_tuning2(self.h, lanes, tile, int(use_graphs), refresh_factor, xcd_split,
                                  rows))
        _check(lib.kt_set_spmv_halo(self.h, halo))
        _check(lib.kt_set_spmv_layout(self.h, layout, block_entries, unroll, row_lanes,
                                      block_threads))
        _check(lib.kt_set_spmv_columns(self.h, col_bytes))
        _check(lib.kt_set_spmv_wide_loads(self.h, int(wide)))

    def set_mixed_scaling(self, on: bool):
        """Mixed precision: FP32 iterations on D^-1/2 A D^-1/2 (default) or on A."""
        _check(lib.kt_set_mixed_scaling(self.h, int(on)))

    def set_mixed_fused_update(self, on: bool):
        """Mixed precision, scaled: two kernels per FP32 iteration (default, where the shape allows) or three."""
        _check(lib.kt_set_mixed_fused_update(self.h, int(on)))

    def kernel_time(self, kernel: int):
        n = C.c_uint64()
        ms = C.c_double()
        _check(lib.kt_kernel_times(self.h, kernel, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def probe_copy(self, nbytes: int, reps: int = 3) -> float:
        """Milliseconds per plain device copy of nbytes (calibration helper)."""
        ms = C.c_double()
        _check(lib.kt_probe_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def reset_kernel_times(self):
        _check(lib.kt_reset_kernel_times(self.h))

    # learned warm start ------------------------------------------------------
    def set_warm_start(self, params: Optional[np.ndarray]):
        """Network for KT_WARM_LEARNED solves (None clears it)."""
        if params is None:
            _check(lib.kt_set_warm_start(self.h, None, 0))
            return
        p = np.ascontiguousarray(params, np.float32)
        _check(lib.kt_set_warm_start(self.h, _lib.ptr(p), len(p)))

    def warm_features(self, with_features: bool = True):
        """(features [n][17] float32 or None, (s_x, s_y)) at the current
        positions and systems."""
        f = np.zeros((self.n, KT_MLP_FEATURES), np.float32) if with_features else None
        sc = np.zeros(2)
        _check(lib.kt_warm_features(self.h, _lib.ptr(f) if f is not None else None, _lib.ptr(sc)))
        return f, (float(sc[0]), float(sc[1]))

    def warm_predict(self):
        dx = np.zeros(self.n)
        dy = np.zeros(self.n)
        _check(lib.kt_warm_predict(self.h, _lib.ptr(dx), _lib.ptr(dy)))
        return dx, dy


class Batch:
    """Many designs' x/y systems solved together on one GPU (a kt_batch)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib.kt_batch_create(device, C.byref(h)))
        self.h = h
        self.rows = []

    def close(self):
        if getattr(self, "h", None):
            lib.kt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def clear(self):
        _check(lib.kt_batch_clear(self.h))
        self.rows = []

    def add(self, sx: SparseSystem, sy: SparseSystem, x0, y0) -> int:
        keep = []

        def view(s):
            arrs = [np.ascontiguousarray(s.row_ptr, np.uint64), np.ascontiguousarray(s.col, np.uint32),
                    _f64(s.val), _f64(s.diag), _f64(s.rhs)]
            keep.extend(arrs)
            return SystemView(len(arrs[3]), *[a.ctypes.data for a in arrs])

        vx, vy = view(sx), view(sy)
        x0, y0 = _f64(x0), _f64(y0)
        idx = C.c_uint32()
        _check(lib.kt_batch_add(self.h, C.byref(vx), C.byref(vy), _lib.ptr(x0), _lib.ptr(y0), C.byref(idx)))
        self.rows.append(len(x0))
        return idx.value

    def add_ctx(self, ctx: Context) -> int:
        idx = C.c_uint32()
        _check(lib.kt_batch_add_ctx(self.h, ctx.h, C.byref(idx)))
        self.rows.append(ctx.system_size(0)[0])
        return idx.value

    def size(self):
        n = C.c_uint32()
        r = C.c_uint64()
        z = C.c_uint64()
        _check(lib.kt_batch_size(self.h, C.byref(n), C.byref(r), C.byref(z)))
        return n.value, r.value, z.value

    def set_tuning(self, lanes: int = 0, tile: int = 0, rows: int = 0, use_graphs: bool = True,
                   layout: int = 0, row_lanes: int = 0, col_bytes: int = 0, wide: int = 0):
        _check(lib.kt_batch_set_tuning(self.h, lanes, tile, rows, int(use_graphs)))
        _check(lib.kt_batch_set_layout(self.h, layout, row_lanes))
        _check(lib.kt_batch_set_columns(self.h, col_bytes))
        _check(lib.kt_batch_set_wide_loads(self.h, int(wide)))

    def reset(self):
        _check(lib.kt_batch_reset(self.h))

    def solve(self, tol: float, max_iter: int, stop_rule: int = KT_STOP_RECURRENCE, check_every: int = 0):
        """(per-instance list of dicts, totals dict)."""
        opts = SolveOpts(tol=tol, max_iter=max_iter, precision=KT_PREC_FP64, stop_rule=stop_rule,
                         check_every=check_every)
        per = (BatchInstanceResult * max(len(self.rows), 1))()
        tot = SolveResult()
        _check(lib.kt_batch_solve(self.h, C.byref(opts), C.addressof(per), C.byref(tot)))
        out = [dict(resid=(r.resid[0], r.resid[1]), true_resid=(r.true_resid[0], r.true_resid[1]),
                    steps=(int(r.steps[0]), int(r.steps[1])), done_ms=r.done_ms)
               for r in per[:len(self.rows)]]
        return out, tot.as_dict()

    def get_positions(self, index: int):
        n = self.rows[index]
        x = np.zeros(n)
        y = np.zeros(n)
        _check(lib.kt_batch_get_positions(self.h, index, _lib.ptr(x), _lib.ptr(y), n))
        return x, y


class LinearSystems:
    """The reference's LinearSystems (simpl_lss.h:68-84) bound to a GPU context."""

    def __init__(self, ctx: Context, floor_x: float, floor_y: float):
        self.ctx = ctx
        self.floorX = floor_x
        self.floorY = floor_y

    def buildB2B(self, x, y, alpha, anchorX, anchorY):
        self.ctx.build_b2b(x, y, alpha, anchorX if len(anchorX) else None,
                           anchorY if len(anchorY) else None, floors=(self.floorX, self.floorY))

    @property
    def sysX(self) -> SparseSystem:
        return self.ctx.get_system(0)

    @property
    def sysY(self) -> SparseSystem:
        return self.ctx.get_system(1)


def solveBothAxes(ctx: Context, x: np.ndarray, y: np.ndarray, tolerance: float,
                  maxIterations: int, observe: Optional[Callable] = None):
    """solveBothAxes on the context's current systems; returns (resid_x, resid_y)."""
    r = ctx.solve_both(x, y, tolerance, maxIterations, observe=observe)
    return r["resid"]
