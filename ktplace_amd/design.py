"""Seeded synthetic designs and SimPL's flat netlist as numpy arrays.

`generate()` drives ktg_generate (the frozen generator in csrc/netgen.cc) and
`collect()` drives ktg_collect, the host restatement of simpl_dm.cc:16-85
(collectCells / collectNets).  A `Design` holds both the raw design (for
re-flattening and for the oracle) and the flat netlist that the C ABI uploads.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import KtgFlat, KtgRaw, KtgSpec, lib, view

# Workloads named by what they are (sizes, degree mix, pads, macros).
WORKLOADS = {
    # 12,500 cells / 14,000 nets / 250 pads, degree tail to 40, 70 % utilisation
    "cells12k": dict(num_cells=12_500, num_nets=14_000, num_pads=250),
    # 210,000 cells / 220,000 nets / 500 pads / 540 macros, 75 % utilisation
    "cells210k": dict(num_cells=210_000, num_nets=220_000, num_pads=500, num_macros=540,
                      utilization=0.75),
    # 1,000,000 cells / 1,100,000 nets / 2,000 pads, 0.1 % of nets with 41..200 pins
    "cells1m": dict(num_cells=1_000_000, num_nets=1_100_000, num_pads=2_000, max_degree=200,
                    high_fanout_frac=0.001),
}


def make_spec(num_cells: int, num_nets: int, num_pads: int, seed: int = 1, num_macros: int = 0,
              utilization: float = 0.70, max_degree: int = 40, high_fanout_frac: float = 0.0,
              local_frac: float = 0.9) -> KtgSpec:
    s = KtgSpec()
    _lib.check_ktg(lib.ktg_default_spec(C.byref(s), num_cells, num_nets, num_pads, seed))
    s.num_macros = num_macros
    s.utilization = utilization
    s.max_degree = max_degree
    s.high_fanout_frac = high_fanout_frac
    s.local_frac = local_frac
    return s


@dataclass
class Raw:
    """Raw design: vertices, nets as pin lists (cells may repeat), rows, die."""
    is_fixed: np.ndarray
    pos_x: np.ndarray
    pos_y: np.ndarray
    width: np.ndarray
    height: np.ndarray
    net_start: np.ndarray
    pin_vertex: np.ndarray
    pin_off_x: np.ndarray
    pin_off_y: np.ndarray
    net_weight: np.ndarray
    row_y: np.ndarray
    row_height: np.ndarray
    row_x0: np.ndarray
    row_x1: np.ndarray
    row_pitch: np.ndarray
    die: np.ndarray

    @property
    def num_vertices(self) -> int:
        return len(self.is_fixed)

    @property
    def num_nets(self) -> int:
        return len(self.net_weight)

    def to_c(self) -> KtgRaw:
        r = KtgRaw()
        r.num_vertices = self.num_vertices
        r.num_movable = int((self.is_fixed == 0).sum())
        r.num_nets = self.num_nets
        r.num_pins = len(self.pin_vertex)
        r.num_rows = len(self.row_y)
        for name in ("is_fixed", "pos_x", "pos_y", "width", "height", "net_start", "pin_vertex",
                     "pin_off_x", "pin_off_y", "net_weight", "row_y", "row_height", "row_x0",
                     "row_x1", "row_pitch"):
            setattr(r, name, _lib.ptr(getattr(self, name)))
        for i in range(4):
            r.die[i] = float(self.die[i])
        r._keep = self  # keep arrays alive with the struct
        return r


@dataclass
class Flat:
    """SimPL's flat netlist (simpl_lss.h:27-36) + movable sizes + obstacles."""
    net_start: np.ndarray      # u64 [N+1]
    pin_cell: np.ndarray       # u32 [P]
    pin_off_x: np.ndarray      # f64 [P]
    pin_off_y: np.ndarray
    net_weight: np.ndarray     # f64 [N]
    var_of_vertex: np.ndarray  # u32 [V]
    mov_vertex: np.ndarray     # u32 [n]
    fixed_x: np.ndarray        # f64 [V]
    fixed_y: np.ndarray
    pin_start: np.ndarray      # u64 [n+1]
    cell_pin_net: np.ndarray   # u32 [Pm]
    cell_pin_idx: np.ndarray   # u32 [Pm]
    width: np.ndarray          # f64 [n]
    height: np.ndarray
    area: np.ndarray
    obstacles: np.ndarray      # f64 [M, 4]

    @property
    def num_movable(self) -> int:
        return len(self.mov_vertex)

    @property
    def num_vertices(self) -> int:
        return len(self.var_of_vertex)

    @property
    def num_nets(self) -> int:
        return len(self.net_weight)

    @property
    def num_pins(self) -> int:
        return len(self.pin_cell)

    @property
    def floors(self) -> tuple[float, float]:
        """B2B length floors: mean movable width and height (simpl_placer.cc:169-179)."""
        n = self.num_movable
        if n == 0:
            return 1.0, 1.0
        # std::accumulate order: sequential left-to-right sum
        fx = 0.0
        fy = 0.0
        for w in self.width.tolist():
            fx += w
        for h in self.height.tolist():
            fy += h
        return fx / n, fy / n


@dataclass
class Design:
    raw: Raw
    flat: Flat
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.flat.num_movable

    @property
    def die(self) -> np.ndarray:
        return self.raw.die


def _raw_from_view(v: KtgRaw) -> Raw:
    V, N, Pn, R = v.num_vertices, v.num_nets, v.num_pins, v.num_rows
    return Raw(
        is_fixed=view(v.is_fixed, V, np.uint8), pos_x=view(v.pos_x, V, np.float64),
        pos_y=view(v.pos_y, V, np.float64), width=view(v.width, V, np.float64),
        height=view(v.height, V, np.float64), net_start=view(v.net_start, N + 1, np.uint64),
        pin_vertex=view(v.pin_vertex, Pn, np.uint32), pin_off_x=view(v.pin_off_x, Pn, np.float64),
        pin_off_y=view(v.pin_off_y, Pn, np.float64), net_weight=view(v.net_weight, N, np.float64),
        row_y=view(v.row_y, R, np.float64), row_height=view(v.row_height, R, np.float64),
        row_x0=view(v.row_x0, R, np.float64), row_x1=view(v.row_x1, R, np.float64),
        row_pitch=view(v.row_pitch, R, np.float64), die=np.array(list(v.die), np.float64))


def _flat_from_view(v: KtgFlat) -> Flat:
    V, n, N, Pn, Pm, M = (v.num_vertices, v.num_movable, v.num_nets, v.num_pins, v.num_cell_pins,
                          v.num_obstacles)
    return Flat(
        net_start=view(v.net_start, N + 1, np.uint64), pin_cell=view(v.pin_cell, Pn, np.uint32),
        pin_off_x=view(v.pin_off_x, Pn, np.float64), pin_off_y=view(v.pin_off_y, Pn, np.float64),
        net_weight=view(v.net_weight, N, np.float64),
        var_of_vertex=view(v.var_of_vertex, V, np.uint32),
        mov_vertex=view(v.mov_vertex, n, np.uint32), fixed_x=view(v.fixed_x, V, np.float64),
        fixed_y=view(v.fixed_y, V, np.float64), pin_start=view(v.pin_start, n + 1, np.uint64),
        cell_pin_net=view(v.cell_pin_net, Pm, np.uint32),
        cell_pin_idx=view(v.cell_pin_idx, Pm, np.uint32), width=view(v.width, n, np.float64),
        height=view(v.height, n, np.float64), area=view(v.area, n, np.float64),
        obstacles=view(v.obstacles, 4 * M, np.float64).reshape(M, 4))


def _take(handle: C.c_void_p, name: str) -> Design:
    try:
        rv = KtgRaw()
        _lib.check_ktg(lib.ktg_raw_view(handle, C.byref(rv)))
        fv = KtgFlat()
        _lib.check_ktg(lib.ktg_flat_view(handle, C.byref(fv)))
        return Design(raw=_raw_from_view(rv), flat=_flat_from_view(fv), name=name)
    finally:
        lib.ktg_free(handle)


def generate(num_cells: int, num_nets: int, num_pads: int, seed: int = 1, **kw) -> Design:
    spec = make_spec(num_cells, num_nets, num_pads, seed=seed, **kw)
    h = C.c_void_p()
    _lib.check_ktg(lib.ktg_generate(C.byref(spec), C.byref(h)))
    d = _take(h, f"synthetic-{num_cells}c-{num_nets}n-seed{seed}")
    d.meta = dict(num_cells=num_cells, num_nets=num_nets, num_pads=num_pads, seed=seed, **kw)
    return d


def workload(name: str, seed: int = 1) -> Design:
    d = generate(seed=seed, **WORKLOADS[name])
    d.name = name
    return d


def collect(raw: Raw, name: str = "") -> Design:
    """Flatten a raw design (simpl_dm.cc:16-85)."""
    h = C.c_void_p()
    c = raw.to_c()
    _lib.check_ktg(lib.ktg_collect(C.byref(c), C.byref(h)))
    return _take(h, name)


def make_raw(cells: list[tuple[float, float, float, float, bool]],
             nets: list[list[tuple[int, float, float]]], die=(0.0, 0.0, 100.0, 100.0),
             rows: list[tuple[float, float, float, float]] | None = None,
             weights: list[float] | None = None) -> Raw:
    """Small hand-made raw design: cells = (x, y, w, h, fixed); nets = [(vertex, offx, offy)];
    rows = (y, height, x0, x1) with site pitch 1."""
    V = len(cells)
    starts = [0]
    for net in nets:
        starts.append(starts[-1] + len(net))
    pins = [p for net in nets for p in net]
    rows = rows or []
    return Raw(
        is_fixed=np.array([1 if c[4] else 0 for c in cells], np.uint8),
        pos_x=np.array([c[0] for c in cells], np.float64),
        pos_y=np.array([c[1] for c in cells], np.float64),
        width=np.array([c[2] for c in cells], np.float64),
        height=np.array([c[3] for c in cells], np.float64),
        net_start=np.array(starts, np.uint64),
        pin_vertex=np.array([p[0] for p in pins], np.uint32).reshape(-1),
        pin_off_x=np.array([p[1] for p in pins], np.float64).reshape(-1),
        pin_off_y=np.array([p[2] for p in pins], np.float64).reshape(-1),
        net_weight=np.array(weights if weights is not None else [1.0] * len(nets), np.float64),
        row_y=np.array([r[0] for r in rows], np.float64),
        row_height=np.array([r[1] for r in rows], np.float64),
        row_x0=np.array([r[2] for r in rows], np.float64),
        row_x1=np.array([r[3] for r in rows], np.float64),
        row_pitch=np.ones(len(rows), np.float64),
        die=np.array(die, np.float64),
    ) if V else None


def read_bookshelf(path) -> Design:
    """A Bookshelf design directory (<dir>/<name>.nodes/.nets/.pl/.scl/.wts,
    each optionally gzipped), read and flattened by the library's C++ reader
    (as the reference's Bookshelf adaptor, src/adaptor/bookshelfToKTAdaptor.cc,
    and getPlacementDieBox, src/datamodel/kt_die.cc:45-103).  The vertex
    names are kept in `meta["names"]`."""
    h = C.c_void_p()
    _lib.check_ktg(lib.ktg_read_bookshelf(str(path).encode(), C.byref(h)))
    blob, size = C.c_void_p(), C.c_uint64()
    try:
        _lib.check_ktg(lib.ktg_vertex_names(h, C.byref(blob), C.byref(size)))
        names = C.string_at(blob.value, size.value).decode().split("\n")[:-1] if size.value else []
    except BaseException:
        lib.ktg_free(h)
        raise
    d = _take(h, Path(str(path).rstrip("/")).name)
    d.meta = dict(names=names, source=str(path))
    return d


def write_pl(d: Design, path, x, y) -> None:
    """A Bookshelf placement file of `d` with the movable cells at (x, y)
    (lower-left corners), as the reference's writer (src/kt_flowMgr.cc:508-521)."""
    names = d.meta.get("names")
    if names is None or len(names) != d.flat.num_vertices:
        raise ValueError("write_pl: the design has no vertex names (read it with read_bookshelf)")
    blob = "".join(n + "\n" for n in names).encode()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    if len(x) != d.n or len(y) != d.n:
        raise ValueError("write_pl: x and y need one entry per movable cell")
    f = d.flat
    _lib.check_ktg(lib.ktg_write_pl(str(path).encode(), f.num_vertices, blob, len(blob),
                                    _lib.ptr(np.ascontiguousarray(f.var_of_vertex)),
                                    _lib.ptr(np.ascontiguousarray(f.fixed_x)),
                                    _lib.ptr(np.ascontiguousarray(f.fixed_y)), _lib.ptr(x), _lib.ptr(y)))


def seed_uniform(d: Design, seed: int = 20240607):
    """Random start placement (simpl_placer.cc:226-237), host C++ in libktplace_amd."""
    n = d.n
    x = np.zeros(n)
    y = np.zeros(n)
    die = np.ascontiguousarray(d.die, np.float64)
    _lib.check_ktg(lib.ktg_seed_uniform(seed, n, _lib.ptr(die), _lib.ptr(d.flat.width),
                                        _lib.ptr(d.flat.height), _lib.ptr(x), _lib.ptr(y)))
    return x, y
