"""ctypes binding of libktplace_amd.so (the C ABI in include/*.h).

The library is loaded from the package directory only.  If it is missing the
import fails loudly: there is no CPU fallback for anything in this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libktplace_amd.so"

NO_VAR = 0xFFFFFFFF


class KtgSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_cells", C.c_uint32),
        ("num_nets", C.c_uint32),
        ("num_pads", C.c_uint32),
        ("num_macros", C.c_uint32),
        ("utilization", C.c_double),
        ("max_degree", C.c_uint32),
        ("high_fanout_frac", C.c_double),
        ("local_frac", C.c_double),
    ]


class KtgRaw(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("num_movable", C.c_uint32),
        ("num_nets", C.c_uint32),
        ("num_pins", C.c_uint64),
        ("num_rows", C.c_uint32),
        ("is_fixed", C.c_void_p),
        ("pos_x", C.c_void_p),
        ("pos_y", C.c_void_p),
        ("width", C.c_void_p),
        ("height", C.c_void_p),
        ("net_start", C.c_void_p),
        ("pin_vertex", C.c_void_p),
        ("pin_off_x", C.c_void_p),
        ("pin_off_y", C.c_void_p),
        ("net_weight", C.c_void_p),
        ("row_y", C.c_void_p),
        ("row_height", C.c_void_p),
        ("row_x0", C.c_void_p),
        ("row_x1", C.c_void_p),
        ("row_pitch", C.c_void_p),
        ("die", C.c_double * 4),
    ]


class KtgFlat(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("num_movable", C.c_uint32),
        ("num_nets", C.c_uint32),
        ("num_pins", C.c_uint64),
        ("num_cell_pins", C.c_uint64),
        ("num_obstacles", C.c_uint32),
        ("net_start", C.c_void_p),
        ("pin_cell", C.c_void_p),
        ("pin_off_x", C.c_void_p),
        ("pin_off_y", C.c_void_p),
        ("net_weight", C.c_void_p),
        ("var_of_vertex", C.c_void_p),
        ("mov_vertex", C.c_void_p),
        ("fixed_x", C.c_void_p),
        ("fixed_y", C.c_void_p),
        ("pin_start", C.c_void_p),
        ("cell_pin_net", C.c_void_p),
        ("cell_pin_idx", C.c_void_p),
        ("width", C.c_void_p),
        ("height", C.c_void_p),
        ("area", C.c_void_p),
        ("obstacles", C.c_void_p),
    ]


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    return C.CDLL(str(LIB_PATH))


lib = _load()


def _proto(name: str, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


P = C.c_void_p
I32 = C.c_int
U32 = C.c_uint32
U64 = C.c_uint64
F64 = C.c_double

_proto("ktg_default_spec", I32, C.POINTER(KtgSpec), U32, U32, U32, U64)
_proto("ktg_generate", I32, C.POINTER(KtgSpec), C.POINTER(P))
_proto("ktg_raw_view", I32, P, C.POINTER(KtgRaw))
_proto("ktg_collect", I32, C.POINTER(KtgRaw), C.POINTER(P))
_proto("ktg_flat_view", I32, P, C.POINTER(KtgFlat))
_proto("ktg_seed_uniform", I32, U64, U64, P, P, P, P, P)
_proto("ktg_read_bookshelf", I32, C.c_char_p, C.POINTER(P))
_proto("ktg_vertex_names", I32, P, C.POINTER(P), C.POINTER(U64))
_proto("ktg_write_pl", I32, C.c_char_p, U32, C.c_char_p, U64, P, P, P, P, P)
_proto("ktg_free", None, P)
_proto("ktg_last_error", C.c_char_p)


def check_ktg(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib.ktg_last_error().decode())


def ptr(a: np.ndarray | None) -> int | None:
    """Address of a contiguous numpy array (None passes a null pointer)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def view(addr: int, count: int, dtype) -> np.ndarray:
    """Copy `count` elements of `dtype` from a C pointer into a new numpy array."""
    dtype = np.dtype(dtype)
    if count == 0:
        return np.zeros(0, dtype)
    buf = (C.c_char * (count * dtype.itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=count).copy()
