"""ctypes binding of the CPU oracle (oracle/dc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by ``bench.py``'s CPU-baseline leg, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdcoracle.so")

RULE_INDEPENDENT = 0
RULE_LDF = 1

PH_DEL, PH_INS, PH_CHANGE, PH_XDEL, PH_XINS, PH_LOOP = 1, 2, 3, 4, 5, 10

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_graph_create.restype = C.c_void_p
        L.orc_graph_create.argtypes = [C.c_int32, C.c_int32, C.c_int64, _i32p, _i32p]
        L.orc_graph_destroy.argtypes = [C.c_void_p]
        L.orc_apply_batch.restype = C.c_int
        L.orc_apply_batch.argtypes = [C.c_void_p, _i32p, C.c_int64, _i32p, _i32p, _u8p]
        L.orc_recolor.restype = C.c_int
        L.orc_recolor.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, C.c_int32,
                                  _i64p, C.c_int32, C.POINTER(C.c_int32)]
        L.orc_recolor_par.restype = C.c_int
        L.orc_recolor_par.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, C.c_int32,
                                      _i64p, C.c_int32, C.POINTER(C.c_int32), C.c_int32]
        L.orc_static_greedy.argtypes = [C.c_void_p, C.c_int, _i32p]
        L.orc_validate.restype = C.c_int64
        L.orc_validate.argtypes = [C.c_void_p, _i32p, C.POINTER(C.c_int32)]
        L.orc_mex.restype = C.c_int32
        L.orc_mex.argtypes = [_i32p, C.c_int64]
        L.orc_live_degree.restype = C.c_int64
        L.orc_live_degree.argtypes = [C.c_void_p, C.c_int32]
        L.orc_neighbors.restype = C.c_int64
        L.orc_neighbors.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int64]
        L.orc_get_prop.argtypes = [C.c_void_p, _i32p]
        _lib = L
    return _lib


def mex(colors) -> int:
    a = np.ascontiguousarray(colors, dtype=np.int32)
    if a.size == 0:
        a = np.zeros(1, dtype=np.int32)
        return int(lib().orc_mex(a, 0))
    return int(lib().orc_mex(a, a.size))


class OracleGraph:
    """One graph + its coloring, replayed on the CPU."""

    def __init__(self, n, u, v, n_owned=None):
        self.n = int(n)
        self.n_owned = self.n if n_owned is None else int(n_owned)
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        if u.size == 0:
            u = np.zeros(1, dtype=np.int32)
            v = np.zeros(1, dtype=np.int32)
            m = 0
        else:
            m = u.size
        self._h = lib().orc_graph_create(self.n, self.n_owned, m, u, v)
        if not self._h:
            raise ValueError("oracle: invalid graph")
        self.color = np.full(self.n, -1, dtype=np.int32)
        self.prev = np.full(self.n, -1, dtype=np.int32)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_graph_destroy(h)
            self._h = None

    def static_color(self, rule):
        lib().orc_static_greedy(self._h, rule, self.color)
        return self.color.copy()

    def set_colors(self, colors):
        self.color[:] = np.asarray(colors, dtype=np.int32)

    def apply_batch(self, u, v, ins):
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        ins = np.ascontiguousarray(ins, dtype=np.uint8)
        k = u.size
        if k == 0:
            u = np.zeros(1, dtype=np.int32)
            v = np.zeros(1, dtype=np.int32)
            ins = np.zeros(1, dtype=np.uint8)
        rc = lib().orc_apply_batch(self._h, self.color, k, u, v, ins)
        if rc != 0:
            raise ValueError(f"oracle: apply_batch failed ({rc})")

    def recolor(self, rule, max_loop=100000, trace_cap=100000):
        trace = np.zeros(4 * trace_cap, dtype=np.int64)
        nph = C.c_int32(0)
        rc = lib().orc_recolor(self._h, rule, self.color, self.prev, max_loop,
                               trace, trace_cap, C.byref(nph))
        phases = trace[: 4 * min(nph.value, trace_cap)].reshape(-1, 4)
        if rc != 0:
            raise RuntimeError(f"oracle: recolor failed ({rc})")
        return phases

    def recolor_par(self, rule, threads, max_loop=100000, trace_cap=100000):
        """The independent rule's rounds split over `threads` OpenMP threads
        (same colours and trace as recolor)."""
        trace = np.zeros(4 * trace_cap, dtype=np.int64)
        nph = C.c_int32(0)
        rc = lib().orc_recolor_par(self._h, rule, self.color, self.prev, max_loop,
                                   trace, trace_cap, C.byref(nph), int(threads))
        phases = trace[: 4 * min(nph.value, trace_cap)].reshape(-1, 4)
        if rc != 0:
            raise RuntimeError(f"oracle: parallel recolor failed ({rc})")
        return phases

    def validate(self):
        mc = C.c_int32(0)
        bad = lib().orc_validate(self._h, self.color, C.byref(mc))
        return int(bad), int(mc.value)

    def neighbors(self, v):
        d = lib().orc_live_degree(self._h, v)
        out = np.zeros(max(d, 1), dtype=np.int32)
        lib().orc_neighbors(self._h, v, out, out.size)
        return out[:d]

    def prop(self):
        p = np.zeros(self.n, dtype=np.int32)
        lib().orc_get_prop(self._h, p)
        return p
