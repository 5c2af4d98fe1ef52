"""dyncolor_amd -- MI355X (gfx950) dynamic graph recoloring.

Python binding (ctypes) of the C ABI in include/dyncolor.h, used by the tests
and the benchmark.  All compute runs in libdyncolor.so's HIP kernels; if the
library is missing this module raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdyncolor.so")

RULE_INDEPENDENT = 0
RULE_LDF = 1
DRIVER_PLAIN = 0
DRIVER_FUSED = 1
DRIVER_GRAPH = 2

OK = 0
ERR_INVALID = -1
ERR_VERTEX = -2
ERR_HIP = -3
ERR_NOMEM = -4
ERR_ROUNDS = -5
ERR_STATE = -6
ERR_ADMIT = -7

# every symbol include/dyncolor.h declares
EXPORTS = [
    "dc_options_default", "dc_graph_create", "dc_partition_create", "dc_graph_destroy",
    "dc_graph_info", "dc_colors_set", "dc_colors_get", "dc_static_color", "dc_apply_batch",
    "dc_batch_create", "dc_batch_destroy", "dc_apply_device_batch", "dc_recolor",
    "dc_process_batch", "dc_validate", "dc_trace_get", "dc_neighbors", "dc_prop_get",
    "dc_compact", "dc_compact_noalloc", "dc_check_consistency",
    "dc_tuning_default", "dc_set_tuning", "dc_get_tuning", "dc_device_info", "dc_set_profiling", "dc_set_driver", "dc_set_parts", "dc_synchronize", "dc_last_error", "dc_device_count",
    "dc_gen_er", "dc_gen_rmat", "dc_gen_batches",
    "dc_server_create", "dc_server_destroy", "dc_server_run", "dc_server_request_bytes",
    "dc_lanes_step",
    "dc_partition_plan_create", "dc_partition_plan_info", "dc_partition_plan_arrays",
    "dc_partition_plan_border", "dc_partition_create_global", "dc_partition_plan_destroy",
]


class Options(C.Structure):
    _fields_ = [("slack_frac", C.c_float), ("slack_min", C.c_int32), ("max_rounds", C.c_int32),
                ("driver", C.c_int32), ("device", C.c_int32), ("spare_slots", C.c_int32),
                ("exact_rounds", C.c_int32), ("list_limit", C.c_int32),
                ("pull_mode", C.c_int32), ("group_deletions", C.c_int32),
                ("heavy_slots", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("loop_blocks", C.c_int32), ("light_blocks", C.c_int32), ("light_c", C.c_int32),
                ("list_c", C.c_int32), ("list_slots", C.c_int64), ("heavy_div", C.c_int32),
                ("group_deletions", C.c_int32), ("ldf_loop", C.c_int32),
                ("reserved0", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class Stats(C.Structure):
    _fields_ = [("updates", C.c_int64), ("insertions", C.c_int64), ("deletions", C.c_int64),
                ("phases", C.c_int32), ("loop_rounds", C.c_int32), ("evaluated", C.c_int64),
                ("changed", C.c_int64), ("frontier_sum", C.c_int64),
                ("slots_scanned", C.c_int64), ("live_scanned", C.c_int64),
                ("recolor_bytes", C.c_int64), ("ms_apply", C.c_double),
                ("ms_recolor", C.c_double), ("ms_recolor_kernel", C.c_double),
                ("recolor_launches", C.c_int32), ("regrows", C.c_int32),
                ("ms_count_pass", C.c_double), ("count_pass_bytes", C.c_int64),
                ("count_pass_launches", C.c_int32), ("loop_settled", C.c_int32),
                ("pull_rounds", C.c_int32), ("heavy_push_rounds", C.c_int32),
                ("heavy_push_listed", C.c_int32), ("graph_captures", C.c_int32),
                ("lc_vertices", C.c_int64), ("lc_slots", C.c_int64), ("del_row_slots", C.c_int64),
                ("ms_round_a", C.c_double), ("moved_slots", C.c_int64),
                ("slots_rewritten", C.c_int32), ("reserved1", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class ServeStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("lanes", C.c_int32), ("served", C.c_int64),
                ("rejected", C.c_int64), ("lines", C.c_int64), ("max_step_lines", C.c_int64),
                ("elapsed_s", C.c_double), ("busy_s", C.c_double), ("budget_bytes", C.c_int64),
                ("max_inflight_bytes", C.c_int64), ("admit_waits", C.c_int64),
                ("reserved", C.c_int64 * 1)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class PartitionInfo(C.Structure):
    _fields_ = [("n_global", C.c_int32), ("rank", C.c_int32), ("n_owned", C.c_int32),
                ("n_ghost", C.c_int32), ("n_border", C.c_int32), ("n_parts", C.c_int32),
                ("m_local", C.c_int64), ("m_internal", C.c_int64), ("m_cross", C.c_int64),
                ("m_dropped", C.c_int64), ("k_local", C.c_int64), ("k_del", C.c_int64),
                ("k_ins", C.c_int64), ("k_xdel", C.c_int64), ("k_xins", C.c_int64),
                ("k_dropped", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DynColorError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dyncolor error {code}: {msg}")
        self.code = code


_lib = None
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    """Load libdyncolor.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run python -m dyncolor_amd.build")
    L = C.CDLL(LIB_PATH)
    sig = {
        "dc_options_default": (None, [C.POINTER(Options)]),
        "dc_graph_create": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp, C.POINTER(Options), C.POINTER(_vp)]),
        "dc_partition_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, _vp, _vp, _vp,
                                          C.POINTER(Options), C.POINTER(_vp)]),
        "dc_graph_destroy": (C.c_int, [_vp]),
        "dc_graph_info": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "dc_colors_set": (C.c_int, [_vp, _i32p]),
        "dc_colors_get": (C.c_int, [_vp, _i32p]),
        "dc_static_color": (C.c_int, [_vp, C.c_int32]),
        "dc_apply_batch": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp]),
        "dc_batch_create": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, C.POINTER(_vp)]),
        "dc_batch_destroy": (C.c_int, [_vp]),
        "dc_apply_device_batch": (C.c_int, [_vp, _vp]),
        "dc_recolor": (C.c_int, [_vp, C.c_int32, C.POINTER(Stats)]),
        "dc_process_batch": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, C.POINTER(Stats)]),
        "dc_validate": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
        "dc_trace_get": (C.c_int, [_vp, _i64p, C.c_int32, C.POINTER(C.c_int32)]),
        "dc_neighbors": (C.c_int, [_vp, C.c_int32, _i32p, C.c_int64, C.POINTER(C.c_int64)]),
        "dc_prop_get": (C.c_int, [_vp, _i32p]),
        "dc_compact": (C.c_int, [_vp]),
        "dc_compact_noalloc": (C.c_int, [_vp]),
        "dc_tuning_default": (None, [C.POINTER(Tuning)]),
        "dc_device_info": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32,
                                     C.POINTER(C.c_int32)]),
        "dc_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
        "dc_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
        "dc_check_consistency": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
        "dc_set_profiling": (C.c_int, [_vp, C.c_int32]),
        "dc_set_driver": (C.c_int, [_vp, C.c_int32, C.c_int32]),
        "dc_set_parts": (C.c_int, [_vp, C.c_int32, _i32p]),
        "dc_synchronize": (C.c_int, [_vp]),
        "dc_last_error": (C.c_char_p, []),
        "dc_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
        "dc_gen_er": (C.c_int, [C.c_int32, C.c_int64, C.c_uint64, _i32p, _i32p]),
        "dc_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                  C.c_uint64, _i32p, _i32p, C.POINTER(C.c_int64)]),
        "dc_gen_batches": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp, C.c_int32, C.c_int64,
                                     C.c_double, C.c_uint64, _i32p, _i32p, _u8p]),
        "dc_server_create": (C.c_int, [C.c_int32, C.POINTER(_vp), C.c_int32, _i32p, _i32p, _i32p,
                                       C.c_int64, C.c_int32, C.POINTER(_vp)]),
        "dc_server_request_bytes": (C.c_int64, [C.c_int64, C.c_int64]),
        "dc_partition_plan_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _vp, _vp, C.c_int64,
                                               _vp, _vp, C.c_int64, _vp, _vp, _vp, C.POINTER(_vp)]),
        "dc_partition_plan_info": (C.c_int, [_vp, C.POINTER(PartitionInfo)]),
        "dc_partition_plan_arrays": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "dc_partition_plan_border": (C.c_int, [_vp, _i64p, C.c_int32]),
        "dc_partition_create_global": (C.c_int, [_vp, _i32p, C.POINTER(Options), C.POINTER(_vp)]),
        "dc_partition_plan_destroy": (C.c_int, [_vp]),
        "dc_server_destroy": (C.c_int, [_vp]),
        "dc_lanes_step": (C.c_int, [C.c_int32, _vp, _vp, C.c_int32, _vp]),
        "dc_server_run": (C.c_int, [_vp, C.c_int64, _i32p, _i64p, _vp, _vp, _vp,
                                    np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"),
                                    np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"),
                                    _i32p, C.POINTER(ServeStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise DynColorError(rc, lib().dc_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _arr(x, dt):
    return np.ascontiguousarray(np.asarray(x, dtype=dt))


def device_count() -> int:
    c = C.c_int32(0)
    rc = lib().dc_device_count(C.byref(c))
    return int(c.value) if rc == OK else 0


def device_info(device=0):
    """(name, architecture, compute units) of a HIP device."""
    name = C.create_string_buffer(256)
    arch = C.create_string_buffer(256)
    cus = C.c_int32(0)
    _check(lib().dc_device_info(device, name, 256, arch, 256, C.byref(cus)))
    return name.value.decode(errors="replace"), arch.value.decode(errors="replace"), int(cus.value)


def tuning(**kw) -> Tuning:
    t = Tuning()
    lib().dc_tuning_default(C.byref(t))
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def options(**kw) -> Options:
    o = Options()
    lib().dc_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ---------------------------------------------------------------- generators
def gen_er(n, m, seed):
    u = np.zeros(max(m, 1), np.int32)
    v = np.zeros(max(m, 1), np.int32)
    _check(lib().dc_gen_er(n, m, seed, u, v))
    return u[:m], v[:m]


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    samples = edge_factor << scale
    u = np.zeros(samples, np.int32)
    v = np.zeros(samples, np.int32)
    m = C.c_int64(0)
    _check(lib().dc_gen_rmat(scale, edge_factor, a, b, c, seed, u, v, C.byref(m)))
    return u[: m.value].copy(), v[: m.value].copy()


def gen_batches(n, u, v, nb, bs, ins_frac, seed):
    u = _arr(u, np.int32)
    v = _arr(v, np.int32)
    tot = max(nb * bs, 1)
    bu = np.zeros(tot, np.int32)
    bv = np.zeros(tot, np.int32)
    bi = np.zeros(tot, np.uint8)
    _check(lib().dc_gen_batches(n, u.size, _ptr(u), _ptr(v), nb, bs, ins_frac, seed, bu, bv, bi))
    return [(bu[i * bs:(i + 1) * bs], bv[i * bs:(i + 1) * bs], bi[i * bs:(i + 1) * bs])
            for i in range(nb)]


# --------------------------------------------------------------------- graph
class Graph:
    """A device-resident graph and its coloring (one dc_graph handle)."""

    def __init__(self, n, u, v, opts: Options | None = None, n_ghost=0, ghost_colors=None):
        u = _arr(u, np.int32)
        v = _arr(v, np.int32)
        h = C.c_void_p()
        o = C.byref(opts) if opts is not None else None
        if n_ghost:
            gc = _arr(ghost_colors, np.int32)
            _check(lib().dc_partition_create(n, n_ghost, u.size, _ptr(u), _ptr(v), _ptr(gc), o,
                                             C.byref(h)))
            self.n = n + n_ghost
            self.n_owned = n
        else:
            _check(lib().dc_graph_create(n, u.size, _ptr(u), _ptr(v), o, C.byref(h)))
            self.n = n
            self.n_owned = n
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().dc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @classmethod
    def _wrap(cls, h, n, n_owned):
        g = cls.__new__(cls)
        g._h = h
        g.n = n
        g.n_owned = n_owned
        return g

    @property
    def handle(self):
        return self._h

    def colors(self) -> np.ndarray:
        out = np.zeros(self.n, np.int32)
        _check(lib().dc_colors_get(self._h, out))
        return out

    def set_colors(self, colors):
        _check(lib().dc_colors_set(self._h, _arr(colors, np.int32)))

    def static_color(self, rule=RULE_INDEPENDENT):
        _check(lib().dc_static_color(self._h, rule))

    def apply_batch(self, u, v, ins):
        u = _arr(u, np.int32)
        v = _arr(v, np.int32)
        ins = _arr(ins, np.uint8)
        _check(lib().dc_apply_batch(self._h, u.size, _ptr(u), _ptr(v), _ptr(ins)))

    def upload_batch(self, u, v, ins) -> "DeviceBatch":
        return DeviceBatch(self, u, v, ins)

    def apply_device_batch(self, b: "DeviceBatch"):
        _check(lib().dc_apply_device_batch(self._h, b._h))

    def recolor(self, rule=RULE_INDEPENDENT) -> Stats:
        st = Stats()
        _check(lib().dc_recolor(self._h, rule, C.byref(st)))
        return st

    def process_batch(self, u, v, ins, rule=RULE_INDEPENDENT) -> Stats:
        self.apply_batch(u, v, ins)
        return self.recolor(rule)

    def validate(self):
        mono = C.c_int64(0)
        mc = C.c_int32(0)
        _check(lib().dc_validate(self._h, C.byref(mono), C.byref(mc)))
        return int(mono.value), int(mc.value)

    def trace(self) -> np.ndarray:
        n = C.c_int32(0)
        buf = np.zeros(4, np.int64)
        _check(lib().dc_trace_get(self._h, buf, 0, C.byref(n)))
        buf = np.zeros(4 * max(n.value, 1), np.int64)
        _check(lib().dc_trace_get(self._h, buf, n.value, C.byref(n)))
        return buf[: 4 * n.value].reshape(-1, 4)

    def neighbors(self, vtx) -> np.ndarray:
        d = C.c_int64(0)
        tmp = np.zeros(1, np.int32)
        _check(lib().dc_neighbors(self._h, vtx, tmp, 0, C.byref(d)))
        out = np.zeros(max(d.value, 1), np.int32)
        _check(lib().dc_neighbors(self._h, vtx, out, out.size, C.byref(d)))
        return out[: d.value]

    def prop(self) -> np.ndarray:
        out = np.zeros(self.n, np.int32)
        _check(lib().dc_prop_get(self._h, out))
        return out

    def compact(self):
        _check(lib().dc_compact(self._h))

    def compact_noalloc(self):
        """dc_compact_noalloc: the same compaction with its slot map in the
        library's static device array (no device allocation)."""
        _check(lib().dc_compact_noalloc(self._h))

    def info(self):
        n = C.c_int32(0)
        no = C.c_int32(0)
        live = C.c_int64(0)
        cap = C.c_int64(0)
        _check(lib().dc_graph_info(self._h, C.byref(n), C.byref(no), C.byref(live), C.byref(cap)))
        return {"n": n.value, "n_owned": no.value, "capacity": cap.value}

    def check_consistency(self) -> int:
        bad = C.c_int64(0)
        _check(lib().dc_check_consistency(self._h, C.byref(bad)))
        return int(bad.value)

    def set_driver(self, driver, blocks_per_cu=0):
        _check(lib().dc_set_driver(self._h, driver, blocks_per_cu))

    def set_parts(self, offsets):
        """dc_set_parts: the graph is a union of disjoint parts (offsets of
        length nparts + 1; an empty array clears)."""
        off = _arr(offsets, np.int32)
        _check(lib().dc_set_parts(self._h, max(off.size - 1, 0), off))

    def set_tuning(self, t: Tuning | None = None, **kw):
        """dc_set_tuning: t, or the current tuning with the given fields changed."""
        if t is None:
            t = self.get_tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().dc_set_tuning(self._h, C.byref(t)))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().dc_get_tuning(self._h, C.byref(t)))
        return t

    def set_profiling(self, on=True):
        _check(lib().dc_set_profiling(self._h, 1 if on else 0))

    def synchronize(self):
        _check(lib().dc_synchronize(self._h))


class PartitionPlan:
    """One rank's share of a partitioned graph from global IDs
    (dc_partition_plan_*; host only until create_graph)."""

    def __init__(self, n_global, rank, owned_ids, part_of, u, v, cu, cv, cins):
        owned = _arr(owned_ids, np.int32)
        part = None if part_of is None else _arr(part_of, np.int32)
        u, v = _arr(u, np.int32), _arr(v, np.int32)
        cu, cv, cins = _arr(cu, np.int32), _arr(cv, np.int32), _arr(cins, np.uint8)
        h = C.c_void_p()
        _check(lib().dc_partition_plan_create(n_global, rank, owned.size, _ptr(owned), _ptr(part),
                                              u.size, _ptr(u), _ptr(v), cu.size, _ptr(cu), _ptr(cv),
                                              _ptr(cins), C.byref(h)))
        self._h = h
        inf = PartitionInfo()
        _check(lib().dc_partition_plan_info(h, C.byref(inf)))
        self.info = inf.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().dc_partition_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """(local_to_global, u, v, bu, bv, bins) in local IDs."""
        i = self.info
        l2g = np.zeros(max(i["n_owned"] + i["n_ghost"], 1), np.int32)
        u = np.zeros(max(i["m_local"], 1), np.int32)
        v = np.zeros_like(u)
        bu = np.zeros(max(i["k_local"], 1), np.int32)
        bv = np.zeros_like(bu)
        bi = np.zeros(max(i["k_local"], 1), np.uint8)
        _check(lib().dc_partition_plan_arrays(self._h, _ptr(l2g), _ptr(u), _ptr(v), _ptr(bu),
                                              _ptr(bv), _ptr(bi)))
        n, m, k = i["n_owned"] + i["n_ghost"], i["m_local"], i["k_local"]
        return l2g[:n], u[:m], v[:m], bu[:k], bv[:k], bi[:k]

    def border(self):
        out = np.zeros(max(self.info["n_parts"], 1), np.int64)
        _check(lib().dc_partition_plan_border(self._h, out, self.info["n_parts"]))
        return out[: self.info["n_parts"]]

    def create_graph(self, colors_global, opts: Options | None = None) -> "Graph":
        h = C.c_void_p()
        o = C.byref(opts) if opts is not None else None
        _check(lib().dc_partition_create_global(self._h, _arr(colors_global, np.int32), o,
                                                C.byref(h)))
        return Graph._wrap(h, self.info["n_owned"] + self.info["n_ghost"], self.info["n_owned"])


class DeviceBatch:
    """An update batch validated and copied to HBM once."""

    def __init__(self, g: Graph, u, v, ins):
        u = _arr(u, np.int32)
        v = _arr(v, np.int32)
        ins = _arr(ins, np.uint8)
        h = C.c_void_p()
        _check(lib().dc_batch_create(g.handle, u.size, _ptr(u), _ptr(v), _ptr(ins), C.byref(h)))
        self._h = h
        self.k = int(u.size)

    def close(self):
        if getattr(self, "_h", None):
            lib().dc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ many graphs ---
def union_edges(graphs):
    """Disjoint union of [(n, u, v), ...]: returns (n_total, u, v, offsets)."""
    offs = np.zeros(len(graphs) + 1, np.int64)
    for i, (n, _, _) in enumerate(graphs):
        offs[i + 1] = offs[i] + n
    if offs[-1] > 0x7ffffffe:
        raise ValueError("union exceeds int32 vertex IDs")
    u = np.concatenate([np.asarray(gu, np.int64) + offs[i] for i, (_, gu, _) in enumerate(graphs)])
    v = np.concatenate([np.asarray(gv, np.int64) + offs[i] for i, (_, _, gv) in enumerate(graphs)])
    return int(offs[-1]), u.astype(np.int32), v.astype(np.int32), offs.astype(np.int32)


def union_batch(batches, offsets):
    """Concatenate per-graph batches [(u, v, ins), ...] into union IDs."""
    u = np.concatenate([np.asarray(b[0], np.int64) + offsets[i] for i, b in enumerate(batches)])
    v = np.concatenate([np.asarray(b[1], np.int64) + offsets[i] for i, b in enumerate(batches)])
    ins = np.concatenate([np.asarray(b[2], np.uint8) for b in batches])
    return u.astype(np.int32), v.astype(np.int32), ins


class Server:
    """Continuous batching over resident graphs (dc_server_*): lanes = arena
    Graphs, graph g at [first[g], first[g] + size[g]) of lanes[lane_of[g]]."""

    def __init__(self, lanes, lane_of, first, size, budget_bytes=0, rule=RULE_INDEPENDENT):
        self.lanes = list(lanes)
        arr = (C.c_void_p * len(self.lanes))(*[g.handle for g in self.lanes])
        self._lane_of = _arr(lane_of, np.int32)
        self._first = _arr(first, np.int32)
        self._size = _arr(size, np.int32)
        h = C.c_void_p()
        _check(lib().dc_server_create(len(self.lanes), arr, self._lane_of.size, self._lane_of,
                                      self._first, self._size, int(budget_bytes), rule,
                                      C.byref(h)))
        self._h = h

    @classmethod
    def one_lane(cls, arena, offsets, budget_bytes=0, rule=RULE_INDEPENDENT):
        """All graphs in one arena with the offsets of union_edges."""
        off = _arr(offsets, np.int64)
        n = off.size - 1
        return cls([arena], np.zeros(n, np.int32), off[:-1], np.diff(off), budget_bytes, rule)

    def close(self):
        if getattr(self, "_h", None):
            lib().dc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, gid, roff, u, v, ins, arrival):
        gid = _arr(gid, np.int32)
        roff = _arr(roff, np.int64)
        u = _arr(u, np.int32)
        v = _arr(v, np.int32)
        ins = _arr(ins, np.uint8)
        arrival = _arr(arrival, np.float64)
        nreq = gid.size
        lat = np.zeros(max(nreq, 1), np.float64)
        step = np.zeros(max(nreq, 1), np.int32)
        st = ServeStats()
        _check(lib().dc_server_run(self._h, nreq, gid, roff, _ptr(u), _ptr(v), _ptr(ins),
                                   arrival, lat, step, C.byref(st)))
        return lat[:nreq], step[:nreq], st


def lanes_step(arenas, batches, rule=RULE_INDEPENDENT):
    """dc_lanes_step: lane l applies batches[l] (device batches of arenas[l])
    and recolors; the lanes run concurrently.  Returns each lane's Stats."""
    if len(arenas) != len(batches) or not arenas:
        raise ValueError("one device batch per arena")
    k = len(arenas)
    ah = (C.c_void_p * k)(*[a.handle for a in arenas])
    bh = (C.c_void_p * k)(*[b._h for b in batches])
    st = (Stats * k)()
    _check(lib().dc_lanes_step(k, ah, bh, rule, st))
    return list(st)


def request_bytes(lines, insertions):
    return int(lib().dc_server_request_bytes(int(lines), int(insertions)))
