"""CPU-side checks of libdyncolor.so: it loads, exports every symbol the C
header declares, validates inputs before touching a device, and its
deterministic generators produce what they promise.  No GPU compute."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dyncolor_amd as d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(dc_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(d.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations parsed"
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(d.EXPORTS) | (declared - set(d.EXPORTS))
    assert set(d.EXPORTS) <= declared


def test_invalid_inputs_rejected_before_device():
    L = d.lib()
    h = C.c_void_p()
    u = np.array([0, 1, 5], np.int32)
    v = np.array([1, 2, 2], np.int32)
    assert L.dc_graph_create(4, 3, d._ptr(u), d._ptr(v), None, C.byref(h)) == d.ERR_VERTEX
    u = np.array([0, 2], np.int32)
    v = np.array([1, 2], np.int32)
    assert L.dc_graph_create(4, 2, d._ptr(u), d._ptr(v), None, C.byref(h)) == d.ERR_VERTEX
    assert b"self loop" in L.dc_last_error()
    assert L.dc_graph_create(-1, 0, None, None, None, C.byref(h)) == d.ERR_INVALID
    assert L.dc_graph_create(4, 2, None, None, None, C.byref(h)) == d.ERR_INVALID
    # ghost-ghost edge in a partition share
    u = np.array([3], np.int32)
    v = np.array([4], np.int32)
    gc = np.array([0, 1, 2], np.int32)
    assert L.dc_partition_create(2, 3, 1, d._ptr(u), d._ptr(v), d._ptr(gc), None,
                                 C.byref(h)) == d.ERR_VERTEX
    assert L.dc_recolor(None, 0, None) == d.ERR_INVALID


def test_er_generator():
    u, v = d.gen_er(10000, 50000, 1)
    assert u.size == 50000
    assert (u != v).all()
    assert ((u >= 0) & (u < 10000) & (v >= 0) & (v < 10000)).all()
    keys = np.minimum(u, v).astype(np.int64) << 32 | np.maximum(u, v)
    assert np.unique(keys).size == 50000
    u2, v2 = d.gen_er(10000, 50000, 1)
    assert (u == u2).all() and (v == v2).all()
    u3, _ = d.gen_er(10000, 50000, 2)
    assert not (u == u3).all()
    # mean degree 10, roughly uniform
    deg = np.bincount(np.concatenate([u, v]), minlength=10000)
    assert abs(deg.mean() - 10) < 1e-9 and deg.max() < 40


def test_rmat_generator():
    u, v = d.gen_rmat(12, 16, seed=1)
    keys = u.astype(np.int64) << 32 | v
    assert (u < v).all()
    assert np.unique(keys).size == keys.size
    assert (np.diff(keys) > 0).all()
    u2, v2 = d.gen_rmat(12, 16, seed=1)
    assert (u == u2).all() and (v == v2).all()
    assert 0.5 * (16 << 12) < u.size < (16 << 12)
    deg = np.bincount(np.concatenate([u, v]), minlength=1 << 12)
    # skew: low IDs are hubs with a=0.57
    assert deg[0] > 20 * np.median(deg[deg > 0])


def test_batch_generator_semantics():
    n = 2000
    u, v = d.gen_er(n, 8000, 3)
    batches = d.gen_batches(n, u, v, 4, 1000, 0.5, 9)
    live = set(zip(np.minimum(u, v).tolist(), np.maximum(u, v).tolist()))
    for bu, bv, bi in batches:
        assert bi.sum() == 500
        ins = set()
        dels = set()
        for a, b, i in zip(bu.tolist(), bv.tolist(), bi.tolist()):
            k = (min(a, b), max(a, b))
            assert a != b
            (ins if i else dels).add(k)
        assert len(ins) == 500 and len(dels) == 500
        assert not (ins & live)          # insertions are non-edges
        assert dels <= live               # deletions are existing edges
        live = (live - dels) | ins
    # deterministic
    b2 = d.gen_batches(n, u, v, 4, 1000, 0.5, 9)
    assert all((x[0] == y[0]).all() and (x[2] == y[2]).all() for x, y in zip(batches, b2))


def test_header_documents_reference_citations():
    src = open(os.path.join(ROOT, "include", "dyncolor.h")).read()
    for sym in ("dc_graph_create", "dc_colors_set", "dc_apply_batch", "dc_recolor", "dc_validate"):
        block = src[: src.index(sym + "(")]
        last = block[block.rfind("/*"):]
        assert re.search(r"\.(cuh|cu|h):\d+", last), sym


def test_struct_layouts_match_header(tmp_path):
    # the ctypes mirrors of dc_options, dc_stats and dc_serve_stats against
    # the C header: size and every field offset
    structs = {"dc_options": d.Options, "dc_stats": d.Stats, "dc_serve_stats": d.ServeStats,
               "dc_tuning": d.Tuning, "dc_partition_info": d.PartitionInfo}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dyncolor.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", exe, str(src)],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, val = line.split()
        py = structs[cname]
        want = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)
