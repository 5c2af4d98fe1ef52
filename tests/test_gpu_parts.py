"""Arena parts (dc_set_parts): the loop checks each part on its own and lets a
settled part leave while the others go on.  Colours must equal those of the
same arena without parts, of each graph alone, and of the CPU oracle."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _graphs():
    out = []
    for i, s in enumerate((11, 9, 12, 10, 8, 11)):
        if i % 2 == 0:
            u, v = d.gen_rmat(s, 8, seed=70 + i)
            out.append((1 << s, u, v))
        else:
            n = 600 * (i + 1)
            u, v = d.gen_er(n, n * 5, 80 + i)
            out.append((n, u, v))
    return out


def test_parts_equal_whole_and_alone(oracle_mod):
    say("parts")
    graphs = _graphs()
    n, U, V, off = d.union_edges(graphs)
    with_parts = d.Graph(n, U, V)
    with_parts.static_color(0)
    with_parts.set_parts(off)
    whole = d.Graph(n, U, V)
    whole.static_color(0)
    og = oracle_mod.OracleGraph(n, U, V)
    og.static_color(0)
    alone = []
    for gn, gu, gv in graphs:
        g = d.Graph(gn, gu, gv)
        g.static_color(0)
        alone.append(g)
    per = [d.gen_batches(gn, gu, gv, 4, max(20, gu.size // 12), 0.5, 90 + i)
           for i, (gn, gu, gv) in enumerate(graphs)]
    left = 0
    for step in range(4):
        bu, bv, bi = d.union_batch([p[step] for p in per], off)
        st_p = with_parts.process_batch(bu, bv, bi, 0)
        st_w = whole.process_batch(bu, bv, bi, 0)
        og.apply_batch(bu, bv, bi)
        og.recolor(0)
        for g, p in zip(alone, per):
            g.process_batch(*p[step], 0)
        got = with_parts.colors()
        assert (got == whole.colors()).all(), f"step {step}: parts differ from the whole arena"
        assert (got == og.color).all(), f"step {step}: parts differ from the oracle"
        ref = np.concatenate([g.colors() for g in alone])
        assert (got == ref).all(), f"step {step}: parts differ from the graphs alone"
        assert with_parts.validate()[0] == 0
        assert with_parts.check_consistency() == 0
        tr = with_parts.trace()
        left += int(tr[tr[:, 0] == 12][:, 2].sum())
        assert st_p.loop_rounds <= st_w.loop_rounds
    assert left > 0, "no part left the loop early"


def test_parts_cross_insert_and_errors():
    say("parts errors")
    graphs = _graphs()[:3]
    n, U, V, off = d.union_edges(graphs)
    g = d.Graph(n, U, V)
    g.static_color(0)
    bad = off.copy()
    bad[-1] += 1
    with pytest.raises(d.DynColorError):
        g.set_parts(bad)
    with pytest.raises(d.DynColorError):
        g.set_parts(np.array([0, 5, 5, n], np.int32))
    # an edge across parts cannot be declared away
    cross = d.Graph(n, np.append(U, np.int32(0)), np.append(V, np.int32(off[1])))
    with pytest.raises(d.DynColorError):
        cross.set_parts(off)
    # a batch inserting one: the parts end, colours still those of the arena
    g.set_parts(off)
    plain = d.Graph(n, U, V)
    plain.static_color(0)
    bu = np.array([0, 1], np.int32)
    bv = np.array([int(off[1]), int(off[2]) + 1], np.int32)
    bi = np.array([1, 1], np.uint8)
    g.process_batch(bu, bv, bi, 0)
    plain.process_batch(bu, bv, bi, 0)
    assert (g.colors() == plain.colors()).all()
    assert g.check_consistency() == 0
    g.set_parts(np.zeros(0, np.int32))
