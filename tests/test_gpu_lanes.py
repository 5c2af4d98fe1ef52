"""Batched steps over concurrent lanes (dc_lanes_step): graphs split over a
few arenas, each lane stepped on its own host thread and stream.  Every
graph's colours equal those it gets alone (the rules read only a vertex's own
neighbourhood) and those of the oracle."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _graphs():
    gs = []
    for i in range(9):
        if i % 2 == 0:
            n = 600 + 400 * i
            u, v = d.gen_er(n, 5 * n, 40 + i)
        else:
            scale = 9 + i // 3
            u, v = d.gen_rmat(scale, 8, seed=40 + i)
            n = 1 << scale
        gs.append((n, u, v))
    return gs


def test_lanes_equal_alone_and_oracle(oracle_mod):
    say("lanes step")
    gs = _graphs()
    steps = 2
    per = [d.gen_batches(n, u, v, steps, max(4, u.size // 20), 0.5, 70 + i)
           for i, (n, u, v) in enumerate(gs)]
    lane_of = [i % 3 for i in range(len(gs))]
    lanes = []
    for k in range(3):
        idx = [i for i in range(len(gs)) if lane_of[i] == k]
        n, U, V, off = d.union_edges([gs[i] for i in idx])
        a = d.Graph(n, U, V)
        a.static_color(0)
        ub = [a.upload_batch(*d.union_batch([per[i][s] for i in idx], off)) for s in range(steps)]
        lanes.append((idx, off, a, ub))
    for s in range(steps):
        st = d.lanes_step([l[2] for l in lanes], [l[3][s] for l in lanes], 0)
        assert len(st) == 3 and all(x.updates == l[3][s].k for x, l in zip(st, lanes))
    for idx, off, a, _ in lanes:
        assert a.validate()[0] == 0
        assert a.check_consistency() == 0
        col = a.colors()
        for j, i in enumerate(idx):
            n, u, v = gs[i]
            g = d.Graph(n, u, v)
            g.static_color(0)
            og = oracle_mod.OracleGraph(n, u, v)
            og.static_color(0)
            for s in range(steps):
                g.process_batch(*per[i][s], 0)
                og.apply_batch(*per[i][s])
                og.recolor(0)
            assert (col[off[j]:off[j + 1]] == g.colors()).all(), f"graph {i}: lane differs from alone"
            assert (g.colors() == og.color).all(), f"graph {i}: differs from the oracle"


def test_lanes_step_errors():
    say("lanes step errors")
    u, v = d.gen_er(300, 900, 3)
    a = d.Graph(300, u, v)
    ub_, vb_ = d.gen_er(200, 600, 4)
    b = d.Graph(200, ub_, vb_)
    a.static_color(0)
    b.static_color(0)
    ba = a.upload_batch(np.array([0], np.int32), np.array([1], np.int32), np.array([1], np.uint8))
    bb = b.upload_batch(np.array([2], np.int32), np.array([3], np.int32), np.array([1], np.uint8))
    # lane 1 gets a batch validated for a graph of another size
    with pytest.raises(d.DynColorError) as e:
        d.lanes_step([a, b], [ba, ba], 0)
    assert e.value.code == d.ERR_INVALID
    assert "lane 1" in str(e.value)
    # the same arena twice
    with pytest.raises(d.DynColorError) as e:
        d.lanes_step([a, a], [ba, ba], 0)
    assert e.value.code == d.ERR_INVALID
    # both handles still work
    d.lanes_step([a, b], [ba, bb], 0)
    assert a.validate()[0] == 0 and b.validate()[0] == 0
