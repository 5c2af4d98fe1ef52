"""dc_compact_noalloc (the compaction with its slot map in the library's
static device array) against dc_compact on identical graphs: the same rows in
the same slot order, consistent caches and twin links, and the same colours
and traces on the next batch.  Two calls in a row, the second on a smaller
graph (the array keeps the first call's contents); a graph past the array's
size is rejected."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _pair(n, m, seed, nb, bs):
    u, v = d.gen_er(n, m, seed)
    batches = d.gen_batches(n, u, v, nb + 1, bs, 0.3, seed + 1)
    gs = []
    for _ in range(2):
        g = d.Graph(n, u, v)
        g.static_color(0)
        for b in batches[:nb]:
            g.process_batch(*b, 0)
        gs.append(g)
    return gs, batches[nb]


def _rows(g, n):
    return [g.neighbors(x).tolist() for x in range(n)]


@pytest.mark.parametrize("n,m,seed", [(20000, 120000, 3), (3000, 9000, 4)])
def test_noalloc_matches_compact(n, m, seed):
    # the larger graph first, then the smaller one on the same static array.
    # Two copies of a graph need not share a slot order (insertions take
    # their slots by atomics), so each compaction is checked against its own
    # graph: every row keeps its live slots in their order; the two copies
    # hold the same rows as multisets and colour the next batch alike.
    say("noalloc", n, m)
    (ga, gb), nxt = _pair(n, m, seed, 2, n // 4)
    before_a, before_b = _rows(ga, n), _rows(gb, n)
    ga.compact()
    gb.compact_noalloc()
    ra, rb = _rows(ga, n), _rows(gb, n)
    assert ra == before_a
    assert rb == before_b
    assert [sorted(r) for r in ra] == [sorted(r) for r in rb]
    assert ga.check_consistency() == 0 and gb.check_consistency() == 0
    sa = ga.process_batch(*nxt, 0)
    sb = gb.process_batch(*nxt, 0)
    assert (ga.colors() == gb.colors()).all()
    assert gb.trace().tolist() == ga.trace().tolist() and sa.loop_rounds == sb.loop_rounds
    assert gb.validate()[0] == 0


def test_noalloc_rejects_graphs_past_the_array():
    u, v = d.gen_er(1000, 3000, 5)
    g = d.Graph(1000, u, v, d.options(spare_slots=1 << 23))
    g.static_color(0)
    assert g.info()["capacity"] >= 1 << 23
    with pytest.raises(d.DynColorError) as e:
        g.compact_noalloc()
    assert e.value.code == d.ERR_INVALID
    # the graph is untouched and the ordinary compaction still works
    g.compact()
    assert g.check_consistency() == 0
