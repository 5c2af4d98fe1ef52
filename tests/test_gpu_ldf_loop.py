"""The heuristic-based (ldf) loop in both forms -- counts kept incrementally
(dc_tuning.ldf_loop = 0, the default that every other ldf test runs) and
C re-derived from L's rows each round (ldf_loop = 1, the form of
HB/Color_main.cu:203-224) -- against the oracle on the same inputs: colours
and per-round traces (|L|, |C|, changed) equal, slot colour caches and
histograms exact afterwards."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu


def _set_ldf_loop(g, form):
    t = g.get_tuning()
    t.ldf_loop = form
    g.set_tuning(t)


def _run(oracle_mod, n, u, v, batches, form, colors=None, opts=None, device_batches=False):
    return run_both(oracle_mod, n, u, v, batches, 1, colors=colors, opts=opts,
                    device_batches=device_batches, tuning={"ldf_loop": form})


@pytest.mark.parametrize("form", [0, 1])
def test_ldf_er(oracle_mod, form):
    say("ldf ER", form)
    u, v = d.gen_er(20000, 100000, 4)
    batches = d.gen_batches(20000, u, v, 4, 4000, 0.5, 9)
    g, _, st = _run(oracle_mod, 20000, u, v, batches, form)
    assert g.get_tuning().ldf_loop == form
    assert g.validate()[0] == 0
    assert sum(s.loop_rounds for s in st) > 0


@pytest.mark.parametrize("form", [0, 1])
def test_ldf_rmat_stream(oracle_mod, form):
    # skewed degrees: hub rows (histogram rows, the wave path) among the
    # members every round
    say("ldf rmat", form)
    u, v = d.gen_rmat(15, 16, seed=3)
    n = 1 << 15
    batches = d.gen_batches(n, u, v, 3, 3000, 0.5, 21)
    g, _, _ = _run(oracle_mod, n, u, v, batches, form, device_batches=True)
    assert g.validate()[0] == 0
    assert g.check_consistency() == 0


def test_ldf_insert_only_and_delete_only(oracle_mod):
    say("ldf one-sided batches")
    u, v = d.gen_er(5000, 30000, 6)
    rng = np.random.default_rng(6)
    a = rng.integers(0, 5000, 3000, dtype=np.int32)
    b = rng.integers(0, 5000, 3000, dtype=np.int32)
    keep = a != b
    ins = (a[keep], b[keep], np.ones(keep.sum(), np.uint8))
    dele = (u[:2000], v[:2000], np.zeros(2000, np.uint8))
    _run(oracle_mod, 5000, u, v, [ins, dele, ins], 0)


def test_ldf_arbitrary_start_colours(oracle_mod):
    # start from colours that are far from a greedy colouring (many equal
    # neighbours): long loops, many colour changes per round
    say("ldf arbitrary colours")
    u, v = d.gen_er(4000, 40000, 12)
    rng = np.random.default_rng(12)
    colors = rng.integers(0, 3, 4000).astype(np.int32)
    batches = d.gen_batches(4000, u, v, 2, 500, 0.5, 13)
    _, _, st = _run(oracle_mod, 4000, u, v, batches, 0, colors=colors)
    assert max(s.loop_rounds for s in st) >= 3


def test_ldf_forms_agree_at_scale():
    # R-MAT scale 18: both forms on copies of one graph, same colours and
    # traces batch after batch (the oracle covers the smaller cases)
    say("ldf forms at scale")
    u, v = d.gen_rmat(18, 16, seed=5)
    n = 1 << 18
    batches = d.gen_batches(n, u, v, 3, 20000, 0.5, 17)
    gs = []
    for form in (0, 1):
        g = d.Graph(n, u, v)
        _set_ldf_loop(g, form)
        g.static_color(1)
        gs.append(g)
    assert (gs[0].colors() == gs[1].colors()).all()
    for bu, bv, bi in batches:
        tr = []
        for g in gs:
            g.process_batch(bu, bv, bi, 1)
            tr.append(g.trace())
        assert (gs[0].colors() == gs[1].colors()).all()
        assert np.array_equal(np.asarray(tr[0]), np.asarray(tr[1]))
    for g in gs:
        assert g.validate()[0] == 0
        assert g.check_consistency() == 0


def test_ldf_max_rounds():
    # the loop stops with DC_ERR_ROUNDS when it needs more rounds than allowed
    say("ldf max rounds")
    u, v = d.gen_er(4000, 40000, 12)
    rng = np.random.default_rng(12)
    g = d.Graph(4000, u, v, d.options(max_rounds=1))
    g.set_colors(np.zeros(4000, np.int32))
    a = rng.integers(0, 4000, 200, dtype=np.int32)
    b = (a + 1 + rng.integers(0, 3998, 200, dtype=np.int32)) % 4000
    with pytest.raises(d.DynColorError) as e:
        g.process_batch(a, b.astype(np.int32), np.ones(200, np.uint8), 1)
    assert e.value.code == d.ERR_ROUNDS


def test_ldf_tuning_range():
    u, v = d.gen_er(100, 300, 1)
    g = d.Graph(100, u, v)
    t = g.get_tuning()
    assert t.ldf_loop == 0
    t.ldf_loop = 2
    with pytest.raises(d.DynColorError):
        g.set_tuning(t)


def test_rules_alternate_on_one_graph(oracle_mod):
    # the heuristic-based loop leaves the histograms, the all-neighbour mex
    # caches and the slot colours of predecessors' rows to one pass after the
    # loop; an independent-rule batch next reads all of them
    say("ldf / independent alternating")
    from traces import same_trace
    u, v = d.gen_rmat(14, 16, seed=9)
    n = 1 << 14
    batches = d.gen_batches(n, u, v, 6, 2000, 0.5, 31)
    g = d.Graph(n, u, v)
    og = oracle_mod.OracleGraph(n, u, v)
    g.static_color(1)
    og.static_color(1)
    for i, (bu, bv, bi) in enumerate(batches):
        rule = 1 if i % 2 == 0 else 0
        g.process_batch(bu, bv, bi, rule)
        og.apply_batch(bu, bv, bi)
        ph = og.recolor(rule)
        assert (g.colors() == og.color).all(), f"batch {i} rule {rule}"
        assert same_trace(g.trace(), ph), f"batch {i} rule {rule}: trace"
        assert g.check_consistency() == 0
    assert g.validate()[0] == 0


def test_ldf_slot_colours_both_ways(oracle_mod):
    # after the loop the neighbours' slot colours, histograms and all-neighbour
    # mexes take the final colours: by the moved vertices' rows when few moved
    # (small batches on a greedy colouring), by a rewrite of every row when
    # many did (from colours far from a greedy colouring most vertices move);
    # both leave them exact, and an independent-rule batch next reads them
    say("ldf slot colours both ways")
    from traces import same_trace
    u, v = d.gen_rmat(17, 16, seed=8)
    n = 1 << 17
    g = d.Graph(n, u, v)
    og = oracle_mod.OracleGraph(n, u, v)
    g.static_color(1)
    og.static_color(1)
    rng = np.random.default_rng(8)
    start = rng.integers(0, 4, n).astype(np.int32)
    steps = [(d.gen_batches(n, u, v, 1, k, 0.5, 40 + k)[0], 1, None) for k in (4, 16, 64, 256)]
    steps += [(d.gen_batches(n, u, v, 1, 1500, 0.5, 33)[0], 1, start),
              (d.gen_batches(n, u, v, 1, 1500, 0.5, 35)[0], 0, None)]
    ways = []
    for i, ((bu, bv, bi), rule, colors) in enumerate(steps):
        if colors is not None:
            g.set_colors(colors)
            og.set_colors(colors)
        st = g.process_batch(bu, bv, bi, rule)
        og.apply_batch(bu, bv, bi)
        ph = og.recolor(rule)
        assert (g.colors() == og.color).all(), f"batch {i} rule {rule}"
        assert same_trace(g.trace(), ph), f"batch {i} rule {rule}: trace"
        assert g.check_consistency() == 0, f"batch {i}"
        if rule == 1:
            ways.append((st.slots_rewritten, st.moved_slots))
    say("ways", ways)
    assert ways[-1][0] == 1
    assert any(w == 0 and m > 0 for w, m in ways[:-1]), ways
