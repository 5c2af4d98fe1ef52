"""The graph driver (round blocks captured once as HIP graphs, then replayed)
and the loop's tuning parameters: colours and traces equal the oracle's for
every driver and every tuning, since neither changes what a round computes."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu


def _stream(scale=15, nb=5, bs=6000, seed=21):
    u, v = d.gen_rmat(scale, 16, seed=3)
    n = 1 << scale
    return n, u, v, d.gen_batches(n, u, v, nb, bs, 0.5, seed)


@pytest.mark.parametrize("exact", [0, 1])
def test_graph_driver_parity(oracle_mod, exact):
    n, u, v, batches = _stream()
    say("graph driver", exact)
    g, _, st = run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(exact_rounds=exact),
                        driver=d.DRIVER_GRAPH, device_batches=True)
    # blocks are captured on first use and replayed after that
    assert st[0].graph_captures > 0
    assert sum(s.graph_captures for s in st[2:]) < sum(s.loop_rounds for s in st[2:]) / 8


def test_graph_driver_partition_and_ldf(oracle_mod):
    from partition_shape import rmat_partition
    p = rmat_partition(13, 16, 5, 4, 2, 4000)
    say("graph driver partition")
    run_both(oracle_mod, p["n"], p["u"], p["v"], p["batches"], 0, colors=p["init"],
             n_owned=p["n_own"], ghost_colors=p["init"][p["n_own"]:], driver=d.DRIVER_GRAPH)
    n, u, v, batches = _stream(12, 2, 1500)
    run_both(oracle_mod, n, u, v, batches, 1, driver=d.DRIVER_GRAPH)


def test_graph_driver_with_parts(oracle_mod):
    # part checks run inside the captured blocks
    say("graph driver parts")
    graphs = []
    for i, s in enumerate((11, 9, 12, 10)):
        if i % 2 == 0:
            u, v = d.gen_rmat(s, 8, seed=70 + i)
            graphs.append((1 << s, u, v))
        else:
            n = 600 * (i + 1)
            u, v = d.gen_er(n, n * 5, 80 + i)
            graphs.append((n, u, v))
    n, U, V, off = d.union_edges(graphs)
    g = d.Graph(n, U, V)
    g.set_driver(d.DRIVER_GRAPH)
    g.static_color(0)
    g.set_parts(off)
    og = oracle_mod.OracleGraph(n, U, V)
    og.static_color(0)
    per = [d.gen_batches(gn, gu, gv, 4, max(20, gu.size // 12), 0.5, 90 + i)
           for i, (gn, gu, gv) in enumerate(graphs)]
    for step in range(4):
        bu, bv, bi = d.union_batch([p[step] for p in per], off)
        g.process_batch(bu, bv, bi, 0)
        og.apply_batch(bu, bv, bi)
        og.recolor(0)
        assert (g.colors() == og.color).all(), f"step {step}"
        assert g.check_consistency() == 0


@pytest.mark.parametrize("kw", [dict(loop_blocks=64), dict(loop_blocks=4096, light_blocks=7),
                                dict(light_c=0), dict(light_c=1 << 30),
                                dict(list_c=0, list_slots=0), dict(list_c=1 << 30, list_slots=1 << 40),
                                dict(heavy_div=1), dict(heavy_div=4096)])
def test_tunings_give_the_same_rounds(oracle_mod, kw):
    n, u, v, batches = _stream(14, 3, 3000)
    say("tuning", kw)
    g = d.Graph(n, u, v, d.options(exact_rounds=1))
    g.set_tuning(**kw)
    got = g.get_tuning().as_dict()
    assert all(got[k] == v for k, v in kw.items())
    g.static_color(0)
    og = oracle_mod.OracleGraph(n, u, v)
    og.static_color(0)
    from traces import same_trace
    for b in batches:
        g.process_batch(*b, 0)
        og.apply_batch(*b)
        tr = og.recolor(0)
        assert (g.colors() == og.color).all()
        assert same_trace(g.trace(), tr)


def test_set_tuning_rejects_out_of_range():
    u, v = d.gen_er(100, 300, 1)
    g = d.Graph(100, u, v)
    for kw in (dict(loop_blocks=0), dict(loop_blocks=5000), dict(heavy_div=0), dict(light_c=-1)):
        with pytest.raises(d.DynColorError) as e:
            g.set_tuning(**kw)
        assert e.value.code == d.ERR_INVALID
    assert g.get_tuning().loop_blocks == 2048


def test_tuner_sweep_keeps_results(oracle_mod):
    # a small sweep on a live graph: it returns a choice from the candidates
    # and every trial's batch + inverse leaves a valid colouring; the tuned
    # graph then colours the stream exactly as the oracle does
    from dyncolor_amd import tune
    say("tuner sweep")
    n, u, v, batches = _stream(13, 3, 2000)
    g = d.Graph(n, u, v)
    g.static_color(0)
    trial = d.gen_batches(n, u, v, 1, 1500, 0.5, 99)
    best, sweep = tune.tune(g, trial, 0, reps=1, params=["light_c", "list_c"])
    assert best["light_c"] in tune.CANDIDATES["light_c"]
    assert best["list_c"] in tune.CANDIDATES["list_c"]
    # each candidate is tried at most once, and every one but the incumbent
    # of the moment (a taken candidate makes the old default a candidate)
    for name in ("light_c", "list_c"):
        tried = [e["value"] for e in sweep if e["param"] == name]
        assert len(tried) == len(set(tried))
        assert set(tried) <= set(tune.CANDIDATES[name])
        assert len(tune.CANDIDATES[name]) - 1 <= len(tried) <= len(tune.CANDIDATES[name])
    assert {e["param"] for e in sweep} <= {"light_c", "list_c"}
    assert g.validate()[0] == 0 and g.check_consistency() == 0
    params = {k: val for k, val in best.items() if k != "driver"}
    run_both(oracle_mod, n, u, v, batches, 0, tuning=params)


def test_lane_tuner_keeps_results(oracle_mod):
    # the tuner over two lanes stepped together: a choice from the candidates,
    # set on both arenas, and the lanes still colour exactly as the oracle
    from dyncolor_amd import tune
    say("lane tuner")
    graphs = [(1 << 12, *d.gen_rmat(12, 8, seed=31)), (3000, *d.gen_er(3000, 15000, 32))]
    arenas, oracles, batches = [], [], []
    for i, (n, u, v) in enumerate(graphs):
        a = d.Graph(n, u, v)
        a.static_color(0)
        o = oracle_mod.OracleGraph(n, u, v)
        o.static_color(0)
        arenas.append(a)
        oracles.append(o)
        batches.append(d.gen_batches(n, u, v, 3, 400, 0.5, 40 + i))
    best, sweep = tune.tune_lanes(arenas, [b[:2] for b in batches], 0, reps=1, params=["light_c"])
    assert best["light_c"] in tune.CANDIDATES["light_c"]
    assert all(a.get_tuning().light_c == best["light_c"] for a in arenas)
    assert {e["param"] for e in sweep} == {"light_c"}
    for a in arenas:
        assert a.validate()[0] == 0 and a.check_consistency() == 0
    # from the tuned state (each trial's batch was undone by its inverse, so
    # the edges are the original ones): one more lanes step against the oracle
    for a, o in zip(arenas, oracles):
        o.set_colors(a.colors())
    ups = [a.upload_batch(*b[2]) for a, b in zip(arenas, batches)]
    d.lanes_step(arenas, ups, 0)
    for a, o, b in zip(arenas, oracles, batches):
        o.apply_batch(*b[2])
        o.recolor(0)
        assert (a.colors() == o.color).all()
        assert a.check_consistency() == 0
