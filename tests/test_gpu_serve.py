"""Many graphs at once: the arena (disjoint union of resident graphs) recolored
in one launch gives every graph exactly the colors it gets alone, both for a
batched step (one batch per graph) and under the serving loop, where each
graph receives the coalesced batches the loop formed for it."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _graphs():
    out = []
    for i, s in enumerate((9, 10, 11, 10, 8)):
        if i % 2 == 0:
            u, v = d.gen_rmat(s, 8, seed=40 + i)
            out.append((1 << s, u, v))
        else:
            n = 700 * (i + 1)
            u, v = d.gen_er(n, n * 6, 50 + i)
            out.append((n, u, v))
    return out


@pytest.mark.parametrize("driver", [d.DRIVER_PLAIN, d.DRIVER_FUSED])
def test_union_step_equals_alone(oracle_mod, driver):
    say("union step", driver)
    graphs = _graphs()
    n, U, V, off = d.union_edges(graphs)
    arena = d.Graph(n, U, V)
    arena.set_driver(driver)
    arena.static_color(0)
    per = [d.gen_batches(gn, gu, gv, 2, max(20, gu.size // 15), 0.5, 60 + i)
           for i, (gn, gu, gv) in enumerate(graphs)]
    for step in range(2):
        bu, bv, bi = d.union_batch([p[step] for p in per], off)
        arena.process_batch(bu, bv, bi, 0)
    got = arena.colors()
    assert arena.validate()[0] == 0
    for i, (gn, gu, gv) in enumerate(graphs):
        og = oracle_mod.OracleGraph(gn, gu, gv)
        og.static_color(0)
        for step in range(2):
            og.apply_batch(*per[i][step])
            og.recolor(0)
        assert (got[off[i]:off[i + 1]] == og.color).all(), f"graph {i}"


def _requests(graphs, nreq, rate, seed):
    rng = np.random.default_rng(seed)
    gid = rng.integers(0, len(graphs), nreq).astype(np.int32)
    sizes = np.clip(np.exp(rng.normal(np.log(40), 1.0, nreq)), 1, 400).astype(np.int64)
    roff = np.zeros(nreq + 1, np.int64)
    roff[1:] = np.cumsum(sizes)
    L = int(roff[-1])
    u = np.zeros(L, np.int32)
    v = np.zeros(L, np.int32)
    ins = np.zeros(L, np.uint8)
    for r in range(nreq):
        gn, gu, gv = graphs[gid[r]]
        for j in range(roff[r], roff[r + 1]):
            if rng.random() < 0.6:
                a, b = rng.choice(gn, 2, replace=False)
                u[j], v[j], ins[j] = a, b, 1
            else:
                e = rng.integers(0, gu.size)
                u[j], v[j], ins[j] = gu[e], gv[e], 0
    arrival = np.cumsum(rng.exponential(1.0 / rate, nreq))
    return gid, roff, u, v, ins, arrival


def _lanes(graphs, nlanes, driver=d.DRIVER_PLAIN):
    """Graph g goes to lane g % nlanes; returns (arenas, lane_of, first, size)."""
    arenas, lane_of, first, size = [], [], [], []
    place = {}
    for l in range(nlanes):
        mine = [g for g in range(len(graphs)) if g % nlanes == l]
        n, U, V, off = d.union_edges([graphs[g] for g in mine])
        a = d.Graph(n, U, V)
        a.set_driver(driver)
        a.static_color(0)
        arenas.append(a)
        for k, g in enumerate(mine):
            place[g] = (l, int(off[k]), int(off[k + 1] - off[k]))
    for g in range(len(graphs)):
        lane_of.append(place[g][0])
        first.append(place[g][1])
        size.append(place[g][2])
    return arenas, lane_of, first, size


def _check_alone(oracle_mod, graphs, arenas, lane_of, first, size, req, step):
    # every graph's colours equal those of its coalesced batches replayed
    # alone on the oracle, in the order its lane ran them
    gid, roff, u, v, ins, _ = req
    for g, (gn, gu, gv) in enumerate(graphs):
        og = oracle_mod.OracleGraph(gn, gu, gv)
        og.static_color(0)
        mine = np.nonzero(gid == g)[0]
        for s_ in sorted(set(step[mine].tolist())):
            rs = [r for r in mine if step[r] == s_]
            idx = np.concatenate([np.arange(roff[r], roff[r + 1]) for r in rs])
            og.apply_batch(u[idx], v[idx], ins[idx])
            og.recolor(0)
        got = arenas[lane_of[g]].colors()[first[g]:first[g] + size[g]]
        assert (got == og.color).all(), f"graph {g}"


@pytest.mark.parametrize("nlanes,driver", [(1, d.DRIVER_FUSED), (2, d.DRIVER_PLAIN),
                                           (3, d.DRIVER_PLAIN)])
def test_serving_isolation(oracle_mod, nlanes, driver):
    # requests arrive faster than steps finish: they are admitted while
    # steps run and coalesced; each graph ends with the colours of its
    # coalesced batches applied alone
    say("serving isolation", nlanes, driver)
    graphs = _graphs()
    arenas, lane_of, first, size = _lanes(graphs, nlanes, driver)
    req = _requests(graphs, 120, 3000.0, 5)
    budget = d.request_bytes(3000, 3000)
    srv = d.Server(arenas, lane_of, first, size, budget_bytes=budget)
    lat, step, st = srv.run(*req)
    assert st.served == req[0].size and st.rejected == 0 and st.lanes == nlanes
    assert (lat >= 0).all()
    assert st.steps >= 2 and st.max_inflight_bytes <= budget
    assert any(np.unique(step[req[0] == g]).size < (req[0] == g).sum() for g in range(len(graphs))), \
        "no coalescing happened"
    for a in arenas:
        assert a.validate()[0] == 0
    _check_alone(oracle_mod, graphs, arenas, lane_of, first, size, req, step)


def test_serving_admission_waits_for_memory(oracle_mod):
    # a budget of about two requests: later arrivals wait until a step frees
    # its bytes; results stay those of the graphs alone
    say("serving admission waits")
    graphs = _graphs()
    arenas, lane_of, first, size = _lanes(graphs, 2)
    req = _requests(graphs, 60, 20000.0, 9)
    budget = d.request_bytes(800, 800)
    srv = d.Server(arenas, lane_of, first, size, budget_bytes=budget)
    lat, step, st = srv.run(*req)
    assert st.served == req[0].size and st.rejected == 0
    assert st.admit_waits > 0 and st.max_inflight_bytes <= budget
    _check_alone(oracle_mod, graphs, arenas, lane_of, first, size, req, step)


def test_serving_rejects_oversized_request():
    say("serving admission")
    graphs = _graphs()[:2]
    n, U, V, off = d.union_edges(graphs)
    arena = d.Graph(n, U, V)
    arena.static_color(0)
    srv = d.Server.one_lane(arena, off, budget_bytes=d.request_bytes(10, 10))
    gid = np.array([0, 1], np.int32)
    roff = np.array([0, 5, 25], np.int64)
    u = np.array([0, 1, 2, 3, 4] + list(range(20)), np.int32)
    v = np.array([5, 6, 7, 8, 9] + list(range(30, 50)), np.int32)
    ins = np.ones(25, np.uint8)
    lat, step, st = srv.run(gid, roff, u, v, ins, np.array([0.0, 0.0]))
    assert st.served == 1 and st.rejected == 1
    assert lat[1] == -1.0 and lat[0] >= 0


def test_server_argument_checks():
    graphs = _graphs()[:2]
    n, U, V, off = d.union_edges(graphs)
    arena = d.Graph(n, U, V)
    with pytest.raises(d.DynColorError):
        d.Server([arena], [0, 0], [0, int(off[1])], [int(off[1]) + 5, 10])   # past the arena
    with pytest.raises(d.DynColorError):
        d.Server([arena], [0, 0], [0, 2], [int(off[1]), 10])                # overlapping graphs
    srv = d.Server.one_lane(arena, off, budget_bytes=1 << 20)
    # a non-empty request with null line arrays
    import ctypes as C
    lat = np.zeros(1)
    st = d.ServeStats()
    rc = d.lib().dc_server_run(srv._h, 1, np.zeros(1, np.int32), np.array([0, 3], np.int64),
                               None, None, None, np.zeros(1), lat, np.zeros(1, np.int32), C.byref(st))
    assert rc == d.ERR_INVALID
