"""Pins for the CPU oracle (oracle/dc_oracle.c).  CPU only.

Each check is chosen so that a plausible mistake in the oracle (a wrong
tie-break, a snapshot read taken mid-round, a filter on the wrong colors, a
deletion that does not win over an insertion) fails at least one of them.
"""
import itertools
import random

import numpy as np
import pytest

import brute


def _rand_graph(rng, n, p):
    return [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]


# ------------------------------------------------------------------ mex ----
def test_mex_exhaustive_subsets(oracle_mod):
    # every subset of {0..11}: the mex is the least absent value
    for mask in range(1 << 12):
        s = [i for i in range(12) if mask >> i & 1]
        expect = next(i for i in range(13) if not (mask >> i & 1))
        # shuffled, with duplicates and out-of-window values mixed in
        vals = s + s[: len(s) // 2] + [100, 57]
        random.Random(mask).shuffle(vals)
        assert oracle_mod.mex(vals) == expect


def test_mex_special_cases(oracle_mod):
    assert oracle_mod.mex([]) == 0
    for k in range(1, 70):
        assert oracle_mod.mex(list(range(k))) == k
        assert oracle_mod.mex(list(range(1, k + 1))) == 0
    assert oracle_mod.mex([5] * 40) == 0


# ----------------------------------------------- reference sample (golden) --
@pytest.mark.parametrize("rule,key", [(0, "independent_hb_colors"), (1, "ldf_hb_colors")])
def test_sample_matches_hand_trace(oracle_mod, sample, rule, key):
    e = np.array(sample["edges"])
    ch = np.array(sample["changes"])
    g = oracle_mod.OracleGraph(sample["n"], e[:, 0], e[:, 1])
    g.set_colors(sample["colors_hb_dir"])
    g.apply_batch(ch[:, 0], ch[:, 1], ch[:, 2])
    ph = g.recolor(rule)
    exp = sample["expected"][key]
    assert g.color.tolist() == exp["colors"]
    assert ph.tolist() == exp["trace"]
    assert g.validate()[0] == 0
    if "prop" in exp:
        assert g.prop().tolist() == exp["prop"]


def test_sample_input_coloring_is_improper(oracle_mod, sample):
    # the shipped coloring has edge 1-9 monochromatic (both color 1)
    e = np.array(sample["edges"])
    g = oracle_mod.OracleGraph(sample["n"], e[:, 0], e[:, 1])
    g.set_colors(sample["colors_hb_dir"])
    assert g.validate()[0] == 1


def test_sample_independent_multiword_colors(oracle_mod, sample):
    # the independent-rule directory's coloring gives vertex 4 color 32
    e = np.array(sample["edges"])
    ch = np.array(sample["changes"])
    g = oracle_mod.OracleGraph(sample["n"], e[:, 0], e[:, 1])
    g.set_colors(sample["colors_hi_dir"])
    g.apply_batch(ch[:, 0], ch[:, 1], ch[:, 2])
    ph = g.recolor(0)
    bg = brute.BruteGraph(sample["n"], [tuple(x) for x in sample["edges"]])
    colors, trace = brute.run_batch(bg, sample["colors_hi_dir"], [tuple(x) for x in sample["changes"]], 0)
    assert g.color.tolist() == colors
    assert ph.tolist() == trace
    assert g.validate()[0] == 0


# ---------------------------------------------------------- CSR model ------
def test_apply_matches_set_model(oracle_mod):
    rng = random.Random(3)
    for trial in range(30):
        n = rng.randint(4, 14)
        edges = _rand_graph(rng, n, 0.35)
        if not edges:
            continue
        live = {}
        for a, b in edges:
            k = (min(a, b), max(a, b))
            live[k] = live.get(k, 0) + 1
        lines = []
        for _ in range(rng.randint(0, 12)):
            a, b = rng.sample(range(n), 2)
            lines.append((a, b, rng.randint(0, 1)))
        for a, b, ins in lines:
            if ins:
                k = (min(a, b), max(a, b))
                live[k] = live.get(k, 0) + 1
        for a, b, ins in lines:   # deletion wins, all copies
            if not ins:
                live[(min(a, b), max(a, b))] = 0
        ea = np.array(edges)
        g = oracle_mod.OracleGraph(n, ea[:, 0], ea[:, 1])
        g.set_colors(np.zeros(n, np.int32))
        la = np.array(lines if lines else [(0, 1, 1)])[: len(lines)]
        g.apply_batch(la[:, 0] if len(lines) else [], la[:, 1] if len(lines) else [],
                      la[:, 2] if len(lines) else [])
        for v in range(n):
            got = sorted(g.neighbors(v).tolist())
            exp = []
            for (a, b), c in live.items():
                if a == v:
                    exp += [b] * c
                elif b == v:
                    exp += [a] * c
            assert got == sorted(exp), (trial, v)
        # live degree = old + inserted - deleted copies, and symmetric
        tot = sum(len(g.neighbors(v)) for v in range(n))
        assert tot == 2 * sum(live.values())


def test_ldf_counter_counts_lines(oracle_mod):
    # duplicate deletion lines decrement more than once
    g = oracle_mod.OracleGraph(4, [0, 1, 2], [1, 2, 3])
    g.set_colors([0, 1, 0, 1])
    g.apply_batch([1, 1, 0], [2, 2, 3], [0, 0, 1])
    assert g.prop().tolist() == [1 + 1, 2 - 2, 2 - 2, 1 + 1]


# ---------------------------------------------------------- static greedy --
def test_static_greedy_is_textbook(oracle_mod):
    rng = random.Random(5)
    for trial in range(40):
        n = rng.randint(2, 30)
        edges = _rand_graph(rng, n, rng.choice([0.1, 0.3, 0.6]))
        if not edges:
            continue
        ea = np.array(edges)
        for rule in (0, 1):
            g = oracle_mod.OracleGraph(n, ea[:, 0], ea[:, 1])
            got = g.static_color(rule).tolist()
            bg = brute.BruteGraph(n, edges)
            assert got == brute.greedy(bg, rule)
            assert g.validate()[0] == 0
            # fixed point: every vertex is the mex of its predecessors
            prop = g.prop()
            for v in range(n):
                if rule == 0:
                    pre = [got[w] for w in g.neighbors(v) if w < v]
                else:
                    pre = [got[w] for w in g.neighbors(v)
                           if (prop[w], -w) > (prop[v], -v)]
                assert got[v] == brute.mex(pre)


# ----------------------------------------------- oracle vs brute simulator --
def _compare(oracle_mod, n, edges, colors, lines, rule, n_owned=None):
    ea = np.array(edges) if edges else np.zeros((0, 2), int)
    g = oracle_mod.OracleGraph(n, ea[:, 0], ea[:, 1], n_owned=n_owned)
    g.set_colors(colors)
    la = np.array(lines) if lines else np.zeros((0, 3), int)
    g.apply_batch(la[:, 0], la[:, 1], la[:, 2])
    ph = g.recolor(rule)
    bg = brute.BruteGraph(n, edges, n_owned=n_owned)
    bc, bt = brute.run_batch(bg, colors, lines, rule)
    assert g.color.tolist() == bc
    assert ph.tolist() == bt
    return g, bg, bc


def test_all_small_batches_exhaustive(oracle_mod):
    # every batch of 1 or 2 lines on a few 6-vertex graphs, both rules
    rng = random.Random(11)
    n = 6
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
    lines_all = [(a, b, i) for a, b in pairs for i in (0, 1)]
    for gi in range(3):
        edges = _rand_graph(rng, n, 0.5)
        bg0 = brute.BruteGraph(n, edges)
        for rule in (0, 1):
            colors = brute.greedy(bg0, rule)
            batches = [[l] for l in lines_all]
            batches += [[l1, l2] for l1, l2 in itertools.product(lines_all, lines_all)
                        if rng.random() < 0.15]
            for lines in batches:
                _, bg, bc = _compare(oracle_mod, n, edges, colors, lines, rule)
                assert brute.monochromatic(bg, bc) == 0


def test_random_batches_vs_brute(oracle_mod):
    rng = random.Random(21)
    for trial in range(150):
        n = rng.randint(3, 12)
        edges = _rand_graph(rng, n, rng.choice([0.2, 0.4, 0.7]))
        rule = trial % 2
        bg0 = brute.BruteGraph(n, edges)
        if trial % 5 == 4:
            colors = [rng.randint(0, 3) for _ in range(n)]   # improper input allowed
        else:
            colors = brute.greedy(bg0, rule)
        lines = []
        for _ in range(rng.randint(0, 8)):
            a, b = rng.sample(range(n), 2)
            lines.append((a, b, rng.randint(0, 1)))
        _, bg, bc = _compare(oracle_mod, n, edges, colors, lines, rule)
        if trial % 5 != 4:
            assert brute.monochromatic(bg, bc) == 0


def test_partition_share_vs_brute(oracle_mod):
    rng = random.Random(31)
    for trial in range(80):
        n = rng.randint(5, 12)
        n_owned = rng.randint(2, n - 1)
        edges = [(a, b) for a, b in _rand_graph(rng, n, 0.45) if a < n_owned or b < n_owned]
        bg0 = brute.BruteGraph(n, edges, n_owned=n_owned)
        ghost = [rng.randint(0, 4) for _ in range(n - n_owned)]
        colors = brute.greedy(bg0, 0, [0] * n_owned + ghost)
        lines = []
        for _ in range(rng.randint(0, 8)):
            a, b = rng.sample(range(n), 2)
            if a >= n_owned and b >= n_owned:
                continue
            lines.append((a, b, rng.randint(0, 1)))
        _, bg, bc = _compare(oracle_mod, n, edges, colors, lines, 0, n_owned=n_owned)
        assert brute.monochromatic(bg, bc) == 0
        assert bc[n_owned:] == ghost


def test_independent_rule_converges_on_er_c1(oracle_mod):
    # the C1 workload: ER(10k, 50k), greedy start, 1k ins + 1k del; both rules
    # terminate far below the round cap and leave no monochromatic edge
    import dyncolor_amd as d
    u, v = d.gen_er(10000, 50000, 1)
    (bu, bv, bi), = d.gen_batches(10000, u, v, 1, 2000, 0.5, 1)
    for rule in (0, 1):
        g = oracle_mod.OracleGraph(10000, u, v)
        g.static_color(rule)
        g.apply_batch(bu, bv, bi)
        ph = g.recolor(rule, max_loop=5000)
        assert g.validate()[0] == 0
        assert len(ph) < 500


def test_partition_sample_fixture(oracle_mod):
    # hand-derived partition case (tests/golden/partition_sample.json): both
    # cross kinds, a cross deletion skipped for a ghost with the larger
    # colour, a cross insertion recolouring the owned endpoint, three loop
    # rounds with frozen ghost colours in every mex
    import json
    import os
    p = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "partition_sample.json")))
    e = np.array(p["edges"], np.int32)
    ch = np.array(p["changes"], np.int32)
    n = p["n_owned"] + p["n_ghost"]
    og = oracle_mod.OracleGraph(n, e[:, 0], e[:, 1], n_owned=p["n_owned"])
    og.set_colors(p["colors"])
    og.apply_batch(ch[:, 0], ch[:, 1], ch[:, 2].astype(np.uint8))
    tr = og.recolor(0)
    assert og.color.tolist() == p["expected"]["colors"]
    assert tr.tolist() == p["expected"]["trace"]
    assert og.validate() == (p["expected"]["monochromatic"], p["expected"]["max_color"])


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_replay_equals_single_thread(oracle_mod, threads):
    # the all-cores CPU baseline: the same colours, prev values and traces as
    # the single-thread oracle, on multi-batch R-MAT streams and on
    # partition shares with frozen ghosts
    import dyncolor_amd as d
    u, v = d.gen_rmat(12, 16, seed=11)
    n = 1 << 12
    batches = d.gen_batches(n, u, v, 3, 1500, 0.5, 12)
    a = oracle_mod.OracleGraph(n, u, v)
    b = oracle_mod.OracleGraph(n, u, v)
    a.static_color(0)
    b.static_color(0)
    for bu, bv, bi in batches:
        a.apply_batch(bu, bv, bi)
        b.apply_batch(bu, bv, bi)
        ta = a.recolor(0)
        tb = b.recolor_par(0, threads)
        assert ta.tolist() == tb.tolist()
        assert (a.color == b.color).all() and (a.prev == b.prev).all()
    rng = random.Random(5 + threads)
    for trial in range(6):
        nn = rng.randint(50, 2000)
        no = rng.randint(10, nn - 5)
        uu, vv = d.gen_er(nn, 4 * nn, 60 + trial)
        keep = (uu < no) | (vv < no)
        uu, vv = uu[keep], vv[keep]
        init = np.array([rng.randint(0, 5) for _ in range(nn)], np.int32)
        lines = []
        while len(lines) < nn // 4:
            x, y = rng.sample(range(nn), 2)
            if x >= no and y >= no:
                continue
            lines.append((x, y, rng.randint(0, 1)))
        la = np.array(lines, np.int32)
        a = oracle_mod.OracleGraph(nn, uu, vv, n_owned=no)
        b = oracle_mod.OracleGraph(nn, uu, vv, n_owned=no)
        for g in (a, b):
            g.set_colors(init)
            g.apply_batch(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))
        ta = a.recolor(0)
        tb = b.recolor_par(0, threads)
        assert ta.tolist() == tb.tolist()
        assert (a.color == b.color).all()
