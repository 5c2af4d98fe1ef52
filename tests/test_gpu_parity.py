"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Colors and the per-phase trace must be bit-identical;
every result must have zero monochromatic edges."""
import random
import sys

import numpy as np
import pytest

import brute
import dyncolor_amd as d
from traces import same_trace

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


def run_both(oracle_mod, n, u, v, batches, rule, colors=None, opts=None, n_owned=None,
             ghost_colors=None, device_batches=False, driver=d.DRIVER_PLAIN, tuning=None):
    """Replay batches on GPU and oracle; assert identical colors + traces.
    tuning: dc_tuning fields to set on the graph before the first batch."""
    u = np.asarray(u, np.int32)
    v = np.asarray(v, np.int32)
    if n_owned is None:
        g = d.Graph(n, u, v, opts)
        og = oracle_mod.OracleGraph(n, u, v)
    else:
        g = d.Graph(n_owned, u, v, opts, n_ghost=n - n_owned, ghost_colors=ghost_colors)
        og = oracle_mod.OracleGraph(n, u, v, n_owned=n_owned)
    g.set_driver(driver)
    if tuning:
        t = g.get_tuning()
        for k, x in tuning.items():
            setattr(t, k, x)
        g.set_tuning(t)
    if colors is None:
        if n_owned is not None:
            init = np.zeros(n, np.int32)
            init[n_owned:] = ghost_colors
            g.set_colors(init)
            og.set_colors(init)
        g.static_color(rule)
        og.static_color(rule)
        assert (g.colors() == og.color).all(), "static coloring differs"
    else:
        g.set_colors(colors)
        og.set_colors(colors)
    stats = []
    for bi, (bu, bv, bins) in enumerate(batches):
        if device_batches:
            db = g.upload_batch(bu, bv, bins)
            g.apply_device_batch(db)
            st = g.recolor(rule)
        else:
            st = g.process_batch(bu, bv, bins, rule)
        og.apply_batch(bu, bv, bins)
        ph = og.recolor(rule)
        got = g.colors()
        tr = g.trace()
        if not (got == og.color).all():
            bad = np.nonzero(got != og.color)[0]
            raise AssertionError(f"batch {bi}: {bad.size} vertices differ, first {bad[:10]}")
        assert same_trace(tr, ph), f"batch {bi}: trace differs"
        # slot colour cache and reverse-slot links stay exact
        assert g.check_consistency() == 0, f"batch {bi}: stale slot colour / twin"
        stats.append(st)
    return g, og, stats


# ------------------------------------------------------------------ sample --
@pytest.mark.parametrize("rule,key", [(0, "independent_hb_colors"), (1, "ldf_hb_colors")])
def test_sample_graph(sample, rule, key):
    say("sample", rule)
    e = np.array(sample["edges"], np.int32)
    ch = np.array(sample["changes"], np.int32)
    g = d.Graph(sample["n"], e[:, 0], e[:, 1])
    g.set_colors(sample["colors_hb_dir"])
    g.process_batch(ch[:, 0], ch[:, 1], ch[:, 2].astype(np.uint8), rule)
    exp = sample["expected"][key]
    assert g.colors().tolist() == exp["colors"]
    assert same_trace(g.trace(), exp["trace"])
    assert g.validate() == (0, max(exp["colors"]))
    if "prop" in exp:
        assert g.prop().tolist() == exp["prop"]


def test_sample_graph_color32(oracle_mod, sample):
    say("sample color 32")
    e = np.array(sample["edges"], np.int32)
    ch = np.array(sample["changes"], np.int32)
    run_both(oracle_mod, sample["n"], e[:, 0], e[:, 1],
             [(ch[:, 0], ch[:, 1], ch[:, 2].astype(np.uint8))], 0,
             colors=np.array(sample["colors_hi_dir"], np.int32))


# ---------------------------------------------------------------- C1 config --
@pytest.mark.parametrize("rule", [0, 1])
def test_c1_er_10k(oracle_mod, rule):
    say("C1", rule)
    u, v = d.gen_er(10000, 50000, 1)
    batches = d.gen_batches(10000, u, v, 1, 2000, 0.5, 1)
    g, og, st = run_both(oracle_mod, 10000, u, v, batches, rule)
    mono, mc = g.validate()
    assert mono == 0
    assert mc == og.validate()[1]
    assert st[0].updates == 2000 and st[0].insertions == 1000


# -------------------------------------------------------------- random small --
def test_random_small_graphs(oracle_mod):
    rng = random.Random(7)
    for trial in range(60):
        n = rng.randint(3, 40)
        p = rng.choice([0.05, 0.2, 0.5])
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
        if not edges:
            continue
        rule = trial % 2
        ea = np.array(edges, np.int32)
        lines = []
        for _ in range(rng.randint(0, 30)):
            a, b = rng.sample(range(n), 2)
            lines.append((a, b, rng.randint(0, 1)))
        la = np.array(lines, np.int32).reshape(-1, 3)
        colors = None
        if trial % 4 == 3:
            colors = np.array([rng.randint(0, 5) for _ in range(n)], np.int32)
        say("small", trial, n, len(edges), len(lines), rule)
        run_both(oracle_mod, n, ea[:, 0], ea[:, 1],
                 [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))], rule, colors=colors)


def test_zero_insert_zero_delete_and_empty_batches(oracle_mod):
    u, v = d.gen_er(500, 2000, 4)
    b = d.gen_batches(500, u, v, 1, 200, 0.5, 2)[0]
    ins_only = (b[0][b[2] == 1], b[1][b[2] == 1], b[2][b[2] == 1])
    del_only = (b[0][b[2] == 0], b[1][b[2] == 0], b[2][b[2] == 0])
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8))
    for rule in (0, 1):
        say("zero-ins/zero-del", rule)
        run_both(oracle_mod, 500, u, v, [ins_only, del_only, empty], rule)


# ------------------------------------------------------------ multi-batch ---
@pytest.mark.parametrize("rule,scale,nb", [(0, 14, 5), (1, 12, 3)])
def test_rmat_batch_stream(oracle_mod, rule, scale, nb):
    say("rmat stream", rule, scale)
    u, v = d.gen_rmat(scale, 16, seed=1)
    n = 1 << scale
    batches = d.gen_batches(n, u, v, nb, max(200, n // 40), 0.5, 5)
    g, og, _ = run_both(oracle_mod, n, u, v, batches, rule, device_batches=True)
    assert g.validate()[0] == 0


@pytest.mark.parametrize("spare", [0, 16])
def test_regrow_path(oracle_mod, spare):
    # no spare row capacity: insertions overflow their rows, which move into
    # the spare tail (spare=0: default size) or, once a 16-slot tail is used
    # up, make every row relaid out
    say("regrow", spare)
    u, v = d.gen_er(3000, 9000, 8)
    batches = d.gen_batches(3000, u, v, 3, 1500, 0.8, 3)
    opts = d.options(slack_frac=0.0, slack_min=1, spare_slots=spare)
    g, og, st = run_both(oracle_mod, 3000, u, v, batches, 0, opts=opts)
    assert sum(s.regrows for s in st) >= 1
    for x in range(0, 3000, 97):
        assert sorted(g.neighbors(x).tolist()) == sorted(og.neighbors(x).tolist())


@pytest.mark.parametrize("rule", [0, 1])
def test_relocation_into_histogram_rows(oracle_mod, rule):
    # rows with one spare slot: insertions move them into the spare tail, rows
    # of ~62 slots cross 64 slots of capacity there and get a colour histogram
    # each, and a hub of ~5 000 slots moves as several copy chunks
    say("relocation into histogram rows", rule)
    rng = np.random.default_rng(17)
    n = 6000
    u, v = d.gen_er(n, 186000, 17)
    hub_nb = rng.choice(np.arange(1, n), 5000, replace=False).astype(np.int32)
    u = np.concatenate([u, np.zeros(5000, np.int32)])
    v = np.concatenate([v, hub_nb])
    batches = []
    for bu, bv, bi in d.gen_batches(n, u, v, 3, 3000, 0.8, 18):
        extra = rng.integers(1, n, 200).astype(np.int32)
        batches.append((np.concatenate([bu, np.zeros(200, np.int32)]), np.concatenate([bv, extra]),
                        np.concatenate([bi, np.ones(200, np.uint8)])))
    opts = d.options(slack_frac=0.0, slack_min=1)
    g, og, st = run_both(oracle_mod, n, u, v, batches, rule, opts=opts)
    assert sum(s.regrows for s in st) >= 1
    for x in [0] + list(range(1, n, 211)):
        assert sorted(g.neighbors(x).tolist()) == sorted(og.neighbors(x).tolist())


def test_hub_rows_and_mid_rows(oracle_mod):
    # star hubs with > 4096 slots (chunked hub path) and 64..4096 (wave path)
    say("hubs")
    rng = np.random.default_rng(3)
    n = 30000
    edges = set()
    for hub, deg in ((0, 20000), (1, 9000), (2, 3000), (3, 700)):
        for w in rng.choice(np.arange(10, n), deg, replace=False):
            edges.add((hub, int(w)))
    for _ in range(40000):
        a, b = rng.integers(10, n, 2)
        if a != b:
            edges.add((int(min(a, b)), int(max(a, b))))
    e = np.array(sorted(edges), np.int32)
    for rule in (0, 1):
        batches = d.gen_batches(n, e[:, 0], e[:, 1], 2, 3000, 0.5, 11 + rule)
        run_both(oracle_mod, n, e[:, 0], e[:, 1], batches, rule)


def test_hub_mex_window_overflow(oracle_mod):
    # a hub whose neighbours carry colors 0..9999: its mex exceeds the
    # 8192-color window and needs the rescan path
    say("hub window overflow")
    n = 10001
    u = np.zeros(10000, np.int32)
    v = np.arange(1, 10001, dtype=np.int32)
    colors = np.zeros(n, np.int32)
    colors[1:] = np.arange(10000)
    colors[0] = 20000
    # delete the edge to the leaf with color 9000 -> hub (larger color) is the
    # target, its mex becomes 9000
    lines = (np.array([0], np.int32), np.array([9001], np.int32), np.array([0], np.uint8))
    g, og, _ = run_both(oracle_mod, n, u, v, [lines], 0, colors=colors)
    assert g.colors()[0] == 9000


def test_partition_share(oracle_mod):
    rng = random.Random(13)
    for trial in range(12):
        n = rng.randint(50, 3000)
        n_owned = rng.randint(10, n - 5)
        m = rng.randint(n, 6 * n)
        u, v = d.gen_er(n, m, 100 + trial)
        keep = (u < n_owned) | (v < n_owned)
        u, v = u[keep], v[keep]
        ghost = np.array([rng.randint(0, 6) for _ in range(n - n_owned)], np.int32)
        lines = []
        while len(lines) < n // 5:
            a, b = rng.sample(range(n), 2)
            if a >= n_owned and b >= n_owned:
                continue
            lines.append((a, b, rng.randint(0, 1)))
        la = np.array(lines, np.int32)
        say("partition", trial, n, n_owned, u.size)
        g, og, _ = run_both(oracle_mod, n, u, v, [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))],
                            0, n_owned=n_owned, ghost_colors=ghost)
        assert (g.colors()[n_owned:] == ghost).all()
        assert g.validate()[0] == 0


def test_compaction_keeps_graph(oracle_mod):
    say("compaction")
    u, v = d.gen_er(4000, 20000, 12)
    batches = d.gen_batches(4000, u, v, 2, 4000, 0.3, 12)
    g, og, _ = run_both(oracle_mod, 4000, u, v, batches, 0)
    before = [sorted(g.neighbors(x).tolist()) for x in range(0, 4000, 37)]
    g.compact()
    after = [sorted(g.neighbors(x).tolist()) for x in range(0, 4000, 37)]
    assert before == after
    assert g.check_consistency() == 0
    # continue both after compaction (lines drawn from the original edge set:
    # some deletions hit absent edges, which is legal)
    for bu, bv, bi in d.gen_batches(4000, u, v, 1, 1000, 0.5, 99):
        g.process_batch(bu, bv, bi, 0)
        og.apply_batch(bu, bv, bi)
        ph = og.recolor(0)
        assert (g.colors() == og.color).all()
        assert same_trace(g.trace(), ph)


def test_determinism_repeat(oracle_mod):
    say("determinism")
    u, v = d.gen_rmat(13, 16, seed=2)
    n = 1 << 13
    batches = d.gen_batches(n, u, v, 2, 2000, 0.5, 4)
    outs = []
    for _ in range(2):
        g = d.Graph(n, u, v)
        g.static_color(0)
        for bu, bv, bi in batches:
            g.process_batch(bu, bv, bi, 0)
        outs.append(g.colors())
    assert (outs[0] == outs[1]).all()


def test_error_paths_on_device():
    say("errors")
    u, v = d.gen_er(100, 300, 1)
    g = d.Graph(100, u, v)
    with pytest.raises(d.DynColorError) as e:
        g.recolor(0)
    assert e.value.code == d.ERR_STATE
    with pytest.raises(d.DynColorError) as e:
        g.apply_batch([0], [100], [1])
    assert e.value.code == d.ERR_VERTEX
    with pytest.raises(d.DynColorError) as e:
        g.apply_batch([3], [3], [1])
    assert e.value.code == d.ERR_VERTEX
    with pytest.raises(d.DynColorError) as e:
        g.static_color(7)
    assert e.value.code == d.ERR_INVALID
    # a failed batch leaves the graph usable
    g.static_color(0)
    g.process_batch([0], [1], [1], 0)
    assert g.validate()[0] == 0


def test_settle_and_exact_rounds(oracle_mod):
    # exact_rounds=1: every round the reference runs, trace equal to the
    # oracle's; default: the loop stops at settled colours in some batches,
    # with the same colours (run_both checks both, and that every skipped
    # oracle round changed nothing)
    say("settle vs exact rounds")
    u, v = d.gen_rmat(15, 16, seed=3)
    n = 1 << 15
    batches = d.gen_batches(n, u, v, 4, 6000, 0.5, 21)
    _, _, st_exact = run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(exact_rounds=1))
    assert all(s.loop_settled == 0 for s in st_exact)
    _, _, st = run_both(oracle_mod, n, u, v, batches, 0)
    assert any(s.loop_settled for s in st)
    assert sum(s.loop_rounds for s in st) < sum(s.loop_rounds for s in st_exact)


@pytest.mark.parametrize("limit", [-1, 1 << 30])
def test_worklist_modes(oracle_mod, limit):
    # the loop's sets by state-array scans only (-1) or from worklists in
    # every round that is not a pull round (1 << 30): same colours and
    # traces as the oracle either way, and run_both's consistency check finds
    # the state arrays cleared between batches
    say("worklists", limit)
    u, v = d.gen_rmat(15, 16, seed=3)
    n = 1 << 15
    batches = d.gen_batches(n, u, v, 4, 6000, 0.5, 21)
    run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(exact_rounds=1, list_limit=limit))
    run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(list_limit=limit), device_batches=True)
    rng = random.Random(29)
    for trial in range(4):
        nn = rng.randint(200, 3000)
        n_owned = rng.randint(100, nn - 5)
        uu, vv = d.gen_er(nn, 5 * nn, 300 + trial)
        keep = (uu < n_owned) | (vv < n_owned)
        uu, vv = uu[keep], vv[keep]
        ghost = np.array([rng.randint(0, 6) for _ in range(nn - n_owned)], np.int32)
        lines = []
        while len(lines) < nn // 5:
            a, b = rng.sample(range(nn), 2)
            if a >= n_owned and b >= n_owned:
                continue
            lines.append((a, b, rng.randint(0, 1)))
        la = np.array(lines, np.int32)
        run_both(oracle_mod, nn, uu, vv, [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))], 0,
                 opts=d.options(list_limit=limit), n_owned=n_owned, ghost_colors=ghost)


@pytest.mark.parametrize("rule,scale", [(0, 17), (1, 16)])
def test_rmat_stream_with_relocations(oracle_mod, rule, scale):
    # R-MAT 16-17 with little spare row capacity and insert-heavy batches: rows
    # of every class (hubs of thousands of slots too) move into the spare tail
    # as several copy chunks, rows cross 64 slots there, and the ldf rule's
    # large cascades rewrite every row's slot colours after the loop
    say("rmat relocations", rule, scale)
    u, v = d.gen_rmat(scale, 16, seed=12)
    n = 1 << scale
    batches = d.gen_batches(n, u, v, 3, 20000 >> (17 - scale), 0.8, 13)
    opts = d.options(slack_frac=0.02, slack_min=1)
    g, og, st = run_both(oracle_mod, n, u, v, batches, rule, opts=opts, device_batches=True)
    assert sum(s.regrows for s in st) >= 1
    assert g.validate()[0] == 0
    for x in list(range(0, 64)) + list(range(64, n, 4099)):
        assert sorted(g.neighbors(x).tolist()) == sorted(og.neighbors(x).tolist())
