"""GPU parity of the fused driver (all rounds of a batch in one persistent
cooperative launch) against the CPU oracle, and against the plain driver."""
import random

import numpy as np
import pytest

import dyncolor_amd as d
from traces import same_trace
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu
FUSED = d.DRIVER_FUSED


@pytest.mark.parametrize("rule,key", [(0, "independent_hb_colors"), (1, "ldf_hb_colors")])
def test_fused_sample(sample, rule, key):
    say("fused sample", rule)
    e = np.array(sample["edges"], np.int32)
    ch = np.array(sample["changes"], np.int32)
    g = d.Graph(sample["n"], e[:, 0], e[:, 1])
    g.set_driver(FUSED)
    g.set_colors(sample["colors_hb_dir"])
    g.process_batch(ch[:, 0], ch[:, 1], ch[:, 2].astype(np.uint8), rule)
    exp = sample["expected"][key]
    assert g.colors().tolist() == exp["colors"]
    assert same_trace(g.trace(), exp["trace"])


@pytest.mark.parametrize("rule", [0, 1])
def test_fused_c1(oracle_mod, rule):
    say("fused C1", rule)
    u, v = d.gen_er(10000, 50000, 1)
    batches = d.gen_batches(10000, u, v, 2, 2000, 0.5, 1)
    g, _, _ = run_both(oracle_mod, 10000, u, v, batches, rule, driver=FUSED)
    assert g.validate()[0] == 0


def test_fused_random_small(oracle_mod):
    rng = random.Random(17)
    for trial in range(40):
        n = rng.randint(3, 40)
        p = rng.choice([0.05, 0.2, 0.5])
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
        if not edges:
            continue
        ea = np.array(edges, np.int32)
        lines = []
        for _ in range(rng.randint(0, 30)):
            a, b = rng.sample(range(n), 2)
            lines.append((a, b, rng.randint(0, 1)))
        la = np.array(lines, np.int32).reshape(-1, 3)
        say("fused small", trial)
        run_both(oracle_mod, n, ea[:, 0], ea[:, 1],
                 [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))], trial % 2, driver=FUSED)


def test_fused_no_loop_batches(oracle_mod):
    # batches whose pre-loop rounds change nothing or whose loop never runs,
    # followed by more batches (the write buffer must be caught up)
    u, v = d.gen_er(500, 2000, 4)
    b = d.gen_batches(500, u, v, 3, 60, 0.5, 2)
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8))
    for rule in (0, 1):
        say("fused no-loop", rule)
        run_both(oracle_mod, 500, u, v, [b[0], empty, b[1], empty, b[2]], rule, driver=FUSED)


def test_fused_rmat_stream_and_hubs(oracle_mod):
    say("fused rmat")
    u, v = d.gen_rmat(14, 16, seed=3)
    n = 1 << 14
    batches = d.gen_batches(n, u, v, 3, 600, 0.5, 8)
    run_both(oracle_mod, n, u, v, batches, 0, driver=FUSED, device_batches=True)
    u, v = d.gen_rmat(12, 16, seed=4)
    batches = d.gen_batches(1 << 12, u, v, 2, 300, 0.5, 9)
    run_both(oracle_mod, 1 << 12, u, v, batches, 1, driver=FUSED)


def test_fused_partition(oracle_mod):
    rng = random.Random(23)
    for trial in range(6):
        n = rng.randint(100, 3000)
        n_owned = rng.randint(20, n - 5)
        u, v = d.gen_er(n, rng.randint(n, 5 * n), 200 + trial)
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
        say("fused partition", trial)
        run_both(oracle_mod, n, u, v, [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))], 0,
                 n_owned=n_owned, ghost_colors=ghost, driver=FUSED)


def test_fused_equals_plain_larger(oracle_mod):
    # scale 16 stream: fused and plain must agree bit for bit (and be valid)
    say("fused vs plain scale 16")
    u, v = d.gen_rmat(16, 16, seed=5)
    n = 1 << 16
    batches = d.gen_batches(n, u, v, 3, 3000, 0.5, 10)
    outs = []
    for drv in (d.DRIVER_PLAIN, FUSED):
        g = d.Graph(n, u, v)
        g.set_driver(drv)
        g.static_color(0)
        traces = []
        for bu, bv, bi in batches:
            g.process_batch(bu, bv, bi, 0)
            traces.append(g.trace().tolist())
        outs.append((g.colors(), traces))
        assert g.validate()[0] == 0
    assert (outs[0][0] == outs[1][0]).all()
    assert outs[0][1] == outs[1][1]


def test_fused_worklists_every_round(oracle_mod):
    # the persistent driver with worklists in every round that is not a pull
    # round: colours and traces equal to the oracle's
    say("fused worklists")
    u, v = d.gen_rmat(14, 16, seed=5)
    n = 1 << 14
    batches = d.gen_batches(n, u, v, 3, 4000, 0.5, 33)
    run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(exact_rounds=1, list_limit=1 << 30),
             driver=d.DRIVER_FUSED)
