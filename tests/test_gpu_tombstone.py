"""Deletion lines grouped by scanned row (dc_options.group_deletions): the
lines are sorted by the row they scan and each row is read once for all of
them, long rows by whole workgroups.  Checked against the oracle and the
per-line pass on streams, on hubs that lose thousands of edges in one batch
(with duplicate lines in both orientations), and on a partition share with
cross deletions."""
import random

import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("group", [1, -1])
def test_grouped_deletions_stream(oracle_mod, group):
    say("grouped deletions stream", group)
    u, v = d.gen_rmat(15, 16, seed=8)
    n = 1 << 15
    batches = d.gen_batches(n, u, v, 3, 8000, 0.4, 17)
    run_both(oracle_mod, n, u, v, batches, 0, opts=d.options(group_deletions=group),
             device_batches=True)
    run_both(oracle_mod, n, u, v, batches[:2], 1, opts=d.options(group_deletions=group))


def test_grouped_deletions_hubs(oracle_mod):
    # two hubs of 20 000 and 6 000 slots (long-row runs) and mid rows; one
    # batch deletes 5 000 of the first hub's edges, 1 500 of the second's
    # (each line once as (hub, w) and some again as (w, hub)), plus edges
    # between the hubs' neighbours, and inserts new hub edges
    say("grouped deletions hubs")
    rng = np.random.default_rng(4)
    n = 40000
    edges = set()
    for hub, deg in ((0, 20000), (1, 6000), (2, 900)):
        for w in rng.choice(np.arange(10, n), deg, replace=False):
            edges.add((hub, int(w)))
    for _ in range(60000):
        a, b = rng.integers(10, n, 2)
        if a != b:
            edges.add((int(min(a, b)), int(max(a, b))))
    e = np.array(sorted(edges), np.int32)
    h0 = e[e[:, 0] == 0][:, 1]
    h1 = e[e[:, 0] == 1][:, 1]
    lines = [(0, int(w), 0) for w in rng.choice(h0, 5000, replace=False)]
    lines += [(int(w), 1, 0) for w in rng.choice(h1, 1500, replace=False)]
    lines += [(w, x, 0) for (x, w, _) in lines[:700]]           # the other orientation
    others = e[e[:, 0] >= 10]
    lines += [(int(a), int(b), 0) for a, b in others[rng.choice(others.shape[0], 4000, replace=False)]]
    lines += [(0, int(w), 1) for w in rng.integers(10, n, 300)]
    random.Random(3).shuffle(lines)
    la = np.array(lines, np.int32)
    b = (la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))
    for group in (1, -1):
        g, og, _ = run_both(oracle_mod, n, e[:, 0], e[:, 1], [b], 0,
                            opts=d.options(group_deletions=group))
        for x in (0, 1, 2, 11, 12345):
            assert sorted(g.neighbors(x).tolist()) == sorted(og.neighbors(x).tolist())


def test_grouped_deletions_partition(oracle_mod):
    say("grouped deletions partition")
    rng = random.Random(21)
    n, n_owned = 3000, 1800
    u, v = d.gen_er(n, 6 * n, 77)
    keep = (u < n_owned) | (v < n_owned)
    u, v = u[keep], v[keep]
    ghost = np.array([rng.randint(0, 6) for _ in range(n - n_owned)], np.int32)
    idx = list(range(u.size))
    rng.shuffle(idx)
    lines = [(int(u[i]), int(v[i]), 0) for i in idx[:2500]]
    while len(lines) < 3500:
        a, b = rng.sample(range(n), 2)
        if a < n_owned or b < n_owned:
            lines.append((a, b, 1))
    la = np.array(lines, np.int32)
    run_both(oracle_mod, n, u, v, [(la[:, 0], la[:, 1], la[:, 2].astype(np.uint8))], 0,
             opts=d.options(group_deletions=1), n_owned=n_owned, ghost_colors=ghost)
