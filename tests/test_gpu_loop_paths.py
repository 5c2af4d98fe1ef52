"""The independent-rule loop's count-pass paths, each forced and checked
against the oracle: every round recounted by pull, no round pulled while
rounds push from more than heavy_slots changed row slots (with and without R
lists), and the automatic choice with a low threshold (both paths in one
batch stream).  Also: a loop that hits max_rounds returns DC_ERR_ROUNDS with
the colours of the rounds it ran, and the graph stays usable."""
import numpy as np
import pytest

import dyncolor_amd as d
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu


def _stream():
    u, v = d.gen_rmat(15, 16, seed=3)
    n = 1 << 15
    return n, u, v, d.gen_batches(n, u, v, 3, 6000, 0.5, 21)


def test_every_round_pulls(oracle_mod):
    n, u, v, batches = _stream()
    say("pull_mode 1")
    _, _, st = run_both(oracle_mod, n, u, v, batches, 0,
                        opts=d.options(exact_rounds=1, pull_mode=1))
    # every round with a non-empty C recounts by pull (the last round's C is
    # empty: nothing to count)
    assert all(s.pull_rounds == max(s.loop_rounds - 1, 0) for s in st)
    assert sum(s.pull_rounds for s in st) > 0


@pytest.mark.parametrize("limit", [0, 1 << 30])
def test_heavy_rounds_push(oracle_mod, limit):
    n, u, v, batches = _stream()
    say("pull_mode -1", limit)
    _, _, st = run_both(oracle_mod, n, u, v, batches, 0,
                        opts=d.options(exact_rounds=1, pull_mode=-1, heavy_slots=2048,
                                       list_limit=limit))
    assert sum(s.pull_rounds for s in st) == 0
    assert sum(s.heavy_push_rounds for s in st) > 0
    assert sum(s.heavy_push_listed for s in st) > 0


def test_automatic_choice_takes_both_paths(oracle_mod):
    n, u, v, batches = _stream()
    say("pull_mode 0, low threshold")
    _, _, st = run_both(oracle_mod, n, u, v, batches, 0,
                        opts=d.options(exact_rounds=1, heavy_slots=4096), device_batches=True)
    assert sum(s.pull_rounds for s in st) > 0
    assert sum(s.heavy_push_rounds for s in st) > 0


def test_max_rounds_error_then_next_batch(oracle_mod):
    # a batch that needs more than 3 loop rounds fails with DC_ERR_ROUNDS
    # after exactly 3 (the oracle's colours after 3 rounds); a later batch
    # that fits in 3 rounds matches the oracle again.  Large and 2-line
    # batches alternate.
    u, v = d.gen_rmat(13, 16, seed=5)
    n = 1 << 13
    cap = 3
    opts = d.options(exact_rounds=1, max_rounds=cap)
    g = d.Graph(n, u, v, opts)
    g.static_color(0)
    og = oracle_mod.OracleGraph(n, u, v)
    og.static_color(0)
    big = d.gen_batches(n, u, v, 20, 1500, 0.5, 33)
    tiny = d.gen_batches(n, u, v, 20, 2, 0.5, 34)
    batches = [b for pair in zip(big, tiny) for b in pair]
    failed = passed = 0
    for i, (bu, bv, bi) in enumerate(batches):
        og.apply_batch(bu, bv, bi)
        try:
            og.recolor(0, max_loop=cap)
            ok = True
        except RuntimeError:
            ok = False
        g.apply_batch(bu, bv, bi)
        if ok:
            g.recolor(0)
            passed += 1
        else:
            with pytest.raises(d.DynColorError) as e:
                g.recolor(0)
            assert e.value.code == d.ERR_ROUNDS
            failed += 1
        got = g.colors()
        assert (got == og.color).all(), f"batch {i} (loop cap hit: {not ok})"
        assert g.check_consistency() == 0
        if failed and passed >= 2 and not ok:
            break
    say("max_rounds", failed, passed)
    assert failed >= 1 and passed >= 1
