"""GPU parity of one partition's share at the R-MAT shape of the multi-GPU
variant (partition 0 of 4 contiguous-ID parts, frozen ghost colours from the
whole graph's static colouring).  Unlike the ER partition tests, owned hub
rows here reach hundreds to thousands of slots with ghost neighbours, so the
loop's colour histograms, mid-row table and hub chunk table run with frozen
ghost colours inside them, and cross insertions land on hubs.  Reference:
Multi-GPU/gpuFunctions_undir.cuh:144-283, Multi-GPU/kernel.cu:296-523."""
import numpy as np
import pytest

import dyncolor_amd as d
from partition_shape import rmat_partition
from test_gpu_parity import run_both, say

pytestmark = pytest.mark.gpu

_CACHE = {}


def _share():
    if "p" not in _CACHE:
        _CACHE["p"] = rmat_partition(15, 16, 7, 4, 3, 24000)
    return _CACHE["p"]


def _shape_checks(p):
    n_own = p["n_own"]
    u, v = p["u"], p["v"]
    deg = np.bincount(np.concatenate([u, v]), minlength=p["n"])[:n_own]
    ghost_nbr = np.bincount(np.concatenate([u[v >= n_own], v[u >= n_own]]), minlength=p["n"])[:n_own]
    hubs = (deg >= 512) & (ghost_nbr > 0)
    assert hubs.sum() >= 8, "no owned hub rows with ghost neighbours"
    assert (deg >= 2048).any(), "no row long enough for several hub chunks"
    # cross lines (one ghost endpoint) of both kinds in every batch
    for bu, bv, bi in p["batches"]:
        cross = (bu >= n_own) | (bv >= n_own)
        assert (cross & (bi == 1)).sum() > 100 and (cross & (bi == 0)).sum() > 100


@pytest.mark.parametrize("driver,limit", [(d.DRIVER_PLAIN, 0), (d.DRIVER_PLAIN, -1),
                                          (d.DRIVER_PLAIN, 1 << 30), (d.DRIVER_FUSED, 0)])
def test_rmat_partition_share(oracle_mod, driver, limit):
    p = _share()
    _shape_checks(p)
    say("rmat partition", driver, limit, p["n_own"], p["n"] - p["n_own"], p["u"].size)
    opts = d.options(list_limit=limit)
    g, og, st = run_both(oracle_mod, p["n"], p["u"], p["v"], p["batches"], 0, colors=p["init"],
                         opts=opts, n_owned=p["n_own"], ghost_colors=p["init"][p["n_own"]:],
                         driver=driver)
    assert (g.colors()[p["n_own"]:] == p["init"][p["n_own"]:]).all(), "a ghost colour moved"
    assert g.validate()[0] == 0
    assert og.validate()[0] == 0
    assert sum(s.loop_rounds for s in st) > 0


def test_rmat_partition_share_exact_rounds(oracle_mod):
    # every round the reference runs, traces equal entry for entry
    p = _share()
    say("rmat partition exact rounds")
    run_both(oracle_mod, p["n"], p["u"], p["v"], p["batches"][:2], 0, colors=p["init"],
             opts=d.options(exact_rounds=1), n_owned=p["n_own"],
             ghost_colors=p["init"][p["n_own"]:])


@pytest.mark.parametrize("driver", [d.DRIVER_PLAIN, d.DRIVER_FUSED])
def test_partition_sample_fixture(driver):
    import json
    import os
    from traces import same_trace
    say("partition fixture", driver)
    p = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "partition_sample.json")))
    e = np.array(p["edges"], np.int32)
    ch = np.array(p["changes"], np.int32)
    g = d.Graph(p["n_owned"], e[:, 0], e[:, 1], n_ghost=p["n_ghost"],
                ghost_colors=np.array(p["ghost_colors"], np.int32))
    g.set_driver(driver)
    g.set_colors(np.array(p["colors"], np.int32))
    g.process_batch(ch[:, 0], ch[:, 1], ch[:, 2].astype(np.uint8), 0)
    assert g.colors().tolist() == p["expected"]["colors"]
    assert same_trace(g.trace(), p["expected"]["trace"])
    assert g.validate() == (p["expected"]["monochromatic"], p["expected"]["max_color"])
    assert g.check_consistency() == 0
