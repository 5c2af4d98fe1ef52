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


def _global_fixture():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "partition_sample.json")) as f:
        return json.load(f)


def test_partition_create_global_fixture():
    # plan from global IDs -> device graph with looked-up colours -> the
    # plan's local change lines: the hand-derived colours
    f = _global_fixture()
    G = f["global"]
    say("partition create global")
    gl = np.array(G["graph_lines"], np.int32)
    cl = np.array(G["change_lines"], np.int32)
    pl = d.PartitionPlan(G["n_global"], G["rank"], G["partition"], G["partition_all"],
                         gl[:, 0], gl[:, 1], cl[:, 0], cl[:, 1], cl[:, 3].astype(np.uint8))
    colors = np.full(G["n_global"], -1, np.int32)
    for x, c in G["color_lines"]:
        colors[x] = c
    g = pl.create_graph(colors)
    l2g, _, _, bu, bv, bi = pl.arrays()
    assert (g.colors() == colors[l2g]).all()
    g.process_batch(bu, bv, bi, 0)
    assert g.colors().tolist() == f["expected"]["colors"]
    assert g.validate() == (0, f["expected"]["max_color"])
    # a vertex the rank holds without a colour is an error
    colors[G["local_to_global"][5]] = -1
    with pytest.raises(d.DynColorError) as e:
        pl.create_graph(colors)
    assert e.value.code == d.ERR_INVALID


@pytest.mark.parametrize("how", ["flag", "env"])
def test_cli_multi_gpu_rank(tmp_path, how):
    # the multi-GPU program's argv (graph, colours, changes) and files
    # (partition<rank>.txt, partitionAll.txt in the working directory)
    import os
    import subprocess
    f = _global_fixture()
    G = f["global"]
    say("cli multi-gpu", how)
    (tmp_path / "partition0.txt").write_text("".join(f"{x}\n" for x in G["partition"]))
    (tmp_path / "partition1.txt").write_text("0\n")
    (tmp_path / "partitionAll.txt").write_text("".join(f"{p}\n" for p in G["partition_all"]))
    (tmp_path / "graph.txt").write_text("".join(f"{a} {b} {w}\n" for a, b, w in G["graph_lines"]))
    (tmp_path / "color.txt").write_text("".join(f"{x} {c}\n" for x, c in G["color_lines"]))
    (tmp_path / "cE.txt").write_text("".join(f"{a} {b} {w} {i}\n" for a, b, w, i in G["change_lines"]))
    cli = os.path.join(os.path.dirname(d.__file__), "dyncolor_cli")
    out = tmp_path / "out.txt"
    args = [cli, "graph.txt", "color.txt", "cE.txt", "--out", str(out)]
    env = dict(os.environ)
    if how == "flag":
        args += ["--rank", "0"]
    else:
        env["OMPI_COMM_WORLD_RANK"] = "0"
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total Time for Vertex Color Update" in r.stdout
    assert "other part vertices 3" in r.stdout
    got = [[int(x) for x in line.split()] for line in out.read_text().splitlines()]
    assert got == G["expected_owned_colors"]


def test_global_ingest_matches_local_share(oracle_mod):
    # the R-MAT 15 partition through the global-ID plan (partition file =
    # owned IDs in order, part vector of 4 contiguous parts, change lines in
    # global IDs) against the oracle on the locally built share: ghost local
    # IDs differ (first appearance vs increasing global ID) but a ghost is
    # never compared by ID, so owned colours agree
    p = _share()
    say("global ingest vs local share")
    N, n_own = p["N"], p["n_own"]
    gu, gv = p["global_edges"]
    part = (np.arange(N) // n_own).astype(np.int32)
    lines = p["global_batches"][0]
    pl = d.PartitionPlan(N, 0, np.arange(n_own), part, gu, gv, lines[0], lines[1], lines[2])
    assert pl.info["n_owned"] == n_own and pl.info["k_dropped"] == 0
    g = pl.create_graph(p["colors_global"])
    _, _, _, bu, bv, bi = pl.arrays()
    g.process_batch(bu, bv, bi, 0)
    og = oracle_mod.OracleGraph(p["n"], p["u"], p["v"], n_owned=n_own)
    og.set_colors(p["init"])
    og.apply_batch(*p["batches"][0])
    og.recolor(0)
    assert (g.colors()[:n_own] == og.color[:n_own]).all()
    assert g.validate()[0] == 0
    assert (pl.border()[1:] > 0).all()
