"""One partition's share of an R-MAT graph, at the shape of the multi-GPU
variant (Multi-GPU/kernel.cu:140-160, Multi-GPU/all_structure_undir.cuh:
229-312): the graph's vertices are split into `parts` contiguous-ID parts;
partition 0 owns IDs [0, N / parts), keeps every edge with an owned endpoint,
and holds every other endpoint as a ghost whose colour is frozen at the whole
graph's static colouring.  Update batches touch owned vertices only: half
insertions (an owned vertex and a uniform other vertex), half deletions of
edges the partition holds.

Local IDs: owned vertices keep their IDs, ghosts follow in increasing global
ID order."""
import numpy as np

import dyncolor_amd as d


def rmat_partition(scale, edge_factor, seed, parts, nb, bs):
    u, v = d.gen_rmat(scale, edge_factor, seed=seed)
    N = 1 << scale
    n_own = N // parts
    full = d.Graph(N, u, v)
    full.static_color(0)
    colors = full.colors()
    full.close()
    keep = (u < n_own) | (v < n_own)
    pu, pv = u[keep].astype(np.int64), v[keep].astype(np.int64)
    rng = np.random.default_rng(seed)
    lines = []
    for _ in range(nb):
        k_ins = bs // 2
        a = rng.integers(0, n_own, k_ins)
        b = (a + rng.integers(1, N, k_ins)) % N
        e = rng.integers(0, pu.size, bs - k_ins)
        bu = np.concatenate([a, pu[e]])
        bv = np.concatenate([b, pv[e]])
        bi = np.concatenate([np.ones(k_ins, np.uint8), np.zeros(bs - k_ins, np.uint8)])
        perm = rng.permutation(bs)
        lines.append((bu[perm], bv[perm], bi[perm]))
    others = np.concatenate([pu[pu >= n_own], pv[pv >= n_own]] +
                            [x[x >= n_own] for b in lines for x in (b[0], b[1])])
    ghosts = np.unique(others)
    lid = np.full(N, -1, np.int64)
    lid[:n_own] = np.arange(n_own)
    lid[ghosts] = n_own + np.arange(ghosts.size)
    n = n_own + ghosts.size
    init = np.zeros(n, np.int32)
    init[:n_own] = colors[:n_own]
    init[n_own:] = colors[ghosts]
    local_batches = [(lid[b[0]].astype(np.int32), lid[b[1]].astype(np.int32), b[2]) for b in lines]
    return {
        "N": N, "n_own": n_own, "n": n, "ghosts": ghosts, "colors_global": colors,
        "u": lid[pu].astype(np.int32), "v": lid[pv].astype(np.int32), "init": init,
        "batches": local_batches, "global_edges": (pu, pv), "global_batches": lines,
    }
