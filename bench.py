"""Throughput benchmark of the dynamic-recoloring hot path on one MI355X.

Default workload (c2, the metric's single-GPU configuration): R-MAT scale 22
(edge factor 16, a/b/c = .57/.19/.19, self loops and duplicates removed, IDs
not permuted), greedy static coloring, then consecutive update batches of 100k
lines (50% insertions of uniform non-edges, 50% deletions of existing edges),
each applied to the graph left by the previous one.  A step = apply one
device-resident batch + run every recolor round until the coloring is final.
The first --warmup batches are untimed; the next --steps batches are timed.

Other workloads (--workload): c1 ER 10k/50k one batch; c3 one partition of
an R-MAT scale-24 graph; c4 512 graphs recolored per step in one launch vs
one plain call per graph; c5 served request stream over 8 resident graphs.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "valid recolored edge updates/sec + p50/p99 batch latency vs plain per-batch"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def base_record(args, value, elapsed_steps, steps, warmup, config):
    return {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": elapsed_steps / max(steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": config,
    }


def roofline(kernel, algo_bytes, kernel_ms, launches):
    kern_s = kernel_ms / 1e3
    achieved = algo_bytes / kern_s / 1e9 if kern_s > 0 else 0.0
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "launches": launches, "avg_launch_us": kernel_ms * 1e3 / max(launches, 1),
            "algorithmic_bytes": int(algo_bytes)}


def traffic_for(kernel):
    """HBM bytes per launch of the kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or
    None when no summary covers it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    ent = t.get(kernel.split(" ")[0])
    return None if ent is None else ent.get("bytes_per_launch")


def set_driver(g, d, args):
    g.set_driver(d.DRIVER_FUSED if args.driver == "fused" else d.DRIVER_PLAIN, args.blocks_per_cu)


# ----------------------------------------------------------------- c1 / c2 --
def run_stream(args, d, rule, n, u, v, workload):
    """One graph, a stream of device-resident batches (c1, c2)."""
    # per-launch hipEvents on the count pass (they cost ~7 % when recorded in
    # the timed batches): by default in an untimed replay of the warm-up and
    # timed batches on a fresh copy of the initial graph (deterministic: the
    # same rounds and launches), with --events-after in 2 batches after the
    # timed ones
    nprof = 2 if args.events_after else 0
    nb = args.warmup + args.steps + nprof
    t = time.perf_counter()
    batches = d.gen_batches(n, u, v, nb, args.batch, 0.5, args.seed + 1000)
    log(f"{nb} batches of {args.batch} ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    g = d.Graph(n, u, v)
    g.static_color(rule)
    g.synchronize()
    mono0, maxc0 = g.validate()
    log(f"device graph + static coloring: {time.perf_counter() - t:.1f}s, "
        f"colors {maxc0 + 1}, monochromatic {mono0}")
    dev_batches = [g.upload_batch(*b) for b in batches]
    g.set_profiling(False)
    set_driver(g, d, args)
    prof = {"ms": 0.0, "bytes": 0, "launches": 0, "kms": 0.0, "rbytes": 0, "rl": 0}

    def add_prof(st):
        prof["ms"] += st.ms_count_pass
        prof["bytes"] += st.count_pass_bytes
        prof["launches"] += st.count_pass_launches
        prof["kms"] += st.ms_recolor_kernel
        prof["rbytes"] += st.recolor_bytes
        prof["rl"] += st.recolor_launches

    colors_after_first = None
    for i in range(args.warmup):
        g.apply_device_batch(dev_batches[i])
        g.recolor(rule)
        if i == 0:
            colors_after_first = g.colors()
    g.synchronize()

    lat = []
    tot = {"bytes": 0, "kms": 0.0, "launches": 0, "rounds": 0, "changed": 0, "evaluated": 0,
           "slots": 0, "regrows": 0, "settled": 0}
    t0 = time.perf_counter()
    last_log = t0
    for i in range(args.warmup, args.warmup + args.steps):
        tb = time.perf_counter()
        g.apply_device_batch(dev_batches[i])
        st = g.recolor(rule)
        lat.append(time.perf_counter() - tb)
        if tb - last_log > 20.0:   # slow workloads (the LDF rule at scale) report progress
            log(f"timed batch {i - args.warmup + 1}/{args.steps}")
            last_log = tb
        tot["bytes"] += st.recolor_bytes
        tot["kms"] += st.ms_recolor_kernel
        tot["launches"] += st.recolor_launches
        tot["rounds"] += st.loop_rounds
        tot["changed"] += st.changed
        tot["evaluated"] += st.evaluated
        tot["slots"] += st.slots_scanned
        tot["regrows"] += st.regrows
        tot["settled"] += st.loop_settled
    g.synchronize()
    elapsed = time.perf_counter() - t0
    mono, maxc = g.validate()
    if mono != 0:
        log(f"INVALID: {mono} monochromatic edges after the timed batches")
    replay_rounds = None
    if args.events_after:
        g.set_profiling(True)
        for i in range(args.warmup + args.steps, nb):
            g.apply_device_batch(dev_batches[i])
            add_prof(g.recolor(rule))
        g.set_profiling(False)
        g.synchronize()
    else:
        colors_timed = g.colors()
        del dev_batches, g
        g = d.Graph(n, u, v)
        g.static_color(rule)
        set_driver(g, d, args)
        dev_batches = [g.upload_batch(*b) for b in batches]
        replay_rounds = 0
        last_log = time.perf_counter()
        for i in range(args.warmup + args.steps):
            if time.perf_counter() - last_log > 20.0:
                log(f"replay batch {i + 1}/{args.warmup + args.steps}")
                last_log = time.perf_counter()
            g.set_profiling(i >= args.warmup)
            g.apply_device_batch(dev_batches[i])
            st = g.recolor(rule)
            if i >= args.warmup:
                add_prof(st)
                replay_rounds += st.loop_rounds
        g.set_profiling(False)
        g.synchronize()
        if replay_rounds != tot["rounds"] or not (g.colors() == colors_timed).all():
            log(f"replay differs: {replay_rounds} rounds against {tot['rounds']}")
            replay_rounds = None
    log(f"timed {args.steps} batches in {elapsed:.3f}s, rounds {tot['rounds']}, colors {maxc + 1}")
    value = args.steps * args.batch / elapsed if mono == 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    out = base_record(args, value, elapsed, args.steps, args.warmup, {
        "workload": workload, "edges": int(u.size), "vertices": int(n), "batch": args.batch,
        "rule": args.rule, "driver": args.driver, "colors": int(maxc + 1),
        "monochromatic": int(mono)})
    out["latency_ms"] = {"p50": pct(lat_ms, 50), "p99": pct(lat_ms, 99), "max": float(lat_ms.max())}
    if args.driver == "plain" and rule == d.RULE_INDEPENDENT and prof["launches"]:
        out["roofline"] = roofline("k_round_c (count pass)", prof["bytes"], prof["ms"],
                                   prof["launches"])
    else:
        out["roofline"] = roofline("k_loop_persistent" if args.driver == "fused" else
                                   "k_recolor", prof["rbytes"], prof["kms"], prof["rl"])
    out["roofline"]["measured_on"] = (
        f"{nprof} batches after the timed ones" if args.events_after else
        f"replay of the warm-up and timed batches (same {replay_rounds} rounds)")
    out["roofline"]["traffic"] = traffic_for(out["roofline"]["kernel"])
    out["work"] = {"loop_rounds": tot["rounds"], "evaluated": tot["evaluated"],
                   "changed": tot["changed"], "row_slots_scanned": tot["slots"],
                   "batches_with_row_growth": tot["regrows"],
                   "batches_settled": tot["settled"]}
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, u, v, batches[0], rule, colors_after_first)
    return out


def cpu_baseline(n, u, v, batch, rule, gpu_colors_after_first):
    """The oracle (single thread) on the first batch of the same stream."""
    from oracle import oracle as O
    try:
        O.lib()
    except Exception as e:  # oracle not built
        log(f"cpu baseline skipped: {e}")
        return None
    t = time.perf_counter()
    og = O.OracleGraph(n, u, v)
    og.static_color(rule)
    log(f"oracle graph + static coloring: {time.perf_counter() - t:.1f}s")
    bu, bv, bi = batch
    t = time.perf_counter()
    og.apply_batch(bu, bv, bi)
    og.recolor(rule)
    dt = time.perf_counter() - t
    same = None
    if gpu_colors_after_first is not None:
        same = bool((og.color == gpu_colors_after_first).all())
    log(f"oracle batch: {dt:.1f}s, colors equal to GPU: {same}")
    return {"value": bu.size / dt, "unit": "updates/s", "cores": 1, "kind": "port",
            "sample": f"first batch ({bu.size} lines) of the same stream from the same initial "
                      f"graph, single-thread oracle replay (apply + all rounds)",
            "seconds": dt, "bit_equal_to_gpu": same}


# ---------------------------------------------------------------------- c4 --
def c4_graphs(d, count, seed):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(math.log(20000), 1.5, count)), 1000, 2_000_000)
    degs = rng.uniform(8, 32, count)
    graphs, kinds = [], []
    for i in range(count):
        if i % 2 == 0:
            n = int(sizes[i])
            m = int(n * degs[i] / 2)
            u, v = d.gen_er(n, m, seed * 100003 + i)
            kinds.append("er")
        else:
            scale = int(math.ceil(math.log2(sizes[i])))
            ef = max(1, int(round(degs[i] / 2)))
            u, v = d.gen_rmat(scale, ef, seed=seed * 100003 + i)
            n = 1 << scale
            kinds.append("rmat")
        graphs.append((n, u, v))
    return graphs, kinds, rng


def run_c4(args, d, rule):
    t = time.perf_counter()
    graphs, kinds, rng = c4_graphs(d, args.graphs, args.seed)
    nsteps = args.warmup + args.steps
    per = []
    for i, (n, u, v) in enumerate(graphs):
        frac = rng.uniform(0.01, 0.10)
        bs = max(2, int(u.size * frac))
        per.append(d.gen_batches(n, u, v, nsteps, bs, 0.5, args.seed * 7919 + i))
    tot_edges = sum(g[2].size for g in graphs)
    step_lines = [sum(p[s][0].size for p in per) for s in range(nsteps)]
    log(f"c4: {len(graphs)} graphs, {sum(g[0] for g in graphs)} vertices, {tot_edges} edges, "
        f"{step_lines[-1]} lines per step ({time.perf_counter() - t:.1f}s)")

    # batched: one arena, one launch per step
    n, U, V, off = d.union_edges(graphs)
    arena = d.Graph(n, U, V)
    arena.static_color(rule)
    set_driver(arena, d, args)
    arena.set_profiling(True)
    ub = [arena.upload_batch(*d.union_batch([p[s] for p in per], off)) for s in range(nsteps)]
    for s in range(args.warmup):
        arena.apply_device_batch(ub[s])
        arena.recolor(rule)
    arena.synchronize()
    lat, kms, byts, rounds = [], 0.0, 0, 0
    t0 = time.perf_counter()
    for s in range(args.warmup, nsteps):
        tb = time.perf_counter()
        arena.apply_device_batch(ub[s])
        st = arena.recolor(rule)
        lat.append(time.perf_counter() - tb)
        kms += st.ms_recolor_kernel
        byts += st.recolor_bytes
        rounds += st.loop_rounds
    arena.synchronize()
    t_batched = time.perf_counter() - t0
    mono, _ = arena.validate()
    lines = sum(step_lines[args.warmup:])
    batched = lines / t_batched if mono == 0 else 0.0
    log(f"c4 batched: {t_batched:.3f}s for {args.steps} steps, rounds {rounds}, mono {mono}")

    # plain per-batch baseline: one graph handle and one plain call per graph
    t_plain = 0.0
    plain_colors = []
    handles = []
    for i, (gn, gu, gv) in enumerate(graphs):
        g = d.Graph(gn, gu, gv)
        g.static_color(rule)
        handles.append((g, [g.upload_batch(*per[i][s]) for s in range(nsteps)]))
    for s in range(nsteps):
        tb = time.perf_counter()
        for g, bl in handles:
            g.apply_device_batch(bl[s])
            g.recolor(rule)
        if s >= args.warmup:
            t_plain += time.perf_counter() - tb
    for g, _ in handles:
        plain_colors.append(g.colors())
    plain = lines / t_plain
    got = arena.colors()
    same = all((got[off[i]:off[i + 1]] == plain_colors[i]).all() for i in range(len(graphs)))
    log(f"c4 plain: {t_plain:.3f}s, speedup {batched / plain:.1f}x, colors equal: {same}")
    out = base_record(args, batched, t_batched, args.steps, args.warmup, {
        "workload": f"c4_{len(graphs)}graphs_lognormal20k_er_rmat_1to10pct",
        "graphs": len(graphs), "vertices": int(n), "edges": int(tot_edges),
        "lines_per_step": int(step_lines[-1]), "rule": args.rule, "driver": args.driver,
        "monochromatic": int(mono)})
    out["latency_ms"] = {"p50": pct(np.array(lat) * 1e3, 50), "p99": pct(np.array(lat) * 1e3, 99)}
    out["plain_per_batch"] = {"value": plain, "seconds": t_plain, "calls_per_step": len(graphs)}
    out["speedup_vs_plain"] = batched / plain
    out["bit_equal_to_plain"] = bool(same)
    out["roofline"] = roofline("k_fused" if args.driver == "fused" else "k_recolor<M_LOOP>",
                               byts, kms, args.steps)
    return out


# ---------------------------------------------------------------------- c5 --
def c5_requests(d, graphs, rate, duration, rng):
    nreq = max(1, int(rng.poisson(rate * duration)))
    arrival = np.sort(rng.uniform(0, duration, nreq))
    gid = rng.integers(0, len(graphs), nreq).astype(np.int32)
    sizes = np.clip(np.exp(rng.normal(math.log(5000), 1.2, nreq)), 10, 500_000).astype(np.int64)
    roff = np.zeros(nreq + 1, np.int64)
    roff[1:] = np.cumsum(sizes)
    L = int(roff[-1])
    u = np.zeros(L, np.int32)
    v = np.zeros(L, np.int32)
    ins = (rng.random(L) < 0.6).astype(np.uint8)
    for r in range(nreq):
        gn, gu, gv = graphs[gid[r]]
        lo, hi = roff[r], roff[r + 1]
        k = hi - lo
        a = rng.integers(0, gn, k)
        b = (a + rng.integers(1, gn, k)) % gn          # uniform pair, no self loop
        e = rng.integers(0, gu.size, k)                 # an original edge (may be gone)
        isi = ins[lo:hi] == 1
        u[lo:hi] = np.where(isi, a, gu[e])
        v[lo:hi] = np.where(isi, b, gv[e])
    return gid, roff, u, v, ins, arrival


def run_c5(args, d, rule):
    scales = [18, 19, 20, 21, 22, 23, 18, 20]
    t = time.perf_counter()
    graphs = []
    for i, s in enumerate(scales):
        u, v = d.gen_rmat(s, 16, seed=args.seed + 31 * i)
        graphs.append((1 << s, u, v))
    n, U, V, off = d.union_edges(graphs)
    arena = d.Graph(n, U, V)
    arena.static_color(rule)
    set_driver(arena, d, args)
    log(f"c5: 8 resident graphs, {n} vertices, {U.size} edges ({time.perf_counter() - t:.1f}s)")
    rng = np.random.default_rng(args.seed)
    srv = d.Server(arena, off, budget_lines=args.budget_lines, rule=rule)
    sweep = []
    for rate in args.rates:
        gid, roff, u, v, ins, arrival = c5_requests(d, graphs, rate, args.duration, rng)
        lat, step, st = srv.run(gid, roff, u, v, ins, arrival)
        ok = lat >= 0
        lines = int(roff[-1])
        rec = {"rate_req_s": rate, "requests": int(gid.size), "served": int(st.served),
               "rejected": int(st.rejected), "lines": lines, "steps": int(st.steps),
               "elapsed_s": st.elapsed_s, "busy_s": st.busy_s,
               "updates_per_s": st.lines / st.elapsed_s,
               "p50_ms": pct(lat[ok] * 1e3, 50), "p99_ms": pct(lat[ok] * 1e3, 99),
               "max_step_lines": int(st.max_step_lines)}
        log(f"c5 rate {rate}/s: {rec}")
        sweep.append(rec)
    mono, _ = arena.validate()
    # plain per-batch capacity: a sample of requests, one plain call each on
    # its own graph handle, no coalescing
    rng2 = np.random.default_rng(args.seed + 1)
    handles = []
    for gn, gu, gv in graphs:
        g = d.Graph(gn, gu, gv)
        g.static_color(rule)
        handles.append(g)
    gid, roff, u, v, ins, _ = c5_requests(d, graphs, 40, 1.0, rng2)
    t0 = time.perf_counter()
    plat = []
    for r in range(gid.size):
        tb = time.perf_counter()
        lo, hi = roff[r], roff[r + 1]
        handles[gid[r]].process_batch(u[lo:hi], v[lo:hi], ins[lo:hi], rule)
        plat.append(time.perf_counter() - tb)
    t_plain = time.perf_counter() - t0
    plain = float(roff[-1]) / t_plain
    best = max(sweep, key=lambda r: r["updates_per_s"])
    out = base_record(args, best["updates_per_s"], best["busy_s"], best["steps"], 0, {
        "workload": "c5_8rmat_s18-23_poisson_lognormal5k_60ins",
        "vertices": int(n), "edges": int(U.size), "budget_lines": args.budget_lines,
        "rule": args.rule, "driver": args.driver, "monochromatic": int(mono)})
    out["latency_ms"] = {"p50": best["p50_ms"], "p99": best["p99_ms"]}
    out["sweep"] = sweep
    out["plain_per_batch"] = {"value": plain, "requests": int(gid.size), "seconds": t_plain,
                              "p50_ms": pct(np.array(plat) * 1e3, 50),
                              "p99_ms": pct(np.array(plat) * 1e3, 99)}
    out["speedup_vs_plain"] = best["updates_per_s"] / plain
    return out


# ---------------------------------------------------------------------- c3 --
def run_c3(args, d, rule):
    """Partition 0 of 4 of an R-MAT scale-24 graph (owned + frozen ghosts)."""
    scale = args.c3_scale
    t = time.perf_counter()
    u, v = d.gen_rmat(scale, 16, seed=args.seed)
    N = 1 << scale
    n_own = N // 4
    log(f"c3: rmat {scale}: {u.size} edges ({time.perf_counter() - t:.1f}s)")
    # ghost colors: the whole graph's static coloring
    full = d.Graph(N, u, v)
    full.static_color(rule)
    colors = full.colors()
    full.close()
    keep = (u < n_own) | (v < n_own)
    pu, pv = u[keep], v[keep]
    rng = np.random.default_rng(args.seed)
    nsteps = args.warmup + args.steps
    bs = args.batch
    # batches: half insertions (owned endpoint + uniform vertex), half deletions
    # of existing edges touching the partition
    blines = []
    for s in range(nsteps):
        k_ins = bs // 2
        a = rng.integers(0, n_own, k_ins)
        b = (a + rng.integers(1, N, k_ins)) % N
        e = rng.integers(0, pu.size, bs - k_ins)
        bu = np.concatenate([a, pu[e]]).astype(np.int64)
        bv = np.concatenate([b, pv[e]]).astype(np.int64)
        bi = np.concatenate([np.ones(k_ins, np.uint8), np.zeros(bs - k_ins, np.uint8)])
        blines.append((bu, bv, bi))
    # ghost IDs: every non-owned vertex named by the edges or the batches
    others = np.concatenate([pu[pu >= n_own], pv[pv >= n_own]] +
                            [x[x >= n_own] for b in blines for x in (b[0], b[1])])
    ghosts = np.unique(others)
    lid = np.full(N, -1, np.int64)
    lid[:n_own] = np.arange(n_own)
    lid[ghosts] = n_own + np.arange(ghosts.size)
    g = d.Graph(n_own, lid[pu].astype(np.int32), lid[pv].astype(np.int32),
                n_ghost=int(ghosts.size), ghost_colors=colors[ghosts])
    init = np.zeros(n_own + ghosts.size, np.int32)
    init[:n_own] = colors[:n_own]
    init[n_own:] = colors[ghosts]
    g.set_colors(init)
    set_driver(g, d, args)
    g.set_profiling(True)
    db = [g.upload_batch(lid[b[0]].astype(np.int32), lid[b[1]].astype(np.int32), b[2])
          for b in blines]
    log(f"c3 partition: {n_own} owned, {ghosts.size} ghosts, {pu.size} edges "
        f"({time.perf_counter() - t:.1f}s)")
    for s in range(args.warmup):
        g.apply_device_batch(db[s])
        g.recolor(rule)
    g.synchronize()
    lat, kms, byts, launches = [], 0.0, 0, 0
    t0 = time.perf_counter()
    for s in range(args.warmup, nsteps):
        tb = time.perf_counter()
        g.apply_device_batch(db[s])
        st = g.recolor(rule)
        lat.append(time.perf_counter() - tb)
        kms += st.ms_recolor_kernel
        byts += st.recolor_bytes
        launches += st.recolor_launches
    g.synchronize()
    el = time.perf_counter() - t0
    mono, maxc = g.validate()
    value = args.steps * bs / el if mono == 0 else 0.0
    out = base_record(args, value, el, args.steps, args.warmup, {
        "workload": f"c3_rmat{scale}_partition0of4_{bs}mixed", "owned": n_own,
        "ghosts": int(ghosts.size), "edges": int(pu.size), "batch": bs, "rule": args.rule,
        "driver": args.driver, "monochromatic": int(mono)})
    out["latency_ms"] = {"p50": pct(np.array(lat) * 1e3, 50), "p99": pct(np.array(lat) * 1e3, 99)}
    out["roofline"] = roofline("k_recolor<M_LOOP>" if args.driver == "plain" else "k_fused",
                               byts, kms, launches)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5"], default="c2")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=int, default=100000)
    ap.add_argument("--rule", choices=["independent", "ldf"], default="independent")
    ap.add_argument("--driver", choices=["plain", "fused"], default="plain")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--graphs", type=int, default=512)
    ap.add_argument("--rates", type=float, nargs="+", default=[20, 80, 320])
    ap.add_argument("--duration", type=float, default=4.0)
    ap.add_argument("--budget-lines", type=int, default=4_000_000)
    ap.add_argument("--c3-scale", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--events-after", action="store_true",
                    help="time the count pass per launch in extra batches after the timed "
                         "ones instead of in the timed batches")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy 1 GiB five times (device to device) and exit: the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 1 << 28
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "copies": 5}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the recolor path has no multi-GPU "
              "collective in this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import dyncolor_amd as d
    rule = d.RULE_INDEPENDENT if args.rule == "independent" else d.RULE_LDF
    if d.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "c1":
        u, v = d.gen_er(10000, 50000, args.seed)
        if args.batch == 100000:
            args.batch = 2000
        out = run_stream(args, d, rule, 10000, u, v, f"c1_er10k_m50k_{args.batch}mixed")
    elif args.workload == "c2":
        t = time.perf_counter()
        u, v = d.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
        log(f"rmat scale {args.scale}: {u.size} edges ({time.perf_counter() - t:.1f}s)")
        out = run_stream(args, d, rule, 1 << args.scale, u, v,
                         f"rmat{args.scale}_ef{args.edge_factor}_{args.steps}x{args.batch}_mixed50")
    elif args.workload == "c3":
        if args.batch == 100000:
            args.batch = 1_000_000
        out = run_c3(args, d, rule)
    elif args.workload == "c4":
        out = run_c4(args, d, rule)
    else:
        out = run_c5(args, d, rule)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
