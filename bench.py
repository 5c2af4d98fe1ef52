"""Throughput benchmark of the dynamic-recoloring hot path on one MI355X.

Default run: the metric's single-GPU configuration (R-MAT scale 22, edge
factor 16, a/b/c = .57/.19/.19, self loops and duplicates removed, IDs not
permuted; greedy static colouring, then consecutive update batches of 100k
lines, 50% insertions of uniform non-edges, 50% deletions of existing edges,
each applied to the graph left by the previous one).  A step = apply one
device-resident batch + run every recolour round until the colouring is
final.  The first --warmup batches are untimed; the next --steps are timed.
The same JSON line carries two compact sub-records: "batched" (512 graphs
recoloured per step in one arena against one plain call per graph) and
"served" (8 resident R-MAT graphs under a Poisson request stream, continuous
batching against one plain call per request); vs_baseline is the batched
speedup over the plain per-batch path (the metric's "vs plain per-batch").

Other workloads alone (--workload): c1 ER 10k/50k, c3 one partition of an
R-MAT scale-24 graph (ingested from global IDs), c4 batched, c5 served.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import platform
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


def base_record(value, elapsed_steps, steps, warmup, config):
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


def traffic_for(kernel, key="bytes_per_launch"):
    """HBM bytes per launch of the kernel (or per executed loop round: key
    "bytes_per_round") from the committed rocprofv3 --pmc summary (FETCH_SIZE
    + WRITE_SIZE, scaled by the copy-kernel factor); kernel may be a list,
    whose figures are summed.  None when the summary does not cover them."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    tot = 0
    for k in ([kernel] if isinstance(kernel, str) else kernel):
        ent = t.get(k.split(" ")[0])
        if ent is None or ent.get(key) is None:
            return None
        tot += ent[key]
    return tot


def roofline(kernel, algo_bytes, kernel_ms, launches, traffic_kernels=None, traffic_key="bytes_per_launch"):
    """Roofline of one kernel: achieved = its algorithmic bytes per launch
    over its mean launch duration (hipEvents on its stream); traffic = the
    measured HBM bytes per launch, both per launch."""
    per_launch = algo_bytes / max(launches, 1)
    avg_s = kernel_ms / 1e3 / max(launches, 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = traffic_for(traffic_kernels or kernel, traffic_key)
    out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "basis": "achieved and frac: the kernel's algorithmic bytes per launch (row slots "
                    "streamed, rows visited, state and list words) / mean launch time; "
                    "traffic: rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch",
           "algorithmic_bytes_per_launch": per_launch, "launches": launches,
           "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_total": int(algo_bytes)}
    if traffic:
        out["traffic_over_algorithmic"] = traffic / per_launch if per_launch else None
        out["achieved_by_traffic"] = traffic / avg_s / 1e9 if avg_s > 0 else None
    return out


TUNING_PATH = os.path.join(ROOT, "profiles", "tuning.json")


def set_driver(g, d, args, workload=None):
    """The requested driver, or the tuned parameters saved for this device
    and workload (driver included) when --driver is not given."""
    tuned = getattr(args, "tuned", {}).get(workload) if workload else None
    if tuned and args.driver is None:
        from dyncolor_amd import tune
        tune.apply(g, tuned)
        return
    drv = {"fused": d.DRIVER_FUSED, "graph": d.DRIVER_GRAPH}.get(args.driver, d.DRIVER_PLAIN)
    g.set_driver(drv, args.blocks_per_cu)


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count(), model


# ----------------------------------------------------------------- c1 / c2 --
def run_stream(args, d, rule, n, u, v, workload):
    """One graph, a stream of device-resident batches (c1, c2)."""
    # per-launch hipEvents on the count pass cost ~7 % when recorded in the
    # timed batches: by default they are taken in an untimed replay of the
    # warm-up and timed batches on a fresh copy of the initial graph
    # (deterministic: the same rounds and launches); with --events-after in 2
    # batches after the timed ones
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
    set_driver(g, d, args, args.stream_key)
    prof = {"ms": 0.0, "ams": 0.0, "bytes": 0, "launches": 0, "kms": 0.0, "rbytes": 0, "rl": 0}
    terms = {"updates": 0, "del_row_slots": 0, "lc_vertices": 0, "lc_slots": 0, "loop_rounds": 0,
             "sum_C": 0, "batches": 0, "kernel_ms": 0.0}

    def add_prof(st, gr):
        prof["ms"] += st.ms_count_pass
        prof["ams"] += st.ms_round_a
        prof["bytes"] += st.count_pass_bytes
        prof["launches"] += st.count_pass_launches
        prof["kms"] += st.ms_recolor_kernel
        prof["rbytes"] += st.recolor_bytes
        prof["rl"] += st.recolor_launches
        tr = gr.trace()
        terms["updates"] += st.updates
        terms["del_row_slots"] += st.del_row_slots
        terms["lc_vertices"] += st.lc_vertices
        terms["lc_slots"] += st.lc_slots
        terms["loop_rounds"] += int((tr[:, 0] == 10).sum()) if tr.size else 0
        terms["sum_C"] += int(tr[tr[:, 0] == 10][:, 2].sum()) if tr.size else 0
        terms["batches"] += 1
        terms["kernel_ms"] += st.ms_recolor_kernel

    if args.tune:
        from dyncolor_amd import tune
        for i in range(args.warmup):
            g.apply_device_batch(dev_batches[i])
            g.recolor(rule)
        tb = d.gen_batches(n, u, v, 4, args.batch, 0.5, args.seed + 2000)
        best, sweep = tune.tune(g, tb, rule, reps=args.tune_reps, log=log)
        tune.save(args.tuning_path, args.stream_key, best, sweep, {"batch": args.batch})
        log(f"{args.stream_key} tuned: {best}")
        return {"tuned": best, "sweep": sweep}
    colors_after_first = None
    for i in range(args.warmup):
        g.apply_device_batch(dev_batches[i])
        g.recolor(rule)
        if i == 0:
            colors_after_first = g.colors()
    g.synchronize()

    lat = []
    tot = {"rounds": 0, "changed": 0, "evaluated": 0, "slots": 0, "regrows": 0, "settled": 0}
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
            add_prof(g.recolor(rule), g)
        g.set_profiling(False)
        g.synchronize()
    else:
        colors_timed = g.colors()
        del dev_batches, g
        gc.collect()
        g = d.Graph(n, u, v)
        g.static_color(rule)
        set_driver(g, d, args, args.stream_key)
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
                add_prof(st, g)
                replay_rounds += st.loop_rounds
        g.set_profiling(False)
        g.synchronize()
        if replay_rounds != tot["rounds"] or not (g.colors() == colors_timed).all():
            log(f"replay differs: {replay_rounds} rounds against {tot['rounds']}")
            replay_rounds = None
    del dev_batches, g
    gc.collect()
    log(f"timed {args.steps} batches in {elapsed:.3f}s, rounds {tot['rounds']}, colors {maxc + 1}")
    value = args.steps * args.batch / elapsed if mono == 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    out = base_record(value, elapsed, args.steps, args.warmup, {
        "workload": workload, "edges": int(u.size), "vertices": int(n), "batch": args.batch,
        "rule": args.rule, "driver": args.driver or "plain", "colors": int(maxc + 1),
        "monochromatic": int(mono), "tuning": args.tuned.get(args.stream_key) or "default"})
    out["latency_ms"] = {"p50": pct(lat_ms, 50), "p99": pct(lat_ms, 99), "max": float(lat_ms.max())}
    if args.driver != "fused" and rule == d.RULE_INDEPENDENT and prof["launches"]:
        # the unit is one loop round: phase A + count pass, which together do
        # the reference's findEligibleNeighbors + recolorNeighbor pair of the
        # round; its algorithmic bytes are the reference's for that round
        rb = (16 * terms["lc_vertices"] + 8 * terms["lc_slots"] + terms["loop_rounds"] * (n / 8.0)
              + 4 * terms["sum_C"])
        out["roofline"] = roofline("loop round (k_round_a + k_round_c)", rb,
                                   prof["ams"] + prof["ms"], prof["launches"],
                                   ["k_round_a", "k_round_c", "k_pull_fix"], "bytes_per_round")
        out["roofline"]["basis"] = (
            "achieved and frac: the reference's bytes per loop round (16 per vertex of L_r u C_r: "
            "list entry, row bounds, colour and prev; 8 per row slot of L_r u C_r: neighbour and its "
            "colour; n/8 flag compaction; 4 per member of C_r) / mean round time (k_round_a + "
            "k_round_c, and k_pull_fix after the count pass of rounds 0-3; hipEvents attached to "
            "each launch); traffic: rocprofv3 FETCH_SIZE + WRITE_SIZE of those kernels per "
            "executed round")
        out["roofline"]["rounds_by_terms"] = terms["loop_rounds"]
        out["roofline"]["round_us"] = {"k_round_a": prof["ams"] * 1e3 / prof["launches"],
                                       "k_round_c": prof["ms"] * 1e3 / prof["launches"]}
        # the count pass alone against this build's own bytes (row slots it
        # streams, rows visited, state and list words)
        out["roofline"]["count_pass_own_model"] = roofline(
            "k_round_c (count pass)", prof["bytes"], prof["ms"], prof["launches"])
    elif args.driver != "fused" and prof["launches"]:
        # ldf rule: the loop's rounds as the unit (four launches each); bytes are
        # the reference's for those rounds (it rescans every row of L_r u C_r)
        out["roofline"] = roofline("ldf loop round", prof["bytes"], prof["ms"], prof["launches"],
                                   ["k_ldf_members", "k_ldf_long_mex", "k_ldf_rows", "k_ldf_commit"],
                                   "bytes_per_round")
        out["roofline"]["basis"] = (
            "achieved and frac: the reference's bytes per loop round (16 per vertex of L_r u C_r, "
            "12 per row slot of L_r u C_r: neighbour, colour, priority; n/8 bitmap; 4 per member "
            "of C_r) / mean round time (k_ldf_members + k_ldf_long_mex + k_ldf_rows + "
            "k_ldf_commit, hipEvents around each block of rounds); traffic: rocprofv3 FETCH_SIZE + "
            "WRITE_SIZE of those kernels per executed round")
    else:
        out["roofline"] = roofline("k_loop_persistent" if args.driver == "fused" else
                                   "k_recolor", prof["rbytes"], prof["kms"], prof["rl"])
    out["roofline"]["measured_on"] = (
        f"{nprof} batches after the timed ones" if args.events_after else
        f"replay of the warm-up and timed batches (same {replay_rounds} rounds)")
    out["reference_model"] = reference_model(terms, n, rule == d.RULE_LDF)
    out["work"] = {"loop_rounds": tot["rounds"], "evaluated": tot["evaluated"],
                   "changed": tot["changed"], "row_slots_scanned": tot["slots"],
                   "batches_with_row_growth": tot["regrows"],
                   "batches_settled": tot["settled"]}
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, n, u, v, batches[0], rule, colors_after_first)
    return out


def reference_model(t, n, ldf=False):
    """Bytes the reference's algorithm moves for the same batches (per batch):
    12 per update record + 8 per adjacency write or tombstone + 4 per row slot
    of both endpoints per deletion (its linear deletion search) + per loop round
    16 + 8 per row slot for every vertex of L_r u C_r (it rescans each of their
    rows; 12 per slot for the ldf rule, whose scans also gather priorities) + a
    bitmap over the vertices and 4 per member of C_r (compaction), counted over
    the rounds this build ran (rounds after settled colours are not counted:
    the figure is a lower bound).  achieved = those bytes over the summed
    recolour-kernel time of the same batches."""
    b = max(t["batches"], 1)
    per_slot = 12 if ldf else 8
    B = (20 * t["updates"] + 4 * t["del_row_slots"] + 16 * t["lc_vertices"]
         + per_slot * t["lc_slots"] + t["loop_rounds"] * (n / 8.0) + 4 * t["sum_C"])
    s = t["kernel_ms"] / 1e3
    ach = B / s / 1e9 if s > 0 else None
    return {"terms_per_batch": {k: t[k] / b for k in ("updates", "del_row_slots", "lc_vertices",
                                                     "lc_slots", "loop_rounds", "sum_C")},
            "bytes_per_batch": B / b, "recolor_kernel_ms_per_batch": t["kernel_ms"] / b,
            "achieved_GBs": ach, "frac_of_peak": ach / HBM_PEAK_GBS if ach else None,
            "formula": f"20*U + 4*del_row_slots + 16*|L u C| + {per_slot}*slots(L u C) + rounds*n/8 + 4*sum|C|"}


def cpu_baseline(args, n, u, v, batch, rule, gpu_colors_after_first):
    """The oracle replaying the first batch of the same stream on this host:
    single-threaded, and with each round's work split over OpenMP threads
    (independent rule); both checked bit-equal to the GPU's colours."""
    from oracle import oracle as O
    try:
        O.lib()
    except Exception as e:  # oracle not built
        log(f"cpu baseline skipped: {e}")
        return None
    nproc, model = cpu_info()
    threads = max(1, min(args.cpu_threads, nproc or 1))
    bu, bv, bi = batch
    res = {}
    for th in ([1, threads] if rule == 0 and threads > 1 else [1]):
        t = time.perf_counter()
        og = O.OracleGraph(n, u, v)
        og.static_color(rule)
        log(f"oracle graph + static coloring: {time.perf_counter() - t:.1f}s")
        t = time.perf_counter()
        og.apply_batch(bu, bv, bi)
        if th == 1:
            og.recolor(rule)
        else:
            og.recolor_par(rule, th)
        dt = time.perf_counter() - t
        same = None
        if gpu_colors_after_first is not None:
            same = bool((og.color == gpu_colors_after_first).all())
        log(f"oracle batch, {th} thread(s): {dt:.1f}s, colors equal to GPU: {same}")
        res[th] = (dt, same)
        del og
    th = max(res)
    dt, same = res[th]
    out = {"value": bu.size / dt, "unit": "updates/s", "cores": th, "kind": "port",
           "sample": f"first batch ({bu.size} lines) of the same stream from the same initial "
                     f"graph, oracle replay (apply + all rounds) on {th} thread(s)",
           "seconds": dt, "bit_equal_to_gpu": same, "nproc": nproc, "cpu_model": model,
           "threads": th}
    if th != 1:
        out["single_thread"] = {"value": bu.size / res[1][0], "seconds": res[1][0],
                                "bit_equal_to_gpu": res[1][1]}
    return out


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


def run_c4(args, d, rule, steps, warmup):
    t = time.perf_counter()
    graphs, kinds, rng = c4_graphs(d, args.graphs, args.seed)
    nsteps = warmup + steps
    per = []
    for i, (n, u, v) in enumerate(graphs):
        frac = rng.uniform(0.01, 0.10)
        bs = max(2, int(u.size * frac))
        per.append(d.gen_batches(n, u, v, nsteps, bs, 0.5, args.seed * 7919 + i))
    tot_edges = sum(g[2].size for g in graphs)
    step_lines = [sum(p[s][0].size for p in per) for s in range(nsteps)]
    log(f"c4: {len(graphs)} graphs, {sum(g[0] for g in graphs)} vertices, {tot_edges} edges, "
        f"{step_lines[-1]} lines per step ({time.perf_counter() - t:.1f}s)")

    # batched: the graphs in K arenas (lanes) balanced by edge count; a step
    # applies every lane's batch and recolours it, the lanes concurrently
    # (dc_lanes_step: one host thread and stream per lane)
    K = max(1, min(args.c4_lanes, len(graphs)))
    edges_of = np.array([g[1].size for g in graphs])
    lane_of = np.zeros(len(graphs), np.int32)
    load = np.zeros(K)
    for i in np.argsort(-edges_of, kind="stable"):
        k = int(np.argmin(load))
        lane_of[i] = k
        load[k] += edges_of[i]
    lanes = []
    n = 0
    for k in range(K):
        idx = np.nonzero(lane_of == k)[0]
        nk, U, V, off = d.union_edges([graphs[i] for i in idx])
        n += nk
        a = d.Graph(nk, U, V)
        a.static_color(rule)
        set_driver(a, d, args, "c4")
        ub = [] if args.tune else [a.upload_batch(*d.union_batch([per[i][s] for i in idx], off))
                                   for s in range(nsteps)]
        lanes.append((idx, off, a, ub))
        del U, V
    arenas = [l[2] for l in lanes]

    if args.tune:
        # the work split is tuned on the lanes stepped together, as they run
        from dyncolor_amd import tune
        first = [a.upload_batch(*d.union_batch([per[i][0] for i in idx], off))
                 for idx, off, a, _ in lanes]
        d.lanes_step(arenas, first, rule)
        lane_batches = [[d.union_batch([per[i][s] for i in idx], off) for s in range(1, min(3, nsteps))]
                        for idx, off, _, _ in lanes]
        best, sweep = tune.tune_lanes(arenas, lane_batches, rule, reps=args.tune_reps, log=log)
        tune.save(args.tuning_path, "c4", best, sweep, {"graphs": len(graphs), "lanes": K,
                                                        "lines_per_step": int(step_lines[-1])})
        log(f"c4 tuned: {best}")
        return {"tuned": best, "sweep": sweep}
    for s in range(warmup):
        d.lanes_step(arenas, [l[3][s] for l in lanes], rule)
    lat, rounds = [], 0
    t0 = time.perf_counter()
    for s in range(warmup, nsteps):
        tb = time.perf_counter()
        sts = d.lanes_step(arenas, [l[3][s] for l in lanes], rule)
        lat.append(time.perf_counter() - tb)
        rounds += max(st.loop_rounds for st in sts)
    t_batched = time.perf_counter() - t0
    mono = sum(a.validate()[0] for a in arenas)
    got = [None] * len(graphs)
    for idx, off, a, _ in lanes:
        col = a.colors()
        for j, i in enumerate(idx):
            got[i] = col[off[j]:off[j + 1]]
    del lanes, arenas
    gc.collect()
    lines = sum(step_lines[warmup:])
    batched = lines / t_batched if mono == 0 else 0.0
    log(f"c4 batched ({K} lanes): {t_batched:.3f}s for {steps} steps, rounds {rounds}, mono {mono}")
    rec = {
        "workload": f"c4_{len(graphs)}graphs_lognormal20k_er_rmat_1to10pct",
        "value": batched, "unit": "updates/s", "steps": steps, "warmup": warmup, "lanes": K,
        "ms_per_step": t_batched / steps * 1e3, "p50_ms": pct(np.array(lat) * 1e3, 50),
        "p99_ms": pct(np.array(lat) * 1e3, 99),
        "graphs": len(graphs), "vertices": int(n), "edges": int(tot_edges),
        "lines_per_step": int(step_lines[-1]), "loop_rounds": rounds, "monochromatic": int(mono),
        "tuning": args.tuned.get("c4") or "default"}
    if args.no_plain:   # profiling the arena alone
        return rec

    # plain per-batch baseline: one graph handle and one plain call per graph
    t_plain = 0.0
    handles = []
    for i, (gn, gu, gv) in enumerate(graphs):
        g = d.Graph(gn, gu, gv)
        g.static_color(rule)
        set_driver(g, d, args, "c4")
        handles.append((g, [g.upload_batch(*per[i][s]) for s in range(nsteps)]))
    for s in range(nsteps):
        tb = time.perf_counter()
        for g, bl in handles:
            g.apply_device_batch(bl[s])
            g.recolor(rule)
        if s >= warmup:
            t_plain += time.perf_counter() - tb
    same = all((got[i] == h[0].colors()).all() for i, h in enumerate(handles))
    del handles
    gc.collect()
    plain = lines / t_plain
    log(f"c4 plain: {t_plain:.3f}s, speedup {batched / plain:.1f}x, colors equal: {same}")
    rec.update({"plain_value": plain, "plain_seconds": t_plain, "plain_calls_per_step": len(graphs),
                "speedup_vs_plain": batched / plain, "bit_equal_to_plain": bool(same)})
    return rec


# ---------------------------------------------------------------------- c5 --
def c5_requests(graphs, rate, duration, max_req, rng):
    nreq = max(1, min(int(rng.poisson(rate * duration)), max_req))
    arrival = np.sort(rng.uniform(0, nreq / rate, nreq))
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


def lane_split(graphs, nlanes):
    """Greedy: largest graph first into the lane with the fewest edges."""
    order = sorted(range(len(graphs)), key=lambda g: -graphs[g][1].size)
    load = [0] * nlanes
    lane_of = [0] * len(graphs)
    for g in order:
        l = int(np.argmin(load))
        lane_of[g] = l
        load[l] += graphs[g][1].size
    return lane_of


def run_c5(args, d, rule):
    scales = [18, 19, 20, 21, 22, 23, 18, 20]
    t = time.perf_counter()
    graphs = []
    for i, s in enumerate(scales):
        u, v = d.gen_rmat(s, 16, seed=args.seed + 31 * i)
        graphs.append((1 << s, u, v))
    nl = max(1, min(args.lanes, len(graphs)))
    lane_of = lane_split(graphs, nl)
    arenas, first, size = [], [0] * len(graphs), [0] * len(graphs)
    for l in range(nl):
        mine = [g for g in range(len(graphs)) if lane_of[g] == l]
        n, U, V, off = d.union_edges([graphs[g] for g in mine])
        a = d.Graph(n, U, V)
        a.static_color(rule)
        set_driver(a, d, args)
        arenas.append(a)
        for k, g in enumerate(mine):
            first[g] = int(off[k])
            size[g] = int(off[k + 1] - off[k])
    tot_edges = sum(g[1].size for g in graphs)
    log(f"c5: 8 resident graphs in {nl} lanes, {tot_edges} edges ({time.perf_counter() - t:.1f}s)")
    rng = np.random.default_rng(args.seed)
    budget = args.budget_mb << 20
    srv = d.Server(arenas, lane_of, first, size, budget_bytes=budget, rule=rule)
    sweep = []
    for rate in args.rates:
        gid, roff, u, v, ins, arrival = c5_requests(graphs, rate, args.duration, args.max_requests, rng)
        lat, step, st = srv.run(gid, roff, u, v, ins, arrival)
        ok = lat >= 0
        rec = {"rate_req_s": rate, "requests": int(gid.size), "served": int(st.served),
               "rejected": int(st.rejected), "lines": int(st.lines), "steps": int(st.steps),
               "elapsed_s": st.elapsed_s, "offered_s": float(arrival[-1]),
               "updates_per_s": st.lines / st.elapsed_s,
               "p50_ms": pct(lat[ok] * 1e3, 50), "p99_ms": pct(lat[ok] * 1e3, 99),
               "max_step_lines": int(st.max_step_lines), "admit_waits": int(st.admit_waits),
               "max_inflight_MB": st.max_inflight_bytes / 2**20}
        log(f"c5 rate {rate}/s: {rec}")
        sweep.append(rec)
    mono = sum(a.validate()[0] for a in arenas)
    srv.close()
    del srv, arenas
    gc.collect()
    # plain per-batch capacity: a sample of requests, one plain call each on
    # its own graph handle, no coalescing
    rng2 = np.random.default_rng(args.seed + 1)
    handles = []
    for gn, gu, gv in graphs:
        g = d.Graph(gn, gu, gv)
        g.static_color(rule)
        handles.append(g)
    gid, roff, u, v, ins, _ = c5_requests(graphs, 40, 1.0, 200, rng2)
    t0 = time.perf_counter()
    plat = []
    for r in range(gid.size):
        tb = time.perf_counter()
        lo, hi = roff[r], roff[r + 1]
        handles[gid[r]].process_batch(u[lo:hi], v[lo:hi], ins[lo:hi], rule)
        plat.append(time.perf_counter() - tb)
    t_plain = time.perf_counter() - t0
    del handles
    gc.collect()
    plain = float(roff[-1]) / t_plain
    best = max(sweep, key=lambda r: r["updates_per_s"])
    return {
        "workload": "c5_8rmat_s18-23_poisson_lognormal5k_60ins",
        "value": best["updates_per_s"], "unit": "updates/s", "capacity_rate_req_s": best["rate_req_s"],
        "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "lanes": nl,
        "budget_MB": args.budget_mb, "edges": int(tot_edges), "monochromatic": int(mono),
        "sweep": sweep,
        "plain_value": plain, "plain_requests": int(gid.size), "plain_seconds": t_plain,
        "plain_p50_ms": pct(np.array(plat) * 1e3, 50), "plain_p99_ms": pct(np.array(plat) * 1e3, 99),
        "speedup_vs_plain": best["updates_per_s"] / plain}


# ---------------------------------------------------------------------- c3 --
def run_c3(args, d, rule):
    """Partition 0 of 4 of an R-MAT scale-24 graph: the rank's share built by
    the global-ID ingest (dc_partition_plan), ghost colours frozen at the
    whole graph's static colouring; the first batch is checked against the
    oracle."""
    scale = args.c3_scale
    t = time.perf_counter()
    u, v = d.gen_rmat(scale, 16, seed=args.seed)
    N = 1 << scale
    n_own = N // 4
    log(f"c3: rmat {scale}: {u.size} edges ({time.perf_counter() - t:.1f}s)")
    full = d.Graph(N, u, v)
    full.static_color(rule)
    colors = full.colors()
    full.close()
    keep = (u < n_own) | (v < n_own)
    pu, pv = u[keep], v[keep]
    rng = np.random.default_rng(args.seed)
    nsteps = args.warmup + args.steps
    bs = args.batch
    # batches: half insertions (owned endpoint + uniform vertex), half
    # deletions of existing edges touching the partition
    cu, cv, ci = [], [], []
    for s in range(nsteps):
        k_ins = bs // 2
        a = rng.integers(0, n_own, k_ins)
        b = (a + rng.integers(1, N, k_ins)) % N
        e = rng.integers(0, pu.size, bs - k_ins)
        cu.append(np.concatenate([a, pu[e]]).astype(np.int32))
        cv.append(np.concatenate([b, pv[e]]).astype(np.int32))
        ci.append(np.concatenate([np.ones(k_ins, np.uint8), np.zeros(bs - k_ins, np.uint8)]))
    part = (np.arange(N, dtype=np.int64) // n_own).astype(np.int32)
    t = time.perf_counter()
    plan = d.PartitionPlan(N, 0, np.arange(n_own, dtype=np.int32), part, u, v,
                           np.concatenate(cu), np.concatenate(cv), np.concatenate(ci))
    l2g, lu, lv, bu, bv, bi = plan.arrays()
    info = plan.info
    g = plan.create_graph(colors)
    log(f"c3 partition (global-ID ingest {time.perf_counter() - t:.1f}s): {info['n_owned']} owned, "
        f"{info['n_ghost']} ghosts, {info['m_local']} edges")
    set_driver(g, d, args)
    db = [g.upload_batch(bu[s * bs:(s + 1) * bs], bv[s * bs:(s + 1) * bs], bi[s * bs:(s + 1) * bs])
          for s in range(nsteps)]
    first_colors = None
    for s in range(args.warmup):
        g.apply_device_batch(db[s])
        g.recolor(rule)
        if s == 0:
            first_colors = g.colors()
    g.synchronize()
    lat = []
    t0 = time.perf_counter()
    for s in range(args.warmup, nsteps):
        tb = time.perf_counter()
        g.apply_device_batch(db[s])
        g.recolor(rule)
        lat.append(time.perf_counter() - tb)
    g.synchronize()
    el = time.perf_counter() - t0
    mono, maxc = g.validate()
    value = args.steps * bs / el if mono == 0 else 0.0
    out = base_record(value, el, args.steps, args.warmup, {
        "workload": f"c3_rmat{scale}_partition0of4_{bs}mixed", "owned": int(info["n_owned"]),
        "ghosts": int(info["n_ghost"]), "edges": int(info["m_local"]), "batch": bs,
        "rule": args.rule, "driver": args.driver, "monochromatic": int(mono)})
    out["latency_ms"] = {"p50": pct(np.array(lat) * 1e3, 50), "p99": pct(np.array(lat) * 1e3, 99)}
    out["border_vertices_per_part"] = plan.border().tolist()
    if first_colors is not None and not args.no_cpu_baseline:
        from oracle import oracle as O
        t = time.perf_counter()
        og = O.OracleGraph(l2g.size, lu, lv, n_owned=info["n_owned"])
        og.set_colors(colors[l2g])
        og.apply_batch(bu[:bs], bv[:bs], bi[:bs])
        og.recolor(rule)
        same = bool((og.color == first_colors).all())
        log(f"c3 oracle first batch: {time.perf_counter() - t:.1f}s, colors equal: {same}")
        out["bit_equal_to_oracle_first_batch"] = same
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
    ap.add_argument("--driver", choices=["plain", "fused", "graph"], default=None,
                    help="recolour driver (default: the tuned one, else plain)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--graphs", type=int, default=512)
    ap.add_argument("--c4-steps", type=int, default=3, help="timed c4 steps of the default run")
    ap.add_argument("--c4-lanes", type=int, default=4,
                    help="c4: arenas the graphs are split over (stepped concurrently)")
    ap.add_argument("--no-plain", action="store_true",
                    help="c4: time the arena only (no plain per-graph leg), for profiling")
    ap.add_argument("--rates", type=float, nargs="+", default=[80, 320, 1280, 5120])
    ap.add_argument("--duration", type=float, default=1.5)
    ap.add_argument("--max-requests", type=int, default=6000)
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--budget-mb", type=int, default=512)
    ap.add_argument("--c3-scale", type=int, default=24)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on the workload (c2: the batch stream; c4: the "
                         "arena) and save its choice and sweep to profiles/tuning.json")
    ap.add_argument("--tune-reps", type=int, default=4)
    ap.add_argument("--tuning-path", default=TUNING_PATH,
                    help="where tuned parameters are read from (and written by --tune)")
    ap.add_argument("--no-tuning", action="store_true", help="ignore profiles/tuning.json")
    ap.add_argument("--no-batched", action="store_true", help="default run: skip the c4 sub-record")
    ap.add_argument("--no-served", action="store_true", help="default run: skip the c5 sub-record")
    ap.add_argument("--events-after", action="store_true",
                    help="time the count pass per launch in extra batches after the timed "
                         "ones instead of in an untimed replay")
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
    args.tuned = {}
    if not args.no_tuning and not args.tune and d.device_count() > 0:
        from dyncolor_amd import tune
        for w in ("c2", "c4"):
            t = tune.load(args.tuning_path, w)
            if t:
                args.tuned[w] = t
                log(f"tuned parameters for {w}: {t}")
    args.stream_key = "c2" if args.workload == "c2" and args.scale == 22 and args.batch == 100000 \
        and rule == d.RULE_INDEPENDENT else "none"
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
        del u, v
        gc.collect()
        if args.tune:
            out = {"c2": out, "c4": run_c4(args, d, rule, 2, 1)}
            print(json.dumps(out), flush=True)
            return
        if not args.no_batched:
            b = run_c4(args, d, rule, args.c4_steps, 1)
            out["batched"] = b
            out["vs_baseline"] = b.get("speedup_vs_plain")
            out["vs_baseline_basis"] = (f"batched (512 graphs per step in {b['lanes']} arenas stepped "
                                        "concurrently) updates/s over the plain per-batch path (one "
                                        "call per graph), same inputs")
        if not args.no_served:
            out["served"] = run_c5(args, d, rule)
    elif args.workload == "c3":
        if args.batch == 100000:
            args.batch = 1_000_000
        out = run_c3(args, d, rule)
    elif args.workload == "c4":
        b = run_c4(args, d, rule, args.steps, args.warmup)
        if args.tune:
            print(json.dumps(b), flush=True)
            return
        out = base_record(b["value"], b["ms_per_step"] / 1e3 * args.steps, args.steps, args.warmup,
                          {"workload": b["workload"], "graphs": b["graphs"], "rule": args.rule,
                           "driver": args.driver})
        out["batched"] = b
        out["vs_baseline"] = b.get("speedup_vs_plain")
    else:
        s = run_c5(args, d, rule)
        out = base_record(s["value"], 0.0, 0, 0, {"workload": s["workload"], "rule": args.rule,
                                                  "driver": args.driver})
        out["served"] = s
        out["vs_baseline"] = s["speedup_vs_plain"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
