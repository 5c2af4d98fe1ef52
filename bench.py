"""Throughput benchmark of the dynamic-recoloring hot path on one MI355X.

Workload (default): R-MAT scale 22 (edge factor 16, a/b/c = .57/.19/.19,
self loops and duplicates removed, IDs not permuted), greedy static coloring,
then consecutive update batches of 100k lines (50% insertions of uniform
non-edges, 50% deletions of existing edges), each applied to the graph left by
the previous one.  A step = apply one device-resident batch + run every
recolor round until the coloring is final.  The first --warmup batches are
untimed; the next --steps batches are timed.

Prints one JSON line with updates/s, batch latency percentiles, the recolor
kernel's HBM roofline fraction and the CPU-oracle baseline.
"""
from __future__ import annotations

import argparse
import json
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=int, default=100000)
    ap.add_argument("--rule", choices=["independent", "ldf"], default="independent")
    ap.add_argument("--driver", choices=["plain", "fused"], default="plain")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the recolor path has no multi-GPU "
              "collective in this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import dyncolor_amd as d
    rule = d.RULE_INDEPENDENT if args.rule == "independent" else d.RULE_LDF
    if d.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    t = time.perf_counter()
    u, v = d.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    n = 1 << args.scale
    log(f"rmat scale {args.scale}: {u.size} edges ({time.perf_counter() - t:.1f}s)")
    nb = args.warmup + args.steps
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
    g.set_profiling(True)
    g.set_driver(d.DRIVER_FUSED if args.driver == "fused" else d.DRIVER_PLAIN,
                 args.blocks_per_cu)

    colors_after_first = None
    for i in range(args.warmup):
        g.apply_device_batch(dev_batches[i])
        g.recolor(rule)
        if i == 0:
            colors_after_first = g.colors()
    g.synchronize()
    log("warmup done")

    lat = []
    tot = {"recolor_bytes": 0, "ms_recolor_kernel": 0.0, "launches": 0, "rounds": 0,
           "changed": 0, "evaluated": 0, "slots": 0}
    t0 = time.perf_counter()
    for i in range(args.warmup, nb):
        tb = time.perf_counter()
        g.apply_device_batch(dev_batches[i])
        st = g.recolor(rule)
        lat.append(time.perf_counter() - tb)
        tot["recolor_bytes"] += st.recolor_bytes
        tot["ms_recolor_kernel"] += st.ms_recolor_kernel
        tot["launches"] += st.recolor_launches
        tot["rounds"] += st.loop_rounds
        tot["changed"] += st.changed
        tot["evaluated"] += st.evaluated
        tot["slots"] += st.slots_scanned
    g.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    mono, maxc = g.validate()
    if mono != 0:
        log(f"INVALID: {mono} monochromatic edges after the timed batches")
    log(f"timed {args.steps} batches in {elapsed:.3f}s, rounds {tot['rounds']}, "
        f"colors {maxc + 1}")

    updates = args.steps * args.batch
    value = updates / elapsed if mono == 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    kern_s = tot["ms_recolor_kernel"] / 1e3
    achieved = tot["recolor_bytes"] / kern_s / 1e9 if kern_s > 0 else 0.0

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, u, v, batches[0], rule, colors_after_first, args)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_ef{args.edge_factor}_{args.steps}x{args.batch}_mixed50",
            "scale": args.scale, "edge_factor": args.edge_factor, "edges": int(u.size),
            "batch": args.batch, "rule": args.rule, "driver": args.driver,
            "colors": int(maxc + 1), "monochromatic": int(mono),
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99)),
                       "max": float(lat_ms.max())},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_recolor<M_LOOP>" if args.driver == "plain" else "k_fused",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches": tot["launches"],
            "avg_launch_us": tot["ms_recolor_kernel"] * 1e3 / max(tot["launches"], 1),
            "algorithmic_bytes": tot["recolor_bytes"],
        },
        "cpu_baseline": cpu,
        "work": {"loop_rounds": tot["rounds"], "evaluated": tot["evaluated"],
                 "changed": tot["changed"], "row_slots_scanned": tot["slots"]},
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(n, u, v, batch, rule, gpu_colors_after_first, args):
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
    return {
        "value": bu.size / dt,
        "unit": "updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first batch ({bu.size} lines) of the same stream from the same initial "
                  f"graph, single-thread oracle replay (apply + all rounds)",
        "seconds": dt,
        "bit_equal_to_gpu": same,
    }


if __name__ == "__main__":
    main()
