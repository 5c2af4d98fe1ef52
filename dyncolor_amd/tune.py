"""Autotuner of the recolour loop's work split (dc_tuning) and driver.

What it tunes: the driver (plain launches or captured HIP graphs per block
of rounds) and the dc_tuning fields: the grid of a loop round (work tiles per
round), the light-round threshold and grid, the worklist thresholds (how
large a frontier may be and still be chunked into lists instead of scanned),
the pull threshold and the size from which deletions are grouped by row.
Results never depend on them; only the time does.

How: coordinate descent from the defaults.  For each parameter in turn, each
candidate value is timed against the incumbent on the same live graph.  A
trial is one update batch followed by its inverse (the same lines with the
insert flags swapped, so the edge multiset and the work per trial stay put
while the colours move on); the graph is compacted after each trial, untimed,
so re-inserted edges do not pile up tombstones.  Each repetition takes
another batch: an untimed trial with the incumbent first (rows that outgrow
their capacity move then, not inside a timed trial), then the incumbent and
the candidate in alternating order (the one that runs second sees no row
growth on even and odd repetitions alike).  A candidate replaces the
incumbent when its median time per trial is lower by more than `margin`.
The choice and every measurement are saved as JSON keyed by device name and
workload; `load` returns the saved choice for this device, `apply` sets it
on a graph.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import DRIVER_GRAPH, DRIVER_PLAIN, device_info, tuning

CANDIDATES = {
    "driver": [DRIVER_PLAIN, DRIVER_GRAPH],
    "loop_blocks": [1024, 2048, 3072, 4096],
    "light_c": [1024, 4096, 16384, 65536],
    "light_blocks": [0, 128, 1024],
    "list_c": [16384, 65536, 131072, 524288],
    "list_slots": [1 << 20, 1 << 22, 1 << 24],
    "heavy_div": [4, 8, 16],
    "group_deletions": [1 << 14, 1 << 16, 1 << 18],
}


def _inverse(batch):
    u, v, ins = batch
    return u, v, (1 - np.asarray(ins, np.uint8)).astype(np.uint8)


def apply(g, params):
    """Set a tuned parameter dict (driver + dc_tuning fields) on a graph."""
    p = dict(params)
    drv = p.pop("driver", None)
    t = tuning()
    for k, v in p.items():
        setattr(t, k, v)
    g.set_tuning(t)
    if drv is not None:
        g.set_driver(drv)


def _trial(g, pair, rule):
    t0 = time.perf_counter()
    for b in pair:
        g.apply_device_batch(b)
        g.recolor(rule)
    g.synchronize()
    dt = time.perf_counter() - t0
    g.compact()
    return dt


def tune(g, batches, rule=0, reps=4, margin=0.03, params=None, log=None):
    """Coordinate descent on the live graph g over the given host batches
    (each used as a batch + its inverse).  Returns (best params, sweep)."""
    pairs = [(g.upload_batch(*b), g.upload_batch(*_inverse(b))) for b in batches]

    def run(p, j):
        apply(g, p)
        return _trial(g, pairs[j % len(pairs)], rule)

    state = {}
    try:
        return _descend(run, reps, margin, params, log, CANDIDATES, state)
    finally:
        if "best" in state:
            apply(g, state["best"])
        for a, b in pairs:
            a.close()
            b.close()


def _lanes_trial(arenas, pair, rule):
    from . import lanes_step
    t0 = time.perf_counter()
    for half in pair:
        lanes_step(arenas, half, rule)
    dt = time.perf_counter() - t0
    for a in arenas:
        a.compact()
    return dt


def tune_lanes(arenas, lane_batches, rule=0, reps=4, margin=0.03, params=None, log=None):
    """The same descent for arenas stepped together (dc_lanes_step): every
    arena takes the candidate, and a trial is one lanes step over each lane's
    batch, then one over the inverses, so the lanes contend for the device as
    they do in service.  lane_batches[l] is lane l's list of host batches (the
    same count per lane).  The driver is not tuned (lanes use the plain one)."""
    nb = min(len(b) for b in lane_batches)
    pairs = [([a.upload_batch(*lb[j]) for a, lb in zip(arenas, lane_batches)],
              [a.upload_batch(*_inverse(lb[j])) for a, lb in zip(arenas, lane_batches)])
             for j in range(nb)]

    def run(p, j):
        for a in arenas:
            apply(a, p)
        return _lanes_trial(arenas, pairs[j % len(pairs)], rule)

    cands = {k: v for k, v in CANDIDATES.items() if k != "driver"}
    state = {}
    try:
        return _descend(run, reps, margin, params, log, cands, state)
    finally:
        for a in arenas:
            if "best" in state:
                apply(a, state["best"])
        for fwd, inv in pairs:
            for b in fwd + inv:
                b.close()


def _descend(run, reps, margin, params, log, cands, state):
    """Coordinate descent over cands from the defaults; run(params, j) times
    trial j.  state["best"] always holds the current choice (the caller sets
    it on its graphs even when a trial fails)."""
    best = {"driver": DRIVER_PLAIN}
    best.update({k: getattr(tuning(), k) for k in CANDIDATES if k != "driver"})
    state["best"] = best
    sweep = []
    run(best, 0)   # warm caches and captures
    j = 0
    for name in (params or list(cands)):
        for val in cands[name]:
            if val == best[name]:
                continue
            cand = dict(best, **{name: val})
            inc_t, cand_t = [], []
            for rep in range(reps):
                run(best, j)   # untimed: row growth of this batch happens here
                if rep % 2 == 0:
                    inc_t.append(run(best, j))
                    cand_t.append(run(cand, j))
                else:
                    cand_t.append(run(cand, j))
                    inc_t.append(run(best, j))
                j += 1
            mi, mc = float(np.median(inc_t)), float(np.median(cand_t))
            rec = {"param": name, "value": val, "incumbent": best[name],
                   "incumbent_ms": mi * 1e3, "candidate_ms": mc * 1e3}
            if mc < mi * (1.0 - margin):
                best = cand
                state["best"] = best
                rec["taken"] = True
            sweep.append(rec)
            if log:
                log(f"tune {name}={val}: {mc * 1e3:.2f} ms vs {mi * 1e3:.2f} ms"
                    f"{' (taken)' if rec.get('taken') else ''}")
    return best, sweep


def _key():
    name, arch, cus = device_info(0)
    return f"{name}|{arch}|{cus}"


def save(path, workload, best, sweep, extra=None):
    """Store the choice for this device and workload (other entries kept)."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data.setdefault(_key(), {})[workload] = dict(best=best, sweep=sweep, **(extra or {}))
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)


def load(path, workload):
    """The saved choice for this device and workload, or None."""
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    ent = data.get(_key(), {}).get(workload)
    return None if ent is None else ent["best"]
