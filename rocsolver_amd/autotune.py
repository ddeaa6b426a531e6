"""Autotuner for the schedule of the variable-size batched routines.

The persistent kernel of rocsolver_amd_{s,d}{getrf,getrs,gesv}_vbatched runs
small matrices (n <= 64) in teams of 16, 32 or 64 lanes (several matrices per
wavefront), hands them out in packs, and runs a number of persistent
workgroups per CU.  This module measures those choices on the GPU for a given
batch and keeps the fastest:

  * team size per size bucket (orders <= 8, 16, 24, 32, 48, 64);
  * pack size per bucket (rounds of teams per work item);
  * persistent workgroups per CU.

The search is coordinate descent from the current table: each parameter in
turn is set to every allowed value with the others fixed, each candidate is
timed (HIP events on the library's stream, inputs restored outside the timed
region, median of `reps` runs), and the best value is kept.  The result can be
saved as JSON (tuning/vbatched_gfx950.json) and is installed at import of
rocsolver_amd when that file exists.
"""
from __future__ import annotations

import json
import os

import rocsolver_amd as ra

BUCKET_ORDERS = [8, 16, 24, 32, 48, 64]
TEAMS = [16, 32, 64]
PACKS = [1, 2, 4, 8, 16]
WG_PER_CU = [1, 2]
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning",
                            "vbatched_gfx950.json")


def _time(handle, vb, A0, B0, op: str, reps: int) -> float:
    import torch
    times = []
    for _ in range(reps):
        vb.A.copy_(A0)
        vb.B.copy_(B0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        getattr(vb, op)(handle)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def tune_vbatched(handle, vb, op: str = "gesv", reps: int = 3, verbose: bool = False):
    """Coordinate-descent search of the schedule table for this batch.

    Returns (best_table, best_ms, log) and leaves the best table installed."""
    A0, B0 = vb.A.clone(), vb.B.clone()
    best = ra.vbatched_get_tuning()
    getattr(vb, op)(handle)  # warm-up (code objects, workspace)
    best_ms = _time(handle, vb, A0, B0, op, reps)
    log = [(list(best), best_ms)]

    def trial(table):
        nonlocal best, best_ms
        ra.vbatched_set_tuning(table)
        ms = _time(handle, vb, A0, B0, op, reps)
        log.append((list(table), ms))
        if verbose:
            print(f"  {table} {ms:.3f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = list(table), ms

    for k, order in enumerate(BUCKET_ORDERS):
        present = any((vb.n > (BUCKET_ORDERS[k - 1] if k else 0)) & (vb.n <= order))
        if not present:
            continue
        for ts in TEAMS:
            if ts >= order and ts != best[k]:
                t = list(best)
                t[k] = ts
                trial(t)
        for pk in PACKS:
            if pk != best[6 + k]:
                t = list(best)
                t[6 + k] = pk
                trial(t)
    for wg in WG_PER_CU:
        if wg != best[12]:
            t = list(best)
            t[12] = wg
            trial(t)
    ra.vbatched_set_tuning(best)
    vb.A.copy_(A0)
    vb.B.copy_(B0)
    return best, best_ms, log


def save(table, path: str = DEFAULT_PATH, note: str = "") -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"table": [int(v) for v in table], "note": note}, f, indent=1)


def load(path: str = DEFAULT_PATH):
    with open(path) as f:
        return json.load(f)["table"]
