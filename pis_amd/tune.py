"""Empirical tile autotuner.

The planner picks each system's tile (cells per workgroup) from a model: lane use of the
256-thread workgroup times the workgroups per CU that the staging class's LDS allows.
This module measures instead: it plans the candidate tiles around the model's pick,
keeps one per distinct (tile, staging class) plan, times a short graph-replayed run of
each through pis_set_tile_override, and returns the fastest.  The override is process-
wide and applies to contexts created afterwards; `apply=True` leaves the winner set.

A batch of systems is tuned per launch class (the systems whose tiles share a staging
class, a tile and a staging mode, so one build and one force launch cover them): each
class is timed as one batch at each candidate tile, and the winner goes into every system
of the class as SystemSpec.tile, which the C ABI carries as pis_system_desc::tile.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import itertools
import time

from ._capi import PisError, check, lib
from .engine import Engine, plan
from .workloads import SystemSpec


def set_tile_override(tile) -> None:
    """Cells per tile for contexts created from now on; None restores the model's pick."""
    arg = (C.c_int32 * 3)(*tile) if tile else None
    check(lib().pis_set_tile_override(arg))


def candidates(spec: SystemSpec, precision: str = "mixed", radius: int = 1, limit: int = 12):
    """Distinct plans reachable by moving each tile side by up to `radius` cells from the
    model's pick: [(tile, stage_capacity)], the model's own pick first."""
    set_tile_override(None)
    spec = dataclasses.replace(spec, tile=None)  # a system's own tile outranks the override
    base = plan(spec, precision)
    seen = {(base["tile"], base["stage_capacity"])}
    out = [(base["tile"], base["stage_capacity"])]
    steps = range(-radius, radius + 1)
    try:
        for d in sorted(itertools.product(steps, steps, steps), key=lambda v: sum(abs(x) for x in v)):
            t = tuple(max(1, base["tile"][k] + d[k]) for k in range(3))
            set_tile_override(t)
            try:
                p = plan(spec, precision)
            except Exception:  # staging beyond LDS: not a candidate
                continue
            key = (p["tile"], p["stage_capacity"])
            if key not in seen:
                seen.add(key)
                out.append(key)
            if len(out) >= limit:
                break
    finally:
        set_tile_override(None)
    return out


def time_tile(spec: SystemSpec, tile, precision: str = "mixed", steps: int = 100, warmup: int = 20) -> float:
    """Seconds per NVE step of `spec` planned with `tile` (graph replay, no thermo)."""
    spec = dataclasses.replace(spec, tile=None)
    set_tile_override(tile)
    try:
        with Engine(spec, precision) as eng:
            eng.load()
            eng.compute_potential()
            eng.set_graph_mode(True)
            eng.run_nve(spec.dt, warmup, record=False)
            t0 = time.perf_counter()
            eng.run_nve(spec.dt, steps, record=False)
            return (time.perf_counter() - t0) / steps
    finally:
        set_tile_override(None)


def autotune_tiles(spec: SystemSpec, precision: str = "mixed", steps: int = 100, radius: int = 1,
                   limit: int = 12, apply: bool = False):
    """(best tile, [(tile, stage_capacity, seconds per step)]) over the candidate plans."""
    results = []
    for tile, scap in candidates(spec, precision, radius, limit):
        results.append((tile, scap, time_tile(spec, tile, precision, steps)))
    best = min(results, key=lambda r: r[2])[0]
    if apply:
        set_tile_override(best)
    return best, results


def ensemble_classes(specs, precision: str = "mixed") -> dict:
    """The planner's model picks of a batch by launch class: {(stage_capacity, tile, mode):
    [system indices]}."""
    set_tile_override(None)
    out: dict = {}
    for i, s in enumerate(specs):
        p = plan(dataclasses.replace(s, tile=None), precision)
        out.setdefault((p["stage_capacity"], p["tile"], p["mode"]), []).append(i)
    return out


def time_batch(specs, precision: str = "mixed", steps: int = 60, warmup: int = 10) -> float:
    """Seconds per NVE step of a batch as planned (graph replay, no thermo)."""
    dts = [s.dt for s in specs]
    with Engine(specs, precision) as eng:
        eng.load()
        eng.compute_potential()
        eng.set_graph_mode(True)
        eng.run_nve(dts, warmup, record=False)
        t0 = time.perf_counter()
        eng.run_nve(dts, steps, record=False)
        return (time.perf_counter() - t0) / steps


def autotune_ensemble(specs, precision: str = "mixed", steps: int = 60, warmup: int = 10, radius: int = 1,
                      limit: int = 6, min_share: float = 0.05, min_gain: float = 0.10):
    """Tiles for a batch, class by class.  Each class holding at least `min_share` of the
    atoms is timed as one batch of its systems at each candidate tile around the model's
    pick for its largest system (a tile at which some system of the class would stage more
    atoms than LDS holds is left out), and takes the fastest if it beats the model's pick by
    more than `min_gain`.  A class's gain alone need not carry over to the whole batch,
    whose classes share the device: on the 512-system Kob-Andersen ensemble one class 9 %
    faster alone left the replayed step unchanged and made directly launched steps 15 %
    slower, hence the 10 % default.  Returns (tiles, report): one tile per
    system (None keeps the model's pick) and, per class timed, the milliseconds per step
    of every candidate."""
    classes = ensemble_classes(specs, precision)
    total = sum(s.n for s in specs)
    tiles = [None] * len(specs)
    report = []
    for (scap, model, mode), idx in sorted(classes.items(), key=lambda kv: -sum(specs[i].n for i in kv[1])):
        atoms = sum(specs[i].n for i in idx)
        if atoms < min_share * total:
            continue
        big = max(idx, key=lambda i: specs[i].n)
        cands = [t for t, _ in candidates(specs[big], precision, radius, limit)]
        if model not in cands:
            cands.insert(0, model)
        timed = []
        for t in cands:
            try:
                timed.append((t, time_batch([dataclasses.replace(specs[i], tile=t) for i in idx], precision,
                                            steps, warmup)))
            except PisError:  # a system of the class stages more than LDS holds at this tile
                if t == model:
                    raise
        model_s = next(sec for t, sec in timed if t == model)
        best, best_s = min(timed, key=lambda r: r[1])
        pick = best if best_s < (1.0 - min_gain) * model_s else model
        if pick != model:
            for i in idx:
                tiles[i] = pick
        report.append({"stage_capacity": scap, "model_tile": model, "mode": mode, "systems": len(idx),
                       "atoms": atoms, "tile": pick,
                       "ms_per_step": [[list(t), round(1e3 * sec, 5)] for t, sec in timed]})
    return tiles, report


__all__ = ["autotune_tiles", "autotune_ensemble", "candidates", "ensemble_classes", "time_batch", "time_tile",
           "set_tile_override"]


def main(argv=None) -> int:
    """`python -m pis_amd.tune --ka-ensemble N --seed S --nmin A --nmax B [--precision P]`:
    tunes that Kob-Andersen ensemble (workloads.ka_ensemble) and prints one JSON line,
    {"tiles": [...], "classes": [...], "seconds": t}.  A separate process leaves the
    caller's device runtime as it was: in one process, directly launched steps of a batch
    created after the tuner's many short-lived contexts ran 15 % slower (replayed steps
    did not change)."""
    import argparse
    import json

    from . import workloads as W

    ap = argparse.ArgumentParser(prog="python -m pis_amd.tune")
    ap.add_argument("--ka-ensemble", type=int, required=True)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--nmin", type=int, default=256)
    ap.add_argument("--nmax", type=int, default=32768)
    ap.add_argument("--precision", default="mixed")
    args = ap.parse_args(argv)
    specs = W.ka_ensemble(args.ka_ensemble, seed=args.seed, nmin=args.nmin, nmax=args.nmax, steps=100)
    t0 = time.perf_counter()
    tiles, report = autotune_ensemble(specs, args.precision)
    print(json.dumps({"tiles": [list(t) if t else None for t in tiles], "classes": report,
                      "seconds": round(time.perf_counter() - t0, 2)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

