"""Empirical tile autotuner.

The planner picks each system's tile (cells per workgroup) from a model: lane use of the
256-thread workgroup times the workgroups per CU that the staging class's LDS allows.
This module measures instead: it plans the candidate tiles around the model's pick,
keeps one per distinct (tile, staging class) plan, times a short graph-replayed run of
each through pis_set_tile_override, and returns the fastest.  The override is process-
wide and applies to contexts created afterwards; `apply=True` leaves the winner set.
"""
from __future__ import annotations

import ctypes as C
import itertools
import time

from ._capi import check, lib
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


__all__ = ["autotune_tiles", "candidates", "time_tile", "set_tile_override"]
