#!/usr/bin/env python3
"""Benchmark of the LJ force + velocity-Verlet hot path on one MI355X.

Default workload: LJ liquid, 1,048,576 atoms (64^3 fcc cells at reduced density 0.8442,
+-0.05 sigma seeded jitter), T* = 0.72, rc = 2.5 sigma, skin 0.3 sigma, dt = 0.005, NVE.
A "step" is one verlet_step_nve of the whole system with state resident in HBM
(position update + wrap + skin check, list rebuild when triggered, LJ forces with the
fused velocity update, per-step PE/KE/virial reduction -- the reference computes thermo
every step).  The headline value is the mixed-precision rate (FP32 pair arithmetic,
FP64 state); the plain FP64 path is timed on the same workload beside it.

Prints ONE JSON line.  Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s at 1M LJ atoms, FP64 vs FP32/FP64 mixed; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from pis_amd import workloads as W

    if name == "c2":
        spec = W.lj_liquid(64, 0.8442, 0.72, seed=2026, rcut=2.5, skin=0.3, dt=0.005)
        desc = ("LJ liquid 1,048,576 atoms (64^3 fcc, rho*=0.8442, +-0.05 sigma jitter), T*=0.72, "
                "rc=2.5 sigma, skin 0.3 sigma, dt=0.005, NVE velocity-Verlet")
    elif name == "c3":
        spec = W.lj_liquid(100, 0.8442, 1.0, seed=2027, rcut=3.0, skin=0.3, dt=0.005)
        desc = ("LJ liquid 4,000,000 atoms (100^3 fcc, rho*=0.8442), T*=1.0, rc=3.0 sigma, "
                "skin 0.3 sigma, dt=0.005, NVE velocity-Verlet")
    else:
        raise SystemExit(f"unknown config {name}")
    return spec, desc


def list_mean(spec):
    rcl = max(p[4] for p in spec.pairs) + spec.skin
    rho = spec.n / spec.box.volume()
    return 4.0 / 3.0 * math.pi * rcl ** 3 * rho


def force_kernel_bytes_per_atom(nbar):
    """Algorithmic HBM bytes of one force launch per atom in a run's steady state: the
    u16 list entries (2 B each) and their count (2 B), position (24 B) and type (1 B) of
    the atom, old force and velocity read (48 B), the skin-check reference position read
    (24 B), new force, velocity and the next step's position written (72 B).  Neighbour
    positions are staged once per tile from L2/MALL and counted at zero."""
    return 2.0 * nbar + 2 + 24 + 1 + 48 + 24 + 72


def measured_traffic(natoms):
    """HBM bytes per force launch from the committed rocprofv3 counter summary
    (FETCH_SIZE / WRITE_SIZE scaled by a streaming copy of known size), rescaled to the
    atom count; None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_force_lj_liquid_1M_traffic.md")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0].endswith(("_kib", "_dispatch", "_bytes", "atoms")):
            kv[parts[0]] = float(parts[1])
    need = ("fetch_kib_per_dispatch", "write_kib_per_dispatch", "copy_fetch_kib", "copy_write_kib", "copy_bytes", "atoms")
    if not all(k in kv for k in need):
        return None
    copy_kib = kv["copy_bytes"] / 1024.0
    b = 1024.0 * (kv["fetch_kib_per_dispatch"] * copy_kib / kv["copy_fetch_kib"]
                  + kv["write_kib_per_dispatch"] * copy_kib / kv["copy_write_kib"])
    return b * natoms / kv["atoms"]


def b_alg(nbar):
    """Algorithmic HBM bytes per atom-step of the step (the survey's byte model):
    4 n (int32 neighbour indices of the full list) + 24 position read + 24 force write +
    120 for the integrator (read x, v, F; write x, v) + 24 for the skin-check reference
    position; neighbour-position gathers are served by L2/MALL and count 0."""
    return 4.0 * nbar + 192.0


def drift_stats(etot, n):
    """Energy drift of an NVE record per atom: least-squares slope of H scaled to 10^4
    steps, H_end - H_start, and the largest |H - H_start|."""
    t = np.arange(etot.shape[0], dtype=np.float64)
    slope = float(np.polyfit(t, etot - etot[0], 1)[0])
    return {
        "fit_per_1e4_steps": abs(slope) * 1e4 / n,
        "end_minus_start": float(abs(etot[-1] - etot[0]) / n),
        "max_abs_dev": float(np.abs(etot - etot[0]).max() / n),
        "steps": int(etot.shape[0]),
        "unit": "epsilon/atom",
    }


def run_gpu(spec, precision, steps, warmup, graph, timing_steps=50):
    """The headline measurement: W warm-up steps with the timed call's own settings
    (thermo recorded every step, graph mode as timed), so the thermo ring and the captured
    step graphs exist before t0; then exactly K steps timed with kernel timing off.  The
    force kernel's launch time comes from a separate pass with HIP events around each force
    launch on the context's stream."""
    from pis_amd import Engine

    eng = Engine(spec, precision)
    eng.load()
    eng.compute_potential()
    _, _, f0 = eng.get_state(0)
    eng.set_graph_mode(graph)
    if warmup:
        eng.run_nve(spec.dt, warmup, record=True)
    r0 = eng.stats()["rebuilds"]
    t0 = time.perf_counter()
    th = eng.run_nve(spec.dt, steps, record=True)[:, 0, :]
    t1 = time.perf_counter()
    rebuilds = eng.stats()["rebuilds"] - r0
    eng.set_kernel_timing(True)
    eng.run_nve(spec.dt, timing_steps, record=True)
    st = eng.stats()
    eng.set_kernel_timing(False)
    eng.close()
    sec = t1 - t0
    return {
        "sec": sec,
        "rate": spec.n * steps / sec,
        "ms_per_step": 1e3 * sec / steps,
        "force_ms": st["force_kernel_ms"] / max(st["timed_force_launches"], 1),
        "rebuilds": rebuilds,
        "max_neighbors": st["max_neighbors"],
        "thermo": th,
        "f0": f0,
    }


def run_drift(spec, precision, warm, steps):
    """Energy drift over a fixed NVE span, independent of --steps: `warm` steps of
    equilibration, then `steps` recorded steps."""
    from pis_amd import Engine

    with Engine(spec, precision) as eng:
        eng.load()
        eng.compute_potential()
        eng.set_graph_mode(True)
        eng.run_nve(spec.dt, warm, record=False)
        r0 = eng.stats()["rebuilds"]
        t0 = time.perf_counter()
        h = eng.run_nve(spec.dt, steps, record=True)[:, 0, 2]
        sec = time.perf_counter() - t0
        rebuilds = eng.stats()["rebuilds"] - r0
    out = drift_stats(h, spec.n)
    out["warmup"] = warm
    # the same recorded, graph-replayed steps as the headline, over the whole span
    out["ms_per_step"] = round(1e3 * sec / steps, 5)
    out["atom_steps_per_s"] = round(spec.n * steps / sec, 1)
    out["rebuilds"] = rebuilds
    return out


def run_c3(spec, steps, warm):
    """C3 (4,000,000 atoms, rc 3.0): mixed-precision throughput over `steps` steps after
    `warm`, its energy drift over the timed span, and the FP64 path's drift over the same
    span; the step-0 forces of both precisions are returned for the oracle check."""
    from pis_amd import Engine

    out = {}
    forces = {}
    for prec in ("mixed", "fp64"):
        with Engine(spec, prec) as eng:
            eng.load()
            eng.compute_potential()
            _, _, forces[prec] = eng.get_state(0)
            eng.set_graph_mode(True)
            eng.run_nve(spec.dt, warm, record=True)
            r0 = eng.stats()["rebuilds"]
            t0 = time.perf_counter()
            h = eng.run_nve(spec.dt, steps, record=True)[:, 0, 2]
            sec = time.perf_counter() - t0
            rb = eng.stats()["rebuilds"] - r0
        out[prec] = {"value": round(spec.n * steps / sec, 1), "unit": "atom-steps/s",
                     "ms_per_step": round(1e3 * sec / steps, 5), "rebuilds": rb,
                     "drift": drift_stats(h, spec.n)}
    return out, forces


def run_c1(steps):
    """C1: the argon FCC crystal (4000 atoms, 100 K, rc 8.5125 A, dt 1 fs) for 1000 FP64
    steps on the GPU (the CPU leg times the oracle on the same run)."""
    from pis_amd import Engine, workloads as W

    spec = W.argon_fcc_c1()
    with Engine(spec, "fp64") as eng:
        eng.load()
        eng.compute_potential()
        eng.set_graph_mode(True)
        eng.run_nve(spec.dt, 20, record=True)
        t0 = time.perf_counter()
        th = eng.run_nve(spec.dt, steps, record=True)[:, 0, :]
        sec = time.perf_counter() - t0
    return spec, {"value": round(spec.n * steps / sec, 1), "unit": "atom-steps/s", "steps": steps,
                  "ms_per_step": round(1e3 * sec / steps, 5), "precision": "fp64",
                  "drift": drift_stats(th[:, 2], spec.n)}


def run_per_call(spec, steps):
    """The drop-in per-call path (LJGpuManager.verlet_step_nve, the reference's call
    pattern): every step ships positions, velocities and forces to the device and back
    over PCIe.  Reported beside the device-resident rate, never as the headline."""
    from pis_amd import Atoms, LennardJones, LJGpuManager

    mgr = LJGpuManager(skin=spec.skin, kb=spec.kb, precision="mixed")
    for (ti, tj, eps, sig, rc) in spec.pairs:
        mgr.insert((ti, tj), LennardJones(eps, sig, rc, True))
    atoms = Atoms(spec.type_ids, spec.masses, spec.positions, spec.velocities, spec.box)
    mgr.compute_potential(atoms)
    mgr.verlet_step_nve(atoms, spec.dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        mgr.verlet_step_nve(atoms, spec.dt)
    sec = time.perf_counter() - t0
    return {"value": round(spec.n * steps / sec, 1), "unit": "atom-steps/s", "steps": steps,
            "ms_per_step": round(1e3 * sec / steps, 3)}


def run_thermostats(spec, nvt_steps, npt_steps):
    """NVT (Nose-Hoover chain, graph replay) and NPT (MTK barostat around the NVT step; the
    box becomes general, and the list is kept while the deformation-aware displacement
    check allows) on the same system, mixed precision.  NPT's target is the starting
    pressure with a heavy barostat (tau_p = 50), so the box stays inside its cell grid."""
    from pis_amd import Engine

    out = {}
    T = 0.72
    if nvt_steps:
        with Engine(spec, "mixed") as eng:
            eng.load()
            eng.compute_potential()
            eng.nvt_init(T, T, 0.5, 3)
            eng.run_nvt(spec.dt, 50, record=False)
            t0 = time.perf_counter()
            eng.run_nvt(spec.dt, nvt_steps, record=False)
            sec = time.perf_counter() - t0
        out["nvt"] = {"value": round(spec.n * nvt_steps / sec, 1), "unit": "atom-steps/s", "steps": nvt_steps,
                      "ms_per_step": round(1e3 * sec / nvt_steps, 4), "tau_t": 0.5}
    if npt_steps:
        with Engine(spec, "mixed") as eng:
            eng.load()
            eng.compute_potential()
            p0 = float(eng.run_nve(spec.dt, 1)[0, 0, 4])
            eng.npt_init(T, T, 0.5, p0, 50.0, 3)
            eng.run_npt(spec.dt, 10, record=False)
            t0 = time.perf_counter()
            eng.run_npt(spec.dt, npt_steps, record=False)
            sec = time.perf_counter() - t0
            h, _ = eng.npt_box(0)
        out["npt"] = {"value": round(spec.n * npt_steps / sec, 1), "unit": "atom-steps/s", "steps": npt_steps,
                      "ms_per_step": round(1e3 * sec / npt_steps, 4), "p_target": round(p0, 6), "tau_p": 50.0,
                      "max_tilt": float(np.abs(h - np.diag(np.diag(h))).max())}
    return out


def run_dump(spec, steps, cadence):
    """Simulation::run's output path (simulation.rs:76-86) on the same system, mixed
    precision: NVE in chunks of `cadence` steps with a thermo record per step and a
    trajectory block (dump_traj.rs:33-74) after each chunk, written to /dev/null so the
    figure is transfer + formatting, not the disk.  Positions leave the device while the
    next chunk runs (pinned double buffer) and are formatted by the native writer on a
    helper thread; the same loop without frames is timed beside it."""
    from pis_amd import Engine
    from pis_amd.script import AsyncFrames, DumpTraj

    chunks = max(1, steps // cadence)
    out = {"cadence": cadence, "frames": chunks, "steps": chunks * cadence,
           "writer_threads": min(16, os.cpu_count() or 1)}
    with Engine(spec, "mixed") as eng:
        eng.load()
        eng.compute_potential()
        eng.run_nve(spec.dt, cadence)
        t0 = time.perf_counter()
        for _ in range(chunks):
            eng.run_nve(spec.dt, cadence)
        plain = time.perf_counter() - t0
        dumper = DumpTraj(os.devnull)
        frames = AsyncFrames(eng, dumper, spec.type_ids)
        t0 = time.perf_counter()
        for c in range(chunks):
            eng.run_nve(spec.dt, cadence)
            frames.frame((c + 1) * cadence, spec.box)
        frames.close()
        sec = time.perf_counter() - t0
        dumper.close()
    n = spec.n * chunks * cadence
    out["value"] = round(n / sec, 1)
    out["unit"] = "atom-steps/s"
    out["without_frames"] = round(n / plain, 1)
    out["s_per_frame"] = round(sec / chunks, 4)
    return out


def run_ensemble(steps, warmup, direct_steps, accuracy=False, tune=True):
    """Batched ensemble: 512 Kob-Andersen systems (256-32,768 atoms, log-uniform from seed 7,
    density 1.2, T* 0.5-1.5) in one context, mixed precision.  Each step is one replay of
    the captured step graph (one launch per kernel per LDS staging class over all systems);
    the same batch with direct launches is timed beside it.  With `tune`, the tiles are the
    autotuner's: every launch class with at least 5 % of the atoms is timed as one batch at
    the candidate tiles around the planner's pick, and moves to another tile only if that
    is 10 % faster (untimed set-up, reported beside)."""
    import dataclasses

    from pis_amd import Engine, workloads as W

    specs = W.ka_ensemble(512, seed=7, nmin=256, nmax=32768, steps=steps)
    natoms = sum(s.n for s in specs)
    dts = [s.dt for s in specs]
    out = {"n_systems": len(specs), "total_atoms": natoms, "steps": steps,
           "min_atoms": min(s.n for s in specs), "max_atoms": max(s.n for s in specs)}
    if tune:
        # the autotuner (pis_amd.tune) in a child process, on the same seeded ensemble
        t0 = time.perf_counter()
        try:
            r = subprocess.run([sys.executable, "-m", "pis_amd.tune", "--ka-ensemble", "512", "--seed", "7",
                                "--nmin", "256", "--nmax", "32768"], capture_output=True, text=True, timeout=600,
                               cwd=os.path.dirname(os.path.abspath(__file__)))
            if r.returncode != 0:
                raise RuntimeError(r.stderr[-2000:])
            tuned = json.loads(r.stdout.strip().splitlines()[-1])
            tiles = [tuple(t) if t else None for t in tuned["tiles"]]
            if len(tiles) != len(specs):
                raise RuntimeError(f"{len(tiles)} tiles for {len(specs)} systems")
        except (RuntimeError, ValueError, IndexError, KeyError, subprocess.TimeoutExpired) as e:
            # the timed run stays on the planner's tiles, and the line says why
            out["autotune"] = {"error": str(e)[-500:]}
            log(f"ensemble autotune failed, model tiles kept: {str(e)[-300:]}")
        else:
            specs = [dataclasses.replace(s, tile=t) for s, t in zip(specs, tiles)]
            out["autotune"] = {"seconds": round(time.perf_counter() - t0, 2), "classes": tuned["classes"],
                               "systems_retiled": sum(t is not None for t in tiles)}
            log(f"ensemble autotune: {len(tuned['classes'])} classes timed in {out['autotune']['seconds']} s, "
                f"{out['autotune']['systems_retiled']} systems off the model's tile")
    with Engine(specs, "mixed") as eng:
        eng.load()
        eng.compute_potential()
        eng.set_graph_mode(True)
        eng.run_nve(dts, warmup, record=False)
        t0 = time.perf_counter()
        eng.run_nve(dts, steps, record=False)
        sec = time.perf_counter() - t0
        out["value"] = round(natoms * steps / sec, 1)
        out["ms_per_step"] = round(1e3 * sec / steps, 5)
        out["rebuilds_per_system"] = round(eng.stats()["rebuilds"] / len(specs), 1)
        if direct_steps:
            eng.set_graph_mode(False)
            t0 = time.perf_counter()
            eng.run_nve(dts, direct_steps, record=False)
            sec = time.perf_counter() - t0
            out["direct_launch_value"] = round(natoms * direct_steps / sec, 1)
            out["direct_launch_steps"] = direct_steps
    out["unit"] = "atom-steps/s"
    if accuracy:
        out["accuracy_sample"] = sample_accuracy(specs, 8, 2030, 10000)
    return out


def sample_accuracy(specs, sample, seed, drift_steps, warm=300):
    """Force RMS error of mixed precision against the oracle, and NVE drift over
    `drift_steps` recorded steps, on `sample` seeded systems of a batch, after `warm` steps
    from the generated lattice (on the lattice itself the forces nearly cancel, and an error
    relative to their RMS says little).  The systems run together in one context per
    precision: a system's arithmetic does not depend on the other systems of its launches,
    so these are the batch's numbers.  The oracle is the full-list gather (orc_forces_subset)
    at the warmed-up positions when every side has at least 3 cells of rc + skin, else the
    all-pairs minimum-image sum (LJManager): with fewer cells the reference's cell stencil
    visits a neighbour cell twice, which the device's whole-box staging does not reproduce."""
    import copy

    from oracle import pyoracle as po
    from pis_amd import Engine
    from pis_amd.engine import plan

    pick = sorted(np.random.default_rng(seed).choice(len(specs), sample, replace=False).tolist())
    sub = [specs[k] for k in pick]
    dts = [s.dt for s in sub]
    errs, refs = [], []
    drift = {}
    for prec in ("mixed", "fp64"):
        with Engine(sub, prec) as eng:
            eng.load()
            eng.compute_potential()
            eng.set_graph_mode(True)
            eng.run_nve(dts, warm, record=False)
            if prec == "mixed":
                for k, s in enumerate(sub):
                    pos, _, f = eng.get_state(k)
                    at = copy.copy(s)
                    at.positions = pos
                    o = po.OSystem.from_spec(at, velocities=False)
                    if min(plan(s, prec)["ncell"]) >= 3:
                        f_ref, _ = o.forces_subset(np.arange(s.n))
                        refs.append("full list")
                    else:
                        o.compute(po.MGR_BRUTE)
                        f_ref = o.force.copy()
                        refs.append("all pairs")
                    rms = math.sqrt(float((f_ref ** 2).sum(axis=1).mean()))
                    errs.append(math.sqrt(float(((f - f_ref) ** 2).sum(axis=1).mean())) / rms)
            th = eng.run_nve(dts, drift_steps)
        drift[prec] = [drift_stats(th[:, k, 2], s.n)["fit_per_1e4_steps"] for k, s in enumerate(sub)]
    return {
        "systems": pick,
        "atoms": [s.n for s in sub],
        "oracle": refs,
        "force_rms_err_mixed": errs,
        "force_rms_err_mixed_max": max(errs),
        "force_state": f"after {warm} NVE steps",
        "drift_fit_per_1e4_steps_mixed": drift["mixed"],
        "drift_fit_per_1e4_steps_fp64": drift["fp64"],
        "drift_fit_per_1e4_steps_max": max(drift["mixed"]),
        "drift_steps": drift_steps,
        "warmup": warm,
        "unit_drift": "epsilon/atom",
    }

def serving_jobs(n_jobs, rate, seed=2028):
    """Serving stream: n_jobs seeded single-species LJ boxes, atoms log-uniform on
    [500, 200000], density U(0.7, 1.1), T* U(0.7, 1.5), steps log-uniform on [100, 5000],
    Poisson arrivals at `rate` jobs/s (exponential gaps)."""
    from pis_amd import workloads as W

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    specs = []
    for k in range(n_jobs):
        n = int(np.exp(rng.uniform(np.log(500), np.log(200000))))
        rho = float(rng.uniform(0.7, 1.1))
        temp = float(rng.uniform(0.7, 1.5))
        steps = int(np.exp(rng.uniform(np.log(100), np.log(5000))))
        spec = W.lj_box(n, rho, temp, seed=seed * 1000 + k)
        spec.steps = steps
        specs.append(spec)
    return specs, arrivals


def offered_load(rate):
    """Mean atom-steps/s the stream asks for: rate x E[atoms] x E[steps], both
    log-uniform: E = (b - a) / ln(b / a)."""
    e_atoms = (200000 - 500) / np.log(200000 / 500)
    e_steps = (5000 - 100) / np.log(5000 / 100)
    return rate * e_atoms * e_steps


def run_serving(n_jobs, rate, budget_gib=64, round_steps=50):
    """Submits the stream at its arrival times (a feeder thread) to one mixed-precision
    server and waits for every job; latency is completion minus submission.  Jobs are
    built before the clock starts, so the feeder only sleeps and submits."""
    import threading

    from pis_amd import Server

    specs, arrivals = serving_jobs(n_jobs, rate)
    ids = [None] * n_jobs
    with Server(hbm_budget_bytes=budget_gib << 30, max_systems=256, round_steps=round_steps,
                precision="mixed") as srv:
        t0 = time.perf_counter()

        def feeder():
            for k, s in enumerate(specs):
                delay = arrivals[k] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(s)

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        res = [srv.wait(j, state=False) for j in ids]
        st = srv.stats()
    bad = [r for r in res if r.status != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} serving jobs failed: {bad[0].error}")
    lat = np.array([r.latency_s for r in res])
    wait = np.array([r.admit_s - r.submit_s for r in res])
    first = min(r.submit_s for r in res)
    last = max(r.done_s for r in res)
    last_arrival = max(r.submit_s for r in res)
    work = sum(r.n_atoms * r.steps for r in res)
    # steady state: the jobs that arrived after the first 20 % and before the last 20 % of
    # the arrival window (no ramp-up, no drain)
    lo, hi = first + 0.2 * (last_arrival - first), first + 0.8 * (last_arrival - first)
    mid = np.array([lo <= r.submit_s <= hi for r in res])
    return {
        "jobs": n_jobs,
        "arrival_rate_jobs_per_s": rate,
        "offered_load": round(float(offered_load(rate)), 1),
        "p50_s": round(float(np.percentile(lat, 50)), 4),
        "p99_s": round(float(np.percentile(lat, 99)), 4),
        "mean_s": round(float(lat.mean()), 4),
        "max_s": round(float(lat.max()), 4),
        "steady_jobs": int(mid.sum()),
        "steady_p50_s": round(float(np.percentile(lat[mid], 50)), 4),
        "steady_p99_s": round(float(np.percentile(lat[mid], 99)), 4),
        "queue_wait_p99_s": round(float(np.percentile(wait, 99)), 4),
        "throughput": round(work / (last - first), 1),
        "unit": "atom-steps/s",
        "makespan_s": round(last - first, 3),
        "arrival_window_s": round(last_arrival - first, 3),
        "offered_sampled": round(work / max(last_arrival - first, 1e-9), 1),
        "drain_s": round(last - last_arrival, 3),
        "atom_steps": int(work),
        "peak_resident_jobs": st["peak_resident"],
        "peak_bytes": st["peak_bytes"],
        "pool_bytes": st["pool_bytes"],
        "graph_instantiations": st["graph_instantiations"],
        "graph_build_s": round(st["graph_build_s"], 4),
        "rounds": st["rounds"],
        "budget_bytes": budget_gib << 30,
        "round_steps": round_steps,
        "precision": "mixed",
    }

def cpu_baseline(spec, steps):
    """The oracle (faithful LJVOffsetManager + verlet_step_nve, 1 core) on `steps` steps
    of the same system; also returns its initial forces for the force-error figures."""
    from oracle import pyoracle as po

    o = po.OSystem.from_spec(spec)
    t0 = time.perf_counter()
    o.compute(po.MGR_OFFSET)
    t1 = time.perf_counter()
    f_ref = o.force.copy()
    t2 = time.perf_counter()
    for _ in range(steps):
        o.verlet_step_nve(spec.dt, po.MGR_OFFSET)
        o.kinetic_energy()
    t3 = time.perf_counter()
    return f_ref, spec.n * steps / (t3 - t2), t1 - t0, t3 - t2


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_c1(spec, steps):
    """The oracle's LJVOffsetManager + verlet_step_nve (1 core) on C1's full run."""
    from oracle import pyoracle as po

    o = po.OSystem.from_spec(spec)
    t0 = time.perf_counter()
    o.run_nve(spec.dt, steps, po.MGR_OFFSET, kb=spec.kb)
    return spec.n * steps / (time.perf_counter() - t0)


def c3_force_errors(spec, forces, sample, seed=2029):
    """Force errors of both precisions at C3 against the oracle on `sample` seeded atoms
    (the oracle's full-list gather for those atoms alone)."""
    from oracle import pyoracle as po

    idx = np.sort(np.random.default_rng(seed).choice(spec.n, sample, replace=False))
    o = po.OSystem.from_spec(spec)
    f_ref, _ = o.forces_subset(idx)
    rms = math.sqrt(float((f_ref ** 2).sum(axis=1).mean()))
    fm, fd = forces["mixed"][idx], forces["fp64"][idx]
    return {
        "sample_atoms": sample,
        "force_rms_err_mixed": math.sqrt(float(((fm - f_ref) ** 2).sum(axis=1).mean())) / rms,
        "force_max_err_fp64": float(np.abs(fd - f_ref).max() / rms),
    }


def cpu_ideal(spec, steps, threads):
    from oracle import pyoracle as po

    o = po.OSystem.from_spec(spec)
    t0 = time.perf_counter()
    o.ideal_run_nve(spec.dt, steps, spec.skin, threads)
    return spec.n * steps / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-graph", action="store_true", help="direct launches instead of step-graph replay")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle force checks")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="recorded NVE steps of the drift legs, after 300 of equilibration (0: skip)")
    ap.add_argument("--c3-steps", type=int, default=2000, help="steps of the 4M-atom C3 leg (0: skip)")
    ap.add_argument("--c1-steps", type=int, default=1000, help="steps of the C1 argon leg (0: skip)")
    ap.add_argument("--per-call-steps", type=int, default=20,
                    help="steps of the PCIe-inclusive per-call path (0: skip)")
    ap.add_argument("--nvt-steps", type=int, default=2000, help="steps of the NVT leg (0: skip)")
    ap.add_argument("--npt-steps", type=int, default=300, help="steps of the NPT leg (0: skip)")
    ap.add_argument("--dump-steps", type=int, default=200,
                    help="steps of the trajectory-output leg, one frame per --dump-cadence steps (0: skip)")
    ap.add_argument("--dump-cadence", type=int, default=20)
    ap.add_argument("--ensemble-steps", type=int, default=5000,
                    help="steps of the 512-system batched ensemble (0: skip)")
    ap.add_argument("--no-ensemble-tune", action="store_true",
                    help="run the ensemble on the planner's model tiles instead of the autotuner's")
    ap.add_argument("--serve-jobs", type=int, default=500,
                    help="jobs of each Poisson serving stream (0: skip)")
    ap.add_argument("--serve-low-rate", type=float, default=25.0,
                    help="arrival rate (jobs/s) of the second, below-capacity serving stream")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30,
                    help="bytes of the streaming-copy calibration run after the timed runs (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(systems shard by spatial domain or by job across GPUs; nothing here launches that).",
              file=sys.stderr)
        return 2

    graph = not args.no_graph
    spec, desc = workload(args.config)
    nbar = list_mean(spec)
    log(f"{args.config}: N={spec.n}, mean list length {nbar:.1f}")

    log(f"mixed precision: warmup {args.warmup}, timed {args.steps} steps, graph replay {graph}")
    mixed = run_gpu(spec, "mixed", args.steps, args.warmup, graph)
    log(f"mixed: {mixed['rate']:.4e} atom-steps/s, {mixed['ms_per_step']:.4f} ms/step, "
        f"force kernel {mixed['force_ms']:.4f} ms, rebuilds {mixed['rebuilds']}")
    fp64 = None
    if not args.no_fp64:
        log(f"fp64: warmup {args.warmup}, timed {args.steps} steps")
        fp64 = run_gpu(spec, "fp64", args.steps, args.warmup, graph)
        log(f"fp64: {fp64['rate']:.4e} atom-steps/s, {fp64['ms_per_step']:.4f} ms/step, "
            f"force kernel {fp64['force_ms']:.4f} ms, rebuilds {fp64['rebuilds']}")

    drift = None
    if args.drift_steps:
        drift = {}
        for prec in ("mixed", "fp64"):
            log(f"drift leg ({prec}): 300 + {args.drift_steps} NVE steps")
            drift[prec] = run_drift(spec, prec, 300, args.drift_steps)
            log(f"drift {prec}: fit {drift[prec]['fit_per_1e4_steps']:.3e}, end-start "
                f"{drift[prec]['end_minus_start']:.3e} eps/atom over {args.drift_steps} steps")

    c3 = c3_forces = c3_spec = None
    if args.c3_steps and args.config == "c2":
        c3_spec, _ = workload("c3")
        log(f"C3 leg: N={c3_spec.n}, rc 3.0, {args.c3_steps} steps per precision")
        c3, c3_forces = run_c3(c3_spec, args.c3_steps, 20)
        log(f"C3 mixed: {c3['mixed']['value']:.4e} atom-steps/s ({c3['mixed']['ms_per_step']} ms/step), "
            f"fp64 {c3['fp64']['value']:.4e}")

    c1 = c1_spec = None
    if args.c1_steps:
        log(f"C1 leg: argon crystal, {args.c1_steps} FP64 steps on the GPU")
        c1_spec, c1 = run_c1(args.c1_steps)
        log(f"C1 GPU: {c1['value']:.4e} atom-steps/s, {c1['ms_per_step']} ms/step")

    per_call = None
    if args.per_call_steps:
        log(f"per-call path (state over PCIe every step): {args.per_call_steps} steps")
        per_call = run_per_call(spec, args.per_call_steps)
        log(f"per-call: {per_call['value']:.4e} atom-steps/s")

    thermostats = None
    if args.nvt_steps or args.npt_steps:
        log(f"NVT {args.nvt_steps} steps, NPT {args.npt_steps} steps (mixed)")
        thermostats = run_thermostats(spec, args.nvt_steps, args.npt_steps)
        for k, v in thermostats.items():
            log(f"{k}: {v['value']:.4e} atom-steps/s, {v['ms_per_step']} ms/step")
    dump = None
    if args.dump_steps:
        log(f"trajectory output: {args.dump_steps} steps, a frame every {args.dump_cadence} (mixed)")
        dump = run_dump(spec, args.dump_steps, args.dump_cadence)
        log(f"dump: {dump['value']:.4e} atom-steps/s with frames ({dump['without_frames']:.4e} without), "
            f"{dump['s_per_frame']} s per frame")

    ensemble = None
    if args.ensemble_steps:
        log(f"batched ensemble: 512 KA systems, {args.ensemble_steps} steps, graph replay")
        ensemble = run_ensemble(args.ensemble_steps, 100, min(args.ensemble_steps, 1000), accuracy=not args.no_cpu,
                                tune=not args.no_ensemble_tune)
        log(f"ensemble: {ensemble['total_atoms']} atoms, {ensemble['value']:.4e} atom-steps/s "
            f"(direct launches {ensemble.get('direct_launch_value', 0):.4e})")

    serving = serving_low = None
    if args.serve_jobs:
        for rate in (50.0, args.serve_low_rate):
            log(f"serving: {args.serve_jobs} Poisson jobs at {rate:g}/s "
                f"(offered {offered_load(rate):.3e} atom-steps/s), continuous batching")
            r = run_serving(args.serve_jobs, rate)
            log(f"serving: p50 {r['p50_s']} s, p99 {r['p99_s']} s (steady {r['steady_p50_s']} / "
                f"{r['steady_p99_s']} s), {r['throughput']:.4e} atom-steps/s over {r['makespan_s']} s, "
                f"drain {r['drain_s']} s, {r['graph_instantiations']} graph builds ({r['graph_build_s']} s)")
            if rate == 50.0:
                serving = r
            else:
                serving_low = r
        if not args.no_cpu:
            log("serving: force error and 10^4-step drift on 8 sampled jobs")
            serving["accuracy_sample"] = sample_accuracy(serving_jobs(args.serve_jobs, 50.0)[0], 8, 2031, 10000)

    copy_gbs = None
    if args.copy_bytes:
        import ctypes as C

        from pis_amd import _capi

        ms = C.c_double(0.0)
        _capi.check(_capi.lib().pis_copy_calibration(args.copy_bytes, 5, C.byref(ms)))
        copy_gbs = 2.0 * args.copy_bytes / (ms.value * 1e-3) / 1e9
        log(f"streaming copy of {args.copy_bytes} B: {copy_gbs:.0f} GB/s (read+write)")

    balg = b_alg(nbar)
    fb = force_kernel_bytes_per_atom(nbar)
    achieved = balg * spec.n / (mixed["force_ms"] * 1e-3) / 1e9
    traffic = measured_traffic(spec.n) if args.config == "c2" else None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": round(traffic) if traffic else None,
        "traffic_source": "profiles/k_force_lj_liquid_1M_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-kernel scaled)",
        "kernel": "k_force (mixed)",
        "bytes_per_atom": round(balg, 1),
        "bytes_model": "B_alg = 4 n + 192 per atom-step (int32 full list, position, force, integrator, skin reference)",
        "force_kernel_ms": round(mixed["force_ms"], 5),
        "kernel_own_bytes_per_atom": round(fb, 1),
        "kernel_own_frac": round(fb * spec.n / (mixed["force_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "step_frac": round(balg * spec.n / (mixed["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "copy_kernel_GBs": round(copy_gbs, 1) if copy_gbs else None,
    }

    out_extra = {}
    cpu = None
    if not args.no_cpu:
        log(f"CPU oracle: initial forces + {args.cpu_steps} faithful steps on 1 core")
        f_ref, cpu_rate, t_force, t_steps = cpu_baseline(spec, args.cpu_steps)
        rms = math.sqrt(float((f_ref ** 2).sum(axis=1).mean()))
        out_extra["force_rms_err_mixed"] = float(
            math.sqrt(float(((mixed["f0"] - f_ref) ** 2).sum(axis=1).mean())) / rms)
        if fp64 is not None:
            out_extra["force_max_err_fp64"] = float(np.abs(fp64["f0"] - f_ref).max() / rms)
        cpu = {
            "value": round(cpu_rate, 1),
            "unit": "atom-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"{args.cpu_steps} NVE steps of the same {spec.n}-atom system with the oracle's "
                       f"LJVOffsetManager + verlet_step_nve restatement ({t_steps:.1f} s), {cpu_model()}"),
        }
        log(f"CPU oracle: {cpu_rate:.4e} atom-steps/s (1 core)")
        # the box's CPU share for one GPU is 16 threads, whatever the host has
        ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        ncores = min(ncores or 1, 16)
        ideal = cpu_ideal(spec, args.cpu_steps, ncores)
        out_extra["cpu_ideal_parallel"] = {
            "value": round(ideal, 1),
            "unit": "atom-steps/s",
            "cores": ncores,
            "kind": "idealised lock-free full-list OpenMP port (not the shipped rayon code)",
            "sample": f"{args.cpu_steps} NVE steps incl. the initial list build",
        }
        out_extra["host_cpu"] = cpu_model()
        out_extra["host_cpus_visible"] = os.cpu_count()
        log(f"CPU ideal parallel: {ideal:.4e} atom-steps/s on {ncores} threads")
        if c3 is not None:
            log("C3 force errors against the oracle on 20000 sampled atoms")
            c3["force_errors_vs_oracle"] = c3_force_errors(c3_spec, c3_forces, 20000)
            log(f"C3: {c3['force_errors_vs_oracle']}")
        if c1 is not None:
            log(f"C1 CPU oracle: {args.c1_steps} steps on 1 core")
            c1["cpu_oracle_1core"] = round(cpu_c1(c1_spec, args.c1_steps), 1)
            c1["gpu_over_cpu"] = round(c1["value"] / c1["cpu_oracle_1core"], 1)
            log(f"C1 CPU: {c1['cpu_oracle_1core']:.4e} atom-steps/s")

    nsteps = args.steps
    res = {
        "metric": METRIC,
        "value": round(mixed["rate"], 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair arithmetic, fp64 state/integration/reductions",
        "data": "synthetic (seeded fcc + jitter, Maxwell-Boltzmann velocities)",
        "config": {
            "workload": desc,
            "n_atoms": spec.n,
            "mean_list_length": round(nbar, 2),
            "precision": "mixed",
            "graph_replay": graph,
        },
        "fp64_value": round(fp64["rate"], 1) if fp64 else None,
        "fp64_ms_per_step": round(fp64["ms_per_step"], 5) if fp64 else None,
        "fp64_force_kernel_ms": round(fp64["force_ms"], 5) if fp64 else None,
        "mixed_over_fp64": round(mixed["rate"] / fp64["rate"], 3) if fp64 else None,
        "nve_drift": drift,
        "rebuilds_in_timed_steps": {"mixed": mixed["rebuilds"], "fp64": fp64["rebuilds"] if fp64 else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "c3_lj_liquid_4M": c3,
        "c1_argon_crystal": c1,
        "thermostat_barostat": thermostats,
        "trajectory_output": dump,
        "batched_ensemble": ensemble,
        "pcie_inclusive_per_call": per_call,
        "job_latency": serving,
        "job_latency_below_capacity": serving_low,
    }
    res.update(out_extra)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
