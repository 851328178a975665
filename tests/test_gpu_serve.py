"""GPU tests of the continuous-batching job server (pis_server_*).

The contract: jobs that arrive while others run are admitted as slots and HBM budget
free up, their steps share launches with the resident jobs, and every job's result is
the same as the job run alone -- bit for bit, since a system's arithmetic does not depend
on the other systems in its launches.
"""
import threading
import time

import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import Engine, PisError, Server, workloads as W
from tests.conftest import force_rms_err

pytestmark = pytest.mark.gpu


def jobs(n=8, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        natoms = int(np.exp(rng.uniform(np.log(500), np.log(6000))))
        spec = W.lj_box(natoms, float(rng.uniform(0.7, 1.1)), float(rng.uniform(0.7, 1.5)), seed=100 + k)
        spec.steps = int(rng.integers(20, 120))
        out.append(spec)
    return out


def alone(spec, precision):
    with Engine(spec, precision) as eng:
        eng.load()
        eng.compute_potential()
        th = eng.run_nve(spec.dt, spec.steps)
        return th[-1, 0, :], eng.get_state(0)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_each_result_equals_the_job_alone(precision):
    specs = jobs()
    results = {}
    with Server(hbm_budget_bytes=8 << 30, max_systems=3, round_steps=16, precision=precision) as srv:
        ids = []
        for k, s in enumerate(specs):
            ids.append(srv.submit(s))
            if k % 3 == 2:
                time.sleep(0.02)  # later jobs arrive while earlier ones are resident
        for k, jid in enumerate(ids):
            results[k] = srv.wait(jid)
        st = srv.stats()
    assert st["completed"] == len(specs) and st["failed"] == 0
    assert st["peak_resident"] <= 3
    for k, s in enumerate(specs):
        r = results[k]
        print(f"job {k}: N={s.n} steps={s.steps} latency={r.latency_s:.3f}s", flush=True)
        assert r.status == 0, r.error
        th, (pos, vel, frc) = alone(s, precision)
        assert np.array_equal(r.positions, pos)
        assert np.array_equal(r.velocities, vel)
        assert np.array_equal(r.forces, frc)
        assert np.array_equal(r.thermo, th)


def test_admission_respects_budget_and_arrivals_overlap():
    specs = jobs(6, seed=9)
    # budget for about two of the largest jobs at a time
    big = max(specs, key=lambda s: s.n)
    with Server(hbm_budget_bytes=1 << 40, max_systems=64, round_steps=8) as srv:
        jid = srv.submit(big, nsteps=8)
        srv.wait(jid, state=False)
        budget = 2 * srv.stats()["peak_bytes"] + 1
    submitted = []

    with Server(hbm_budget_bytes=budget, max_systems=64, round_steps=8) as srv:
        def feeder():
            for s in specs:
                submitted.append(srv.submit(s))
                time.sleep(0.01)

        t = threading.Thread(target=feeder)
        t.start()
        t.join()
        res = [srv.wait(j, state=False) for j in submitted]
        st = srv.stats()
    assert all(r.status == 0 for r in res)
    assert st["peak_bytes"] <= budget
    assert st["peak_resident"] >= 1
    # somebody had to wait for memory
    assert any(r.admit_s - r.submit_s > 0.0 for r in res)


def test_oversized_job_rejected_and_bad_job_fails_alone():
    specs = jobs(3, seed=11)
    with Server(hbm_budget_bytes=4096, max_systems=4, round_steps=8) as srv:
        with pytest.raises(PisError) as e:
            srv.submit(specs[0])
        assert e.value.status == 8
    with Server(hbm_budget_bytes=8 << 30, max_systems=4, round_steps=8) as srv:
        bad = W.lj_box(800, 0.9, 1.0, seed=3)
        bad.positions = bad.positions.copy()
        bad.positions[1] = bad.positions[0]  # coincident atoms: infinite energy
        ok_ids = [srv.submit(s) for s in specs]
        bad_id = srv.submit(bad, nsteps=10)
        r_bad = srv.wait(bad_id)
        r_ok = [srv.wait(j) for j in ok_ids]
    assert r_bad.status == 3
    assert all(r.status == 0 for r in r_ok)


def test_wait_timeout_then_collect():
    spec = W.lj_box(2000, 0.9, 1.0, seed=4)
    with Server(hbm_budget_bytes=8 << 30, max_systems=4, round_steps=50) as srv:
        jid = srv.submit(spec, nsteps=3000)
        with pytest.raises(PisError) as e:
            srv.wait(jid, timeout_s=0.0)
        assert e.value.status == 12
        r = srv.wait(jid)
        assert r.status == 0 and r.steps == 3000
        with pytest.raises(PisError) as e:
            srv.wait(jid)
        assert e.value.status == 11


def test_large_job_among_small_ones_equals_alone_and_its_forces_match_the_oracle():
    """A 131,072-atom liquid job served next to small jobs that come and go: its final
    state is the job run alone, bit for bit (mixed precision), and the job's step-0
    forces (the admission's force launch, the same kernels) are within the mixed-precision
    bound of the oracle on a seeded sample of 4,000 atoms."""
    big = W.lj_liquid(32, 0.8442, 0.72, seed=31)  # 4 * 32^3 = 131,072 atoms
    big.steps = 60
    small = jobs(6, seed=13)
    with Server(hbm_budget_bytes=16 << 30, max_systems=8, round_steps=16, precision="mixed") as srv:
        ids = [srv.submit(small[0]), srv.submit(small[1])]
        big_id = srv.submit(big)
        for s in small[2:]:
            time.sleep(0.01)
            ids.append(srv.submit(s))
        r = srv.wait(big_id)
        rs = [srv.wait(j, state=False) for j in ids]
        st = srv.stats()
    print(f"N={big.n}: latency {r.latency_s:.3f}s, stats {st}", flush=True)
    assert r.status == 0, r.error
    assert all(x.status == 0 for x in rs)
    th, (pos, vel, frc) = alone(big, "mixed")
    assert np.array_equal(r.positions, pos)
    assert np.array_equal(r.velocities, vel)
    assert np.array_equal(r.forces, frc)
    assert np.array_equal(r.thermo, th)
    with Engine(big, "mixed") as eng:
        eng.load()
        eng.compute_potential()
        f0 = eng.get_state(0)[2]
    idx = np.sort(np.random.default_rng(31).choice(big.n, 4000, replace=False))
    f_ref, _ = po.OSystem.from_spec(big).forces_subset(idx)
    err = force_rms_err(f0[idx], f_ref)
    print(f"step-0 force RMS error vs oracle: {err:.3e}", flush=True)
    assert err <= 1e-5


def test_jobs_reusing_released_memory_equal_the_jobs_alone():
    """Jobs of nearly one shape (one staging class, so every round has the same launch
    topology) pass through three slots: each newly admitted job lands in pool memory that a
    retired job of another size used, and the resident set changes at every round.  Every
    job still equals the job run alone, and the round's step graph was captured again after
    each change."""
    rng = np.random.default_rng(61)
    specs = []
    for k in range(10):
        s = W.lj_box(int(rng.integers(1400, 1700)), 0.85, 1.0, seed=70 + k)
        s.steps = int(rng.integers(1, 5)) * 16
        specs.append(s)
    with Server(hbm_budget_bytes=8 << 30, max_systems=3, round_steps=16, precision="fp64") as srv:
        ids = [srv.submit(s) for s in specs]
        res = [srv.wait(j) for j in ids]
        st = srv.stats()
    print(f"stats {st}", flush=True)
    assert all(r.status == 0 for r in res)
    assert st["graph_instantiations"] >= 4
    for s, r in zip(specs, res):
        th, (pos, vel, frc) = alone(s, "fp64")
        assert np.array_equal(r.positions, pos)
        assert np.array_equal(r.velocities, vel)
        assert np.array_equal(r.forces, frc)
        assert np.array_equal(r.thermo, th)
