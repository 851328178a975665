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

from pis_amd import Engine, PisError, Server, workloads as W

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
