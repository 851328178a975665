"""GPU tests of the batched path: many systems of uneven size in one context.

Every kernel of a step covers all systems of a context at once (one launch per LDS
staging class), and each system's arithmetic is the same whether it runs alone or in a
batch, so a system's trajectory must be bit-identical in both; the same holds between
the directly launched step and its replayed hipGraph.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
import pis_amd
from pis_amd import Engine, workloads as W
from tests.conftest import force_err

pytestmark = pytest.mark.gpu


def small_ensemble():
    """Eight KA systems spanning the thin-box (whole-box staging) and tiled modes, plus a
    single-species box with another cutoff, so several staging classes share a batch."""
    specs = W.ka_ensemble(8, seed=3, nmin=256, nmax=9000, steps=100)
    specs.append(W.lj_box(3000, 0.9, 1.2, seed=21, rcut=3.0))
    return specs


def run(specs, precision, nsteps, graph=False):
    with Engine(specs, precision) as eng:
        eng.set_graph_mode(graph)
        eng.load()
        pe0 = eng.compute_potential()
        th = eng.run_nve([s.dt for s in specs], nsteps)
        states = [eng.get_state(k) for k in range(len(specs))]
        st = eng.stats()
    return pe0, th, states, st


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_alone(precision):
    specs = small_ensemble()
    nsteps = 150
    pe0, th, states, st = run(specs, precision, nsteps)
    assert st["rebuilds"] > 0
    for k, s in enumerate(specs):
        print(f"system {k}: N={s.n}", flush=True)
        a_pe0, a_th, a_states, _ = run([s], precision, nsteps)
        assert a_pe0[0] == pe0[k]
        assert np.array_equal(a_th[:, 0, :], th[:, k, :])
        for a, b in zip(a_states[0], states[k]):
            assert np.array_equal(a, b)


def test_graph_replay_equals_direct_launches():
    specs = small_ensemble()
    nsteps = 200
    _, th_d, st_d, _ = run(specs, "mixed", nsteps, graph=False)
    _, th_g, st_g, stats = run(specs, "mixed", nsteps, graph=True)
    assert stats["rebuilds"] > 0
    assert np.array_equal(th_d, th_g)
    for k in range(len(specs)):
        for a, b in zip(st_d[k], st_g[k]):
            assert np.array_equal(a, b)


def test_graph_replays_across_calls_and_eager_work():
    """Replays of one captured graph, interleaved with eager work (state reads, a force
    evaluation) between calls, match the same sequence with direct launches."""
    specs = small_ensemble()[:4]
    out = []
    for graph in (False, True):
        with Engine(specs, "fp64") as eng:
            eng.set_graph_mode(graph)
            eng.load()
            eng.compute_potential()
            recs = []
            for _ in range(4):
                recs.append(eng.run_nve([s.dt for s in specs], 40))
                recs.append(eng.compute_potential()[None, :])
                recs.append(np.concatenate([np.ravel(a) for a in eng.get_state(1)]))
            out.append(recs)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_batched_fp64_forces_match_oracle():
    specs = small_ensemble()
    with Engine(specs, "fp64") as eng:
        eng.load()
        pe = eng.compute_potential()
        forces = [eng.get_state(k)[2] for k in range(len(specs))]
    for k, s in enumerate(specs):
        o = po.OSystem.from_spec(s)
        thin = any(int(L // (max(p[4] for p in s.pairs) + s.skin)) < 3 for L in s.box.lengths)
        pe_ref = o.compute(po.MGR_BRUTE if thin else po.MGR_OFFSET)
        print(f"system {k}: N={s.n} thin={thin}", flush=True)
        assert force_err(forces[k], o.force) <= 1e-10
        assert abs(pe[k] - pe_ref) <= 1e-10 * abs(pe_ref)


def test_dense_halo_grows_staging_class():
    """A KA box whose fcc planes sit on cell faces puts ~2,180 atoms into some tile halos,
    past the 2,048-atom class its uniform-density plan picks: loading the state grows the
    class from the measured halo counts, and forces still match the oracle."""
    spec = W.ka_mixture(14365, 1.2, 1.0, seed=5)
    with Engine(spec, "fp64") as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        f = eng.get_state(0)[2]
        th = eng.run_nve(spec.dt, 50)
    o = po.OSystem.from_spec(spec)
    pe_ref = o.compute(po.MGR_OFFSET)
    assert force_err(f, o.force) <= 1e-10
    assert abs(pe - pe_ref) <= 1e-10 * abs(pe_ref)
    assert np.all(np.isfinite(th))


def test_nvt_batch_equals_alone_and_graph_equals_direct():
    specs = small_ensemble()[:5]
    nsteps = 120

    def run_nvt(ss, graph):
        with Engine(ss, "mixed") as eng:
            eng.set_graph_mode(graph)
            eng.load()
            eng.compute_potential()
            eng.nvt_init(1.1, 0.9, 0.3, 3)
            th = eng.run_nvt([s.dt for s in ss], nsteps)
            return th, [eng.get_state(k) for k in range(len(ss))]

    th_b, st_b = run_nvt(specs, False)
    th_g, st_g = run_nvt(specs, True)
    assert np.array_equal(th_b, th_g)
    for k, s in enumerate(specs):
        th_a, st_a = run_nvt([s], False)
        assert np.array_equal(th_a[:, 0, :], th_b[:, k, :])
        for a, b, c in zip(st_a[0], st_b[k], st_g[k]):
            assert np.array_equal(a, b) and np.array_equal(b, c)


def test_tile_autotuner_times_candidates_and_restores_the_model_pick():
    from pis_amd import tune

    spec = W.lj_liquid(12, 0.8442, 0.72, seed=31)
    model = pis_amd.plan(spec, "mixed")["tile"]
    best, results = tune.autotune_tiles(spec, "mixed", steps=10, limit=3)
    assert best in [r[0] for r in results] and len(results) >= 2
    assert all(r[2] > 0 for r in results)
    assert pis_amd.plan(spec, "mixed")["tile"] == model


def test_ensemble_autotuner_covers_every_class_and_retiled_systems_match_the_oracle():
    """autotune_ensemble times every launch class (min_share 0) at its candidate tiles and
    returns one tile per system from those candidates; a batch planned with tiles other
    than the model's (pis_system_desc::tile) still gives FP64 forces within 1e-10 of
    LJVOffsetManager and the model's planner is left in force."""
    import dataclasses

    from pis_amd import tune

    specs = W.ka_ensemble(6, seed=11, nmin=3000, nmax=12000, steps=50)  # 3,121-10,863 atoms
    models = [pis_amd.plan(s, "mixed")["tile"] for s in specs]
    tiles, report = tune.autotune_ensemble(specs, "mixed", steps=10, warmup=2, limit=3, min_share=0.0)
    classes = tune.ensemble_classes(specs, "mixed")
    assert len(report) == len(classes) and len(tiles) == len(specs)
    for r in report:
        cands = [tuple(t) for t, ms in r["ms_per_step"]]
        assert r["model_tile"] in cands and r["tile"] in cands and all(ms > 0 for _, ms in r["ms_per_step"])
    for t, m in zip(tiles, models):
        assert t is None or t != m
    assert [pis_amd.plan(s, "mixed")["tile"] for s in specs] == models
    # every system on a candidate tile that is not the model's pick
    retiled = []
    for s in specs:
        m = pis_amd.plan(s, "fp64")["tile"]
        other = next(t for t, _ in tune.candidates(s, "fp64", limit=4) if t != m)
        retiled.append(dataclasses.replace(s, tile=other))
        assert pis_amd.plan(retiled[-1], "fp64")["tile"] == other
    with Engine(retiled, "fp64") as eng:
        eng.load()
        pe = eng.compute_potential()
        forces = [eng.get_state(k)[2] for k in range(len(retiled))]
    for k, s in enumerate(retiled):
        o = po.OSystem.from_spec(s)
        pe_ref = o.compute(po.MGR_OFFSET)
        print(f"retiled system {k}: N={s.n} tile={s.tile}", flush=True)
        assert force_err(forces[k], o.force) <= 1e-10
        assert abs(pe[k] - pe_ref) <= 1e-10 * abs(pe_ref)

