"""The tile autotuner's host side: candidate plans around the model's pick (no GPU)."""
import pis_amd
from pis_amd import tune, workloads as W


def test_candidates_are_distinct_plans_led_by_the_model_pick():
    spec = W.lj_liquid(64, 0.8442, 0.72, seed=1)
    model = pis_amd.plan(spec, "mixed")
    cands = tune.candidates(spec, "mixed", radius=1, limit=12)
    assert cands[0] == (model["tile"], model["stage_capacity"])
    assert len(cands) == len(set(cands)) and 2 <= len(cands) <= 12
    for tile, scap in cands:
        assert all(1 <= t <= n - 2 for t, n in zip(tile, model["ncell"]))
        assert scap in (1024, 1280, 1536, 1792, 2048, 3072, 4096)
    # the override is process-wide: the helper leaves the model's pick in force
    assert pis_amd.plan(spec, "mixed")["tile"] == model["tile"]


def test_a_systems_own_tile_is_planned_and_outranks_the_override():
    import dataclasses

    spec = W.lj_liquid(64, 0.8442, 0.72, seed=1)
    model = pis_amd.plan(spec, "mixed")
    own = dataclasses.replace(spec, tile=(2, 2, 3))
    assert pis_amd.plan(own, "mixed")["tile"] == (2, 2, 3)
    tune.set_tile_override((4, 1, 1))
    try:
        assert pis_amd.plan(spec, "mixed")["tile"] == (4, 1, 1)
        assert pis_amd.plan(own, "mixed")["tile"] == (2, 2, 3)
    finally:
        tune.set_tile_override(None)
    # as with the override: each side clamped to ncell - 2, and a tile whose halo would
    # exceed the staging tables leaves the model's pick
    small = W.lj_liquid(8, 0.8442, 0.72, seed=1)
    assert pis_amd.plan(small, "mixed")["ncell"] == (4, 4, 4)
    assert pis_amd.plan(dataclasses.replace(small, tile=(5, 1, 1)), "mixed")["tile"] == (2, 1, 1)
    assert pis_amd.plan(dataclasses.replace(spec, tile=(200, 1, 1)), "mixed")["tile"] == model["tile"]
    assert pis_amd.plan(dataclasses.replace(spec, tile=(0, 0, 0)), "mixed")["tile"] == model["tile"]


def test_ensemble_classes_group_systems_by_their_model_plan():
    specs = W.ka_ensemble(24, seed=5, nmin=256, nmax=32768, steps=10)
    classes = tune.ensemble_classes(specs, "mixed")
    assert sorted(i for idx in classes.values() for i in idx) == list(range(len(specs)))
    for (scap, tile, mode), idx in classes.items():
        for i in idx:
            p = pis_amd.plan(specs[i], "mixed")
            assert (p["stage_capacity"], p["tile"], p["mode"]) == (scap, tile, mode)
