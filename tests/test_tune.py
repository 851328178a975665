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
