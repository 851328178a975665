"""GPU checks at the full benchmark sizes: the 1,048,576-atom LJ liquid (rc 2.5) and the
4,000,000-atom liquid (rc 3.0), through properties that do not depend on size.

* FP64 forces against the oracle's LJVOffsetManager on the whole system (<= 1e-10
  normwise) and the PE against its exactly rounded sum;
* mixed-precision force RMS error <= 1e-5 against the same oracle forces;
* Newton's third law: the total force vanishes to rounding;
* the neighbour list is symmetric (j lists i exactly when i lists j) and its size equals
  the oracle's pair count at rc + skin;
* a short NVE run conserves energy and the mixed and FP64 runs agree.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import Engine, workloads as W
from tests.conftest import force_err, force_rms_err

pytestmark = pytest.mark.gpu

CONFIGS = {
    "lj_liquid_1M_rc2.5": lambda: W.lj_liquid(64, 0.8442, 0.72, seed=2026, rcut=2.5, skin=0.3, dt=0.005),
    "lj_liquid_4M_rc3.0": lambda: W.lj_liquid(100, 0.8442, 1.0, seed=2027, rcut=3.0, skin=0.3, dt=0.005),
}


@pytest.fixture(scope="module", params=sorted(CONFIGS))
def system(request):
    spec = CONFIGS[request.param]()
    print(f"{request.param}: N={spec.n}", flush=True)
    o = po.OSystem.from_spec(spec)
    pe_ref = o.compute(po.MGR_OFFSET)
    return request.param, spec, pe_ref, o.force.copy()


def gpu(spec, precision, with_list=False):
    with Engine(spec, precision) as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
        off, idx = eng.neighbor_list() if with_list else (None, None)
    return pe, f, off, idx


def test_fp64_forces_and_energy_at_full_size(system):
    name, spec, pe_ref, f_ref = system
    with_list = spec.n <= 1 << 20  # the 4M list (0.5 G entries) is too big to sort on the host
    pe, f, off, idx = gpu(spec, "fp64", with_list)
    assert force_err(f, f_ref) <= 1e-10
    # the oracle's serial sum over ~10^7 pairs carries ~1e-13..1e-12 relative rounding
    assert abs(pe - pe_ref) <= 1e-10 * abs(pe_ref)
    # Newton's third law on the GPU's own forces: rounding-level net force
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).sum()
    if not with_list:
        return
    # symmetric list; its length is twice the number of pairs within rc + skin
    n = spec.n
    counts = np.diff(off)
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    key = rows * n + idx
    rkey = idx * n + rows
    assert np.array_equal(np.sort(key), np.sort(rkey))
    rcl = max(p[4] for p in spec.pairs) + spec.skin
    nbar = 4.0 / 3.0 * math.pi * rcl ** 3 * n / spec.box.volume()
    assert abs(counts.mean() - nbar) < 0.02 * nbar


def test_mixed_force_error_at_full_size(system):
    name, spec, pe_ref, f_ref = system
    pe, f, _, _ = gpu(spec, "mixed")
    assert force_rms_err(f, f_ref) <= 1e-5
    assert abs(pe - pe_ref) <= 1e-6 * abs(pe_ref)


def test_short_nve_run_conserves_energy_at_full_size(system):
    name, spec, _, _ = system
    out = {}
    for prec in ("fp64", "mixed"):
        with Engine(spec, prec) as eng:
            eng.load()
            eng.compute_potential()
            out[prec] = eng.run_nve(spec.dt, 200)[:, 0, :]
    for prec, th in out.items():
        drift = abs(th[-1, 2] - th[0, 2]) / spec.n
        assert drift < 2e-4, (prec, drift)
    # the two precisions follow the same trajectory over a short run
    assert np.abs(out["mixed"][:, 0] - out["fp64"][:, 0]).max() <= 1e-5 * np.abs(out["fp64"][:, 0]).max()
