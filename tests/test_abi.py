"""C ABI surface and host-side logic, no GPU needed.

* libpis_hip.so loads and exports every function include/pis_hip.h declares;
* the host-only planner (pis_plan_system) validates descriptors with the reference's
  error mapping (PisError variants) and produces the grid/tile/capacity plan;
* the exact cutoff thresholds (largest r^2 whose rounded sqrt <= rcut) are right;
* the synthetic workload generators produce the configured systems.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import pis_amd
from pis_amd import _capi, workloads as W
from pis_amd.engine import _DescHolder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pis_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(pis_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = pis_amd.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert set(names) <= set(_capi.SIGNATURES)


def test_abi_version_and_status_strings():
    lib = pis_amd.lib()
    assert lib.pis_abi_version() == 3
    for code in range(14):
        assert lib.pis_status_string(code) != b"unknown status"


def plan_status(spec, **kw):
    h = _DescHolder(spec, "fp64")
    for k, v in kw.items():
        setattr(h.desc, k, v)
    info = _capi.PlanInfo()
    return pis_amd.lib().pis_plan_system(C.byref(h.desc), C.byref(info)), info


def test_plan_error_mapping_mirrors_pis_error():
    spec = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    assert plan_status(spec)[0] == 0
    assert plan_status(spec, n_atoms=0)[0] == 1        # NoAtomsDefined
    assert plan_status(spec, n_pairs=0)[0] == 2        # PotentialNotInitialized
    bad = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    bad.type_ids = bad.type_ids.copy()
    bad.type_ids[0] = 0
    assert plan_status(bad)[0] == 4                    # InvalidAtomType
    s, _ = plan_status(spec, flags=1)
    assert s == 5


def test_general_boxes_plan_tile_relative_staging():
    # a tilted box (xy) stages tile-relative positions (its tiles leave two cells between
    # a tile and its periodic image); cells follow the perpendicular widths, so a tilt that
    # thins a side below 3 cells stages the whole box with a per-pair image
    big = W.lj_liquid(8, 0.8442, 0.72, seed=1)
    h = list(big.box.h)
    h[3] = 0.1 * h[0]
    big.box = pis_amd.SimulationBox(h)
    st, info = plan_status(big)
    assert st == 0 and info.mode == 0
    assert all(info.ncell[k] - info.tile[k] >= 2 for k in range(3))
    L = big.box.h[0]
    assert list(info.ncell)[:2] == [int(L / (2.5 + 0.3)), int(L / (2.5 + 0.3))]
    small = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    h = list(small.box.h)
    h[3] = 0.5
    small.box = pis_amd.SimulationBox(h)
    st, info = plan_status(small)
    assert st == 0 and info.mode == 2
    sing = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    holder = _DescHolder(sing, "fp64")
    holder.desc.h[0] = float("nan")
    info = _capi.PlanInfo()
    assert pis_amd.lib().pis_plan_system(C.byref(holder.desc), C.byref(info)) == 5


def test_plans_for_benchmark_configurations():
    c2 = pis_amd.plan(W.lj_liquid(64, 0.8442, 0.72, seed=1))
    assert c2["ncell"] == (38, 38, 38)
    assert c2["mode"] == 0 and c2["stage_capacity"] <= 2048
    assert c2["r2_list_max"] == W_r2max(2.8)
    c1 = pis_amd.plan(W.argon_fcc_c1())
    assert c1["ncell"] == (5, 5, 5)
    small = pis_amd.plan(W.ka_mixture(256, 1.2, 1.0, seed=1))
    assert small["mode"] == 2  # fewer than 3 cells: whole system staged, pairwise minimum image


def W_r2max(rc):
    r2 = rc * rc
    while math.sqrt(r2) > rc:
        r2 = math.nextafter(r2, 0.0)
    while math.sqrt(math.nextafter(r2, math.inf)) <= rc:
        r2 = math.nextafter(r2, math.inf)
    return r2


@pytest.mark.parametrize("rc", [2.5, 2.8, 3.3, 8.5125 + 1.0215, 9.534])
def test_exact_threshold_is_the_rounded_sqrt_boundary(rc):
    r2 = W_r2max(rc)
    assert math.sqrt(r2) <= rc < math.sqrt(math.nextafter(r2, math.inf)) or math.sqrt(
        math.nextafter(r2, math.inf)) > rc
    spec = W.lj_liquid(8, 0.8442, 0.72, seed=1)
    spec.pairs = [(1, 1, 1.0, 1.0, rc - spec.skin)]
    assert pis_amd.plan(spec)["r2_list_max"] == W_r2max((rc - spec.skin) + spec.skin)


def test_lj_liquid_generator_matches_configuration():
    s = W.lj_liquid(8, 0.8442, 0.72, seed=3)
    assert s.n == 4 * 8 ** 3
    assert s.n / s.box.volume() == pytest.approx(0.8442, rel=1e-12)
    L = s.box.lengths
    assert np.all(s.positions >= 0) and np.all(s.positions < L)
    v = s.velocities
    assert np.abs(v.sum(axis=0)).max() < 1e-10  # drift removed (velocities.rs:35-50)
    ke = 0.5 * (v ** 2).sum()
    assert 2 * ke / (3 * s.n) == pytest.approx(0.72, rel=1e-12)  # 3N dof (properties.rs:41-43)


def test_ka_mixture_composition_and_density():
    s = W.ka_mixture(1000, 1.2, 0.8, seed=5)
    assert s.n == 1000
    assert np.count_nonzero(s.type_ids == 2) == 200
    assert s.n / s.box.volume() == pytest.approx(1.2, rel=1e-12)
    assert len(s.pairs) == 3


def test_ensemble_sizes_are_log_uniform_in_range():
    ens = W.ka_ensemble(64, seed=3, steps=10)
    ns = np.array([s.n for s in ens])
    assert ns.min() >= 256 and ns.max() <= 32768
    assert abs(np.median(np.log(ns)) - 0.5 * (np.log(256) + np.log(32768))) < 0.7


def test_argon_configuration_units():
    s = W.argon_fcc_c1()
    assert s.n == 4000
    assert s.box.lengths[0] == pytest.approx(52.6)
    eps = s.pairs[0][2]
    assert eps == pytest.approx(0.0104 * 96.4853)
    assert s.pairs[0][4] == pytest.approx(2.5 * 3.405)
    assert s.skin == pytest.approx(0.3 * 3.405)


def test_server_needs_a_device_and_valid_arguments():
    lib = pis_amd.lib()
    srv = C.c_void_p()
    assert lib.pis_server_create(1 << 30, 0, 10, 1, C.byref(srv)) == 5  # no slots
    assert lib.pis_server_create(1 << 30, 4, 0, 1, C.byref(srv)) == 5   # no round length
    assert lib.pis_server_create(1 << 30, 4, 10, 7, C.byref(srv)) == 5  # unknown precision
    if lib.pis_device_count(C.byref(C.c_int32())) != 0:
        assert lib.pis_server_create(1 << 30, 4, 10, 1, C.byref(srv)) == 9


def test_atoms_own_their_arrays():
    # Atoms::new takes ownership: stepping the mirror must never write into the caller's
    # arrays (a workload spec reused afterwards, e.g. as the CPU reference's input)
    from pis_amd import Atoms

    spec = W.lj_box(64, 0.8, 1.0, seed=1)
    pos0, vel0 = spec.positions.copy(), spec.velocities.copy()
    atoms = Atoms(spec.type_ids, spec.masses, spec.positions, spec.velocities, spec.box)
    atoms.positions += 1.0
    atoms.velocities *= 2.0
    assert np.array_equal(spec.positions, pos0) and np.array_equal(spec.velocities, vel0)
