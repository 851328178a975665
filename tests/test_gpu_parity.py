"""GPU parity: the gfx950 path through the C ABI against the CPU oracle.

Bars (from the project's correctness contract):
  * neighbour pair set bit-identical to build_neighbour_list (lennard_jones.rs:344-416)
    with every rcut raised by the skin (brute-force LJManager semantics when a box side
    holds fewer than 3 cells, where the reference's stencil double counts);
  * FP64 per-atom forces within 1e-10 of LJVOffsetManager (lennard_jones.rs:186-244),
    measured normwise max_i|dF_i| / RMS|F|; PE within 1e-12 of an exactly rounded sum;
  * the FP64 path tracks the oracle's NVE energies over 1000 steps (C1 argon crystal).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import Engine, PisError, workloads as W
from tests.conftest import force_err, force_rms_err

pytestmark = pytest.mark.gpu


def gpu_forces(spec, precision="fp64"):
    with Engine(spec, precision) as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
    return pe, f


def oracle_forces(spec, manager=po.MGR_OFFSET):
    o = po.OSystem.from_spec(spec)
    pe = o.compute(manager)
    return pe, o.force.copy()


def check_pe(pe, spec, pe_serial):
    """PE against the oracle two ways.  (1) 1e-12 relative to an exactly rounded sum
    (math.fsum of the oracle's per-atom energies, halved): the GPU's fixed-order tree
    sum has nothing but rounding to answer for.  (2) Against the reference's own serial
    half-list sum (lennard_jones.rs:229), within that sum's recursive-summation error
    bound n*u*sum|U| -- on ordered lattices its rounding errors do not cancel."""
    import math

    o = po.OSystem.from_spec(spec)
    _, pe_atom = o.compute_fulllist(0.0, True, brute=True)
    exact = math.fsum(pe_atom) / 2.0
    assert abs(pe - exact) <= 1e-12 * abs(exact)
    st, _ = o.list(0.0, brute=True)
    nterms = int(st[-1]) // 2
    bound = nterms * 2.0 ** -53 * (np.abs(pe_atom).sum() / 2.0)
    assert abs(pe - pe_serial) <= max(bound, 1e-12 * abs(exact))


SPECS = {
    "lj_liquid_4000": lambda: W.lj_liquid(10, 0.8442, 0.72, seed=3),
    "argon_c1_jittered": lambda: _jittered_argon(),
    "lj_rc3_6912": lambda: W.lj_liquid(12, 0.8442, 1.0, seed=5, rcut=3.0),
    "ka_2048": lambda: W.ka_mixture(2048, 1.2, 1.0, seed=9),
    "ka_5000": lambda: W.ka_mixture(5000, 1.2, 0.7, seed=10),
    "lj_dilute_3000": lambda: W.lj_box(3000, 0.3, 1.5, seed=4),
    # a long cutoff: about 520 list entries per atom, 4096-slot tiles of one cell
    "lj_rc5_16384": lambda: W.lj_liquid(16, 0.8442, 1.0, seed=6, rcut=5.0),
    # an orthorhombic box of unequal sides: cells 3.36 x 2.94 x 2.84 sigma (4 x 8 x 13)
    "lj_aniso_9856": lambda: _anisotropic_liquid(),
}


def _anisotropic_liquid():
    """FCC 8 x 14 x 22 unit cells at rho* = 0.8442, jittered by 0.1 sigma: a box of
    13.4 x 23.5 x 37.0 sigma whose list cells differ in width along each axis."""
    rng = np.random.default_rng(88)
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    lengths = [8 * a, 14 * a, 22 * a]
    pos = W.fcc_sites(8, 14, 22, a) + rng.uniform(-0.1, 0.1, (4 * 8 * 14 * 22, 3))
    pos = W.wrap_positions(pos, lengths)
    n = pos.shape[0]
    return W.SystemSpec(positions=pos, type_ids=np.ones(n, dtype=np.uint32), masses=np.array([1.0]),
                        pairs=[(1, 1, 1.0, 1.0, 2.5)], box=W.SimulationBox.from_lammps_data(
                            0.0, lengths[0], 0.0, lengths[1], 0.0, lengths[2]),
                        velocities=W.start_velocities(np.ones(n), 0.9, 1.0, rng), skin=0.3, kb=1.0, dt=0.005,
                        steps=100, name="lj_aniso")


def _jittered_argon():
    s = W.argon_fcc_c1()
    rng = np.random.default_rng(77)
    s.positions = W.wrap_positions(s.positions + rng.uniform(-0.1, 0.1, s.positions.shape), s.box.lengths)
    return s


@pytest.mark.parametrize("name", sorted(SPECS))
def test_fp64_forces_match_offset_manager(name):
    spec = SPECS[name]()
    pe, f = gpu_forces(spec)
    pe_ref, f_ref = oracle_forces(spec)
    assert force_err(f, f_ref) <= 1e-10
    check_pe(pe, spec, pe_ref)


@pytest.mark.parametrize("name", sorted(SPECS))
def test_pair_set_bit_identical(name):
    spec = SPECS[name]()
    with Engine(spec, "fp64") as eng:
        eng.load()
        off, idx = eng.neighbor_list()
    o = po.OSystem.from_spec(spec)
    st, ix = o.list(rcut_add=spec.skin)
    got = po.csr_to_sorted_pairs(off, idx)
    want = po.csr_to_sorted_pairs(st, ix)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [256, 500])
def test_small_box_pair_set_matches_brute_force(n):
    spec = W.ka_mixture(n, 1.2, 1.0, seed=n)
    with Engine(spec, "fp64") as eng:
        eng.load()
        off, idx = eng.neighbor_list()
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
    o = po.OSystem.from_spec(spec)
    st, ix = o.list(rcut_add=spec.skin, brute=True)
    assert np.array_equal(po.csr_to_sorted_pairs(off, idx), po.csr_to_sorted_pairs(st, ix))
    pe_ref, f_ref = oracle_forces(spec, po.MGR_BRUTE)
    assert force_err(f, f_ref) <= 1e-10
    check_pe(pe, spec, pe_ref)


@pytest.mark.parametrize("name", sorted(SPECS))
def test_mixed_force_rms_error(name):
    """Mixed precision (FP32 pair arithmetic) against the FP64 oracle: force RMS error
    <= 1e-5 relative and PE within 2e-6 relative, on every parity system (single species
    and Kob-Andersen mixtures, tile-relative and raw staging)."""
    spec = SPECS[name]()
    pe, f = gpu_forces(spec, "mixed")
    pe_ref, f_ref = oracle_forces(spec)
    err = force_rms_err(f, f_ref)
    print(f"{name}: mixed force RMS error {err:.3e}, PE {pe} vs {pe_ref}", flush=True)
    assert err <= 1e-5
    assert abs(pe - pe_ref) <= 2e-6 * abs(pe_ref)


def test_slab_denser_than_its_box_average_replans_its_tiles():
    """A liquid slab filling a third of its box: tiles planned for the box's average density
    meet halos of about 4,000 atoms, more than the largest LDS staging class holds even after
    growing; loading the state plans the tiles again for the measured peak, and forces,
    energy and a short run match the oracle."""
    from pis_amd import SimulationBox

    spec = W.lj_liquid(12, 0.8442, 0.72, seed=31)
    L = spec.box.lengths
    spec.box = SimulationBox((L[0], 0.0, 0.0, 0.0, L[1], 0.0, 0.0, 0.0, 3.0 * L[2]))
    pe_ref, f_ref = oracle_forces(spec)
    o = po.OSystem.from_spec(spec)
    ref = o.run_nve(spec.dt, 50, po.MGR_OFFSET, kb=spec.kb)
    for precision in ("fp64", "mixed"):
        with Engine(spec, precision) as eng:
            eng.load()
            pe = float(eng.compute_potential()[0])
            _, _, f = eng.get_state(0)
            th = eng.run_nve(spec.dt, 50)[:, 0, :]
        if precision == "fp64":
            assert force_err(f, f_ref) <= 1e-10
            check_pe(pe, spec, pe_ref)
            assert np.abs(th[:, 2] - ref[:, 2]).max() <= 1e-9 * np.abs(ref[:, 0]).max()
        else:
            assert force_rms_err(f, f_ref) <= 1e-5
            assert np.abs(th[:, 2] - ref[:, 2]).max() <= 1e-5 * np.abs(ref[:, 0]).max()


def test_nve_pressure_and_virials_track_oracle():
    """Every step of a 100-step FP64 run of the C1 argon crystal: pressure
    (properties.rs:61-65), the absolute virial sum_i r_i.F_i (properties.rs:49-51) and the
    periodic pair virial against the oracle's own trajectory, to 1e-10."""
    spec = _jittered_argon()
    nsteps = 100
    with Engine(spec, "fp64") as eng:
        eng.load()
        eng.compute_potential()
        th = eng.run_nve(spec.dt, nsteps)[:, 0, :]
    o = po.OSystem.from_spec(spec)
    o.compute(po.MGR_OFFSET)
    vol = spec.box.volume()
    for k in range(nsteps):
        o.verlet_step_nve(spec.dt, po.MGR_OFFSET)
        ke = o.kinetic_energy()
        w_abs = o.virial_trace()
        scale_abs = float(np.abs((o.pos * o.force).sum(axis=1)).sum())
        p_ref = o.pressure(ke)
        w_pair = o.pair_virial()
        assert abs(th[k, 5] - w_abs) <= 1e-10 * scale_abs, k
        assert abs(th[k, 4] - p_ref) <= 1e-10 * (2 * ke + scale_abs) / (3 * vol), k
        assert abs(th[k, 6] - w_pair) <= 1e-10 * abs(w_pair), k


def test_argon_c1_tracks_oracle_energies():
    spec = W.argon_fcc_c1()
    nsteps = 1000
    with Engine(spec, "fp64") as eng:
        eng.load()
        eng.compute_potential()
        th = eng.run_nve(spec.dt, nsteps)[:, 0, :]
    o = po.OSystem.from_spec(spec)
    ref = o.run_nve(spec.dt, nsteps, po.MGR_OFFSET, kb=spec.kb)
    scale = np.abs(ref[:, 0]).max()
    assert np.abs(th[:, 0] - ref[:, 0]).max() <= 1e-9 * scale  # PE
    assert np.abs(th[:, 1] - ref[:, 1]).max() <= 1e-9 * scale  # KE
    assert np.abs(th[:, 2] - ref[:, 2]).max() <= 1e-9 * scale  # H
    assert np.abs(th[:, 3] - ref[:, 3]).max() <= 1e-7 * np.abs(ref[:, 3]).max()  # T


def test_repeat_runs_bit_identical():
    spec = W.lj_liquid(10, 0.8442, 0.72, seed=8)
    out = []
    for _ in range(2):
        with Engine(spec, "fp64") as eng:
            eng.load()
            eng.compute_potential()
            th = eng.run_nve(spec.dt, 200)
            out.append((th, eng.get_state(0)))
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


def test_list_order_repeatable_in_a_box_four_cells_wide():
    """A box four list cells wide is still staged tile-relative, so a halo cell can hold an
    atom at the periodic image boundary that is staged on the far side of its cell (its x
    out of order).  The list build's row windows must not depend on thread order there:
    two builds give the same list entry for entry, the pair set of the oracle, and the
    same 100-step trajectory."""
    spec = W.lj_box(3000, 0.9, 1.2, seed=21, rcut=3.0)
    out = []
    for _ in range(2):
        with Engine(spec, "fp64") as eng:
            eng.load()
            off, idx = eng.neighbor_list()
            eng.compute_potential()
            th = eng.run_nve(spec.dt, 100)
            out.append((off, idx, th))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    st, ix = po.OSystem.from_spec(spec).list(rcut_add=spec.skin)
    assert np.array_equal(po.csr_to_sorted_pairs(out[0][0], out[0][1]), po.csr_to_sorted_pairs(st, ix))


def test_errors_map_to_status():
    spec = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    bad = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    bad.type_ids = bad.type_ids.copy()
    bad.type_ids[3] = 5
    with pytest.raises(PisError) as e:
        Engine(bad)
    assert e.value.status == 4
    nopair = W.lj_liquid(4, 0.8442, 0.72, seed=1)
    nopair.pairs = []
    with pytest.raises(PisError) as e:
        Engine(nopair)
    assert e.value.status == 2
    with Engine(spec) as eng:
        pos = spec.positions.copy()
        pos[1] = pos[0]  # coincident atoms -> infinite energy
        eng.set_state(0, pos, spec.velocities)
        with pytest.raises(PisError) as e:
            eng.compute_potential()
        assert e.value.status == 3


def test_nvt_tracks_oracle_nose_hoover_chain():
    """verlet_step_nvt_nhc with the target ramp (potential.rs:35-58, simulation.rs:53-57):
    FP64 GPU thermo against the oracle over 400 steps of the argon crystal heated from
    100 K toward 140 K (tau = 100 fs); etot is the reference's NVT Hamiltonian (PE + KE +
    chain KE)."""
    spec = W.argon_fcc_c1()
    nsteps = 400
    with Engine(spec, "fp64") as eng:
        eng.load()
        eng.compute_potential()
        eng.nvt_init(100.0, 140.0, 1.0, 3)
        th = eng.run_nvt(spec.dt, nsteps)[:, 0, :]
        pos, vel, _ = eng.get_state(0)
        xi, target = eng.nvt_chain(0)
    o = po.OSystem.from_spec(spec)
    ref = o.run_nvt(spec.dt, nsteps, 100.0, 140.0, 1.0, 3, po.MGR_OFFSET, kb=spec.kb)
    scale = np.abs(ref[:, 0]).max()
    for col in (0, 1, 2):
        assert np.abs(th[:, col] - ref[:, col]).max() <= 1e-9 * scale
    assert np.abs(th[:, 3] - ref[:, 3]).max() <= 1e-7 * np.abs(ref[:, 3]).max()
    assert ref[-100:, 3].mean() > ref[:100, 3].mean()  # the ramp heated it
    # after the last step the ramp stands at step nsteps-1 (nvt.rs:109-118)
    assert target == pytest.approx(100.0 + 40.0 / nsteps * (nsteps - 1), rel=1e-15)
    assert np.all(np.isfinite(xi)) and xi[0] != 0.0
    L = spec.box.lengths
    d = (pos - o.pos + L / 2) % L - L / 2
    assert np.abs(d).max() <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_state_comes_back_in_the_callers_atom_order_after_steps(precision):
    # two atom types: after set_state's sort the set written by the first step must carry
    # the sorted types and original indices (else types and read-back order are scrambled)
    spec = W.ka_mixture(4000, 1.2, 1.0, seed=71)
    o = po.OSystem.from_spec(spec)
    ref = o.run_nve(spec.dt, 3, po.MGR_OFFSET, kb=spec.kb)
    with Engine(spec, precision) as eng:
        eng.load()
        eng.compute_potential()
        th = eng.run_nve(spec.dt, 3)[:, 0, :]
        pos, vel, _ = eng.get_state(0)
    L = np.asarray(spec.box.lengths)
    d = (pos - o.pos + L / 2) % L - L / 2
    tol = 1e-10 if precision == "fp64" else 1e-6
    assert np.abs(d).max() <= tol
    assert np.abs(vel - o.vel).max() <= tol * 10
    assert abs(th[-1, 0] - ref[-1, 0]) <= 1e-6 * abs(ref[-1, 0])
