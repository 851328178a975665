"""Pinning the CPU oracle (oracle/pis_oracle.c) without running the reference.

The reference (Rust) cannot be built here and ships no golden vectors, so every oracle
function is checked against something that fixes it independently: closed forms,
brute force, an independent numpy restatement, invariants, and the reference's own input
files.  Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) fails it.
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import box as pbox
from pis_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def lj_np(r, eps, sigma, rcut, shift=True):
    """Textbook LJ 12-6 in numpy: U(r) = 4 eps ((s/r)^12 - (s/r)^6) - U(rc); F_r = -dU/dr."""
    sr6 = (sigma / r) ** 6
    u = 4 * eps * (sr6 * sr6 - sr6)
    if shift:
        sc6 = (sigma / rcut) ** 6
        u = u - 4 * eps * (sc6 * sc6 - sc6)
    fr = 24 * eps * (2 * sr6 * sr6 - sr6) / r  # radial force, positive = repulsive
    return u, fr


# ---------------------------------------------------------------------- pair function
# lennard_jones.rs:32-55


def test_pair_zero_at_sigma_unshifted():
    u, f = po.lj_pair(1.3, 2.0, 5.0, False, [2.0, 0.0, 0.0])
    assert u == 0.0
    # at r = sigma the force is repulsive: 24 eps / sigma along +rij (force on j)
    assert f[0] == pytest.approx(24 * 1.3 / 2.0, rel=1e-15)


def test_pair_minimum_at_two_sixth_sigma():
    eps, sig = 0.7, 1.1
    rmin = 2 ** (1 / 6) * sig
    d = np.array([0.3, -0.4, 0.5])
    d = d / np.linalg.norm(d) * rmin
    u, f = po.lj_pair(eps, sig, 10.0, False, d)
    assert u == pytest.approx(-eps, rel=1e-14)
    assert np.abs(f).max() <= 1e-13 * 24 * eps / sig


def test_pair_shift_makes_energy_vanish_at_cutoff():
    eps, sig, rc = 1.0, 1.0, 2.5
    u, _ = po.lj_pair(eps, sig, rc, True, [rc, 0.0, 0.0])
    assert abs(u) <= 1e-16
    u_unshifted, _ = po.lj_pair(eps, sig, rc, False, [rc, 0.0, 0.0])
    assert u_unshifted == pytest.approx(4 * (rc ** -12 - rc ** -6), rel=1e-14)


@pytest.mark.parametrize("r", [0.95, 1.0, 1.12, 1.5, 2.2, 2.49])
def test_pair_force_is_minus_gradient(r):
    eps, sig, rc = 1.0, 1.0, 2.5
    h = 1e-6
    up, _ = po.lj_pair(eps, sig, rc, True, [r + h, 0.0, 0.0])
    um, _ = po.lj_pair(eps, sig, rc, True, [r - h, 0.0, 0.0])
    _, f = po.lj_pair(eps, sig, rc, True, [r, 0.0, 0.0])
    fr = -(up - um) / (2 * h)
    # f is the force on j for rij = (r,0,0): F = -dU/dr along +x
    assert f[0] == pytest.approx(fr, rel=1e-7, abs=1e-9)


def test_pair_matches_textbook_formula_random():
    rng = np.random.default_rng(0)
    for _ in range(200):
        eps, sig = rng.uniform(0.1, 2.0), rng.uniform(0.5, 3.5)
        rc = 2.5 * sig
        d = rng.normal(size=3)
        d *= rng.uniform(0.85 * sig, rc) / np.linalg.norm(d)
        u, f = po.lj_pair(eps, sig, rc, True, d)
        r = np.linalg.norm(d)
        u_ref, fr = lj_np(r, eps, sig, rc)
        assert u == pytest.approx(u_ref, rel=1e-12, abs=1e-12 * eps)
        np.testing.assert_allclose(f, fr * d / r, rtol=1e-12, atol=1e-12 * eps / sig)


# ---------------------------------------------------------------------- box
# simulation_box.rs:12-69


def test_box_inverse_diagonal_within_one_ulp_of_reciprocal():
    for L in [52.6, 54.1, 107.4942, 167.9596, 5.975, 1.0, 3.0]:
        h = [L, 0, 0, 0, L * 1.1, 0, 0, 0, L * 0.9]
        hi = po.box_inverse(h)
        for k, Lk in enumerate([L, L * 1.1, L * 0.9]):
            assert abs(hi[4 * k] - 1.0 / Lk) <= np.spacing(1.0 / Lk)
        assert all(hi[q] == 0.0 for q in (1, 2, 3, 5, 6, 7))


def test_box_inverse_triclinic_is_an_inverse():
    rng = np.random.default_rng(1)
    for _ in range(50):
        h = np.zeros(9)
        h[0], h[4], h[8] = rng.uniform(5, 50, size=3)
        h[3], h[6], h[7] = rng.uniform(-2, 2, size=3)
        hi = po.box_inverse(h)
        H = h.reshape(3, 3).T  # column-major -> matrix
        Hi = np.asarray(hi).reshape(3, 3).T
        np.testing.assert_allclose(H @ Hi, np.eye(3), atol=1e-14)


def test_product_and_oracle_box_inverse_agree_bit_for_bit():
    """pis_amd.box (the product's host side) and the oracle restate the same nalgebra
    formula independently; they must produce identical bits."""
    rng = np.random.default_rng(2)
    for _ in range(100):
        h = np.zeros(9)
        h[0], h[4], h[8] = rng.uniform(1, 200, size=3)
        if rng.random() < 0.5:
            h[3], h[6], h[7] = rng.uniform(-3, 3, size=3)
        a = np.array(pbox.box_inverse(h))
        b = po.box_inverse(h)
        assert a.tobytes() == b.tobytes()


def _box(L):
    return po.make_box([L, 0, 0, 0, L, 0, 0, 0, L])


def test_min_image_closed_form_orthorhombic():
    L = 10.0
    b = _box(L)
    rng = np.random.default_rng(3)
    for _ in range(500):
        d = rng.uniform(-3 * L, 3 * L, size=3)
        got = po.min_image(b, d)
        s = d * b.hinv[0]
        want = L * (s - np.sign(s) * np.floor(np.abs(s) + 0.5))
        np.testing.assert_array_equal(got, want)
        assert np.all(np.abs(got) <= L / 2 + 1e-12)


def test_min_image_rounds_half_away_from_zero():
    # Rust f64::round: half away from zero, so +L/2 -> -L/2 and -L/2 -> +L/2
    b = _box(8.0)
    assert po.min_image(b, [4.0, 0.0, 0.0])[0] == -4.0
    assert po.min_image(b, [-4.0, 0.0, 0.0])[0] == 4.0
    assert po.min_image(b, [12.0, 0.0, 0.0])[0] == -4.0  # 1.5 rounds to 2


def test_min_image_norm_is_sign_symmetric():
    b = _box(7.3)
    rng = np.random.default_rng(4)
    for _ in range(300):
        d = rng.uniform(-20, 20, size=3)
        assert po.orc_norm_py(po.min_image(b, d)) == po.orc_norm_py(po.min_image(b, -d))


def test_wrap_into_box_and_periodicity():
    L = 9.0
    b = _box(L)
    rng = np.random.default_rng(5)
    for _ in range(300):
        r = rng.uniform(-30, 30, size=3)
        w = po.wrap(b, r)
        assert np.all(w >= 0.0) and np.all(w <= L)
        np.testing.assert_allclose(np.mod(w - r + L / 2, L) - L / 2, 0.0, atol=1e-12)


def test_wrap_tiny_negative_maps_to_box_length():
    # s = -tiny: floor(s) = -1, s - floor(s) rounds to 1.0, r = h s = L (the reference quirk)
    b = _box(5.0)
    assert po.wrap(b, [-1e-300, 1.0, 1.0])[0] == 5.0


def test_divide_into_cells_matches_reference_examples():
    # neighbour_list.rs:45-57 on the reference's example boxes
    assert po.divide_into_cells(_box(52.6), 8.5125) == (6, 6, 6)   # C1 reference grid
    assert po.divide_into_cells(_box(54.1), 8.5) == (6, 6, 6)      # example/argon4000.txt
    assert po.divide_into_cells(_box(27.05), 8.5) == (3, 3, 3)     # example/example.txt
    assert po.divide_into_cells(_box(5.0), 8.5) == (1, 1, 1)       # example/data.txt


# ---------------------------------------------------------------------- managers
# lennard_jones.rs:62-244, 340-456


def brute_numpy(spec):
    """Independent vectorised all-pairs restatement (LJManager semantics)."""
    x = spec.positions
    n = x.shape[0]
    L = spec.box.lengths
    d = x[None, :, :] - x[:, None, :]
    d = d - L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    nt = len(spec.masses)
    eps = np.zeros((nt, nt))
    sig = np.ones((nt, nt))
    rc = np.zeros((nt, nt))
    for (ti, tj, e, s, c) in spec.pairs:
        for a, b in ((ti - 1, tj - 1), (tj - 1, ti - 1)):
            eps[a, b], sig[a, b], rc[a, b] = e, s, c
    t = spec.type_ids.astype(int) - 1
    E, S, RC = eps[t][:, t], sig[t][:, t], rc[t][:, t]
    inside = r <= RC
    u, fr = lj_np(np.where(inside, r, 1.0), E, S, RC)
    u = np.where(inside, u, 0.0)
    fr = np.where(inside, fr, 0.0)
    f = -(fr / np.where(inside, r, 1.0))[:, :, None] * d  # force on i from j: -F_ij
    return 0.5 * u.sum(), f.sum(axis=1)


@pytest.mark.parametrize("seed", [1, 2])
def test_managers_agree_with_independent_numpy(seed):
    spec = W.lj_liquid(5, 0.8442, 0.72, seed=seed)
    pe_np, f_np = brute_numpy(spec)
    o = po.OSystem.from_spec(spec)
    rms = np.sqrt((f_np ** 2).sum(1).mean())
    for mgr in (po.MGR_OFFSET, po.MGR_BRUTE, po.MGR_VERLET27, po.MGR_FULLLIST):
        pe = o.compute(mgr)
        assert pe == pytest.approx(pe_np, rel=1e-12)
        assert np.abs(o.force - f_np).max() <= 1e-11 * rms


def test_managers_agree_on_binary_mixture():
    spec = W.ka_mixture(1500, 1.2, 1.0, seed=4)
    pe_np, f_np = brute_numpy(spec)
    o = po.OSystem.from_spec(spec)
    rms = np.sqrt((f_np ** 2).sum(1).mean())
    for mgr in (po.MGR_OFFSET, po.MGR_BRUTE, po.MGR_VERLET27):
        pe = o.compute(mgr)
        assert pe == pytest.approx(pe_np, rel=1e-12)
        assert np.abs(o.force - f_np).max() <= 1e-11 * rms


def test_offset_manager_overcounts_with_one_cell_reference_quirk():
    """With a 1x1x1 grid all 14 forward offsets hit the self cell, so every pair is
    visited 14 times (lennard_jones.rs:200-210).  The oracle keeps the quirk."""
    spec = W.read_lammps_data(os.path.join(GOLDEN, "data.txt"))
    o = po.OSystem.from_spec(spec, velocities=False)
    pe_brute = o.compute(po.MGR_BRUTE)
    f_brute = o.force.copy()
    pe_off = o.compute(po.MGR_OFFSET)
    assert pe_off == pytest.approx(14 * pe_brute, rel=1e-13)
    np.testing.assert_allclose(o.force, 14 * f_brute, rtol=1e-13)


def test_verlet27_is_correct_for_small_grids():
    spec = W.ka_mixture(300, 1.2, 1.0, seed=7)  # 2 cells per side at rc_max 2.5
    o = po.OSystem.from_spec(spec)
    assert po.divide_into_cells(o.s.box, 2.5)[0] == 2
    pe_b = o.compute(po.MGR_BRUTE)
    fb = o.force.copy()
    pe_v = o.compute(po.MGR_VERLET27)
    assert pe_v == pytest.approx(pe_b, rel=1e-13)
    np.testing.assert_allclose(o.force, fb, rtol=0, atol=1e-11 * np.abs(fb).max())


def test_newtons_third_law():
    spec = W.lj_liquid(6, 0.8442, 1.0, seed=9)
    o = po.OSystem.from_spec(spec)
    o.compute(po.MGR_OFFSET)
    rms = np.sqrt((o.force ** 2).sum(1).mean())
    assert np.abs(o.force.sum(axis=0)).max() <= 1e-11 * rms * np.sqrt(spec.n)


def fcc_lattice_energy_per_atom(a, eps, sig, rc, shift=True):
    """Half the sum of U over FCC lattice vectors within rc (independent shell sum)."""
    m = int(np.ceil(rc / a)) + 1
    rng = np.arange(-m, m + 1)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    cells = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    R = a * (cells[:, None, :] + basis[None]).reshape(-1, 3)
    r = np.linalg.norm(R, axis=1)
    r = r[(r > 0) & (r <= rc)]
    u, _ = lj_np(r, eps, sig, rc, shift)
    return 0.5 * u.sum()


def check_lattice(o, spec, a, eps, sig, rc):
    """Per-atom energies summed exactly (fsum) equal the independent shell sum to 1e-13;
    the shipped manager's serial sum (lennard_jones.rs:229) stays within its own
    recursive-summation bound (its rounding errors do not cancel on a perfect lattice)."""
    want = fcc_lattice_energy_per_atom(a, eps, sig, rc)
    _, pe_atom = o.compute_fulllist(0.0, True, brute=False)
    assert math.fsum(pe_atom) / 2.0 / spec.n == pytest.approx(want, rel=1e-13)
    pe = o.compute(po.MGR_OFFSET)
    nterms = int(o.list(0.0)[0][-1]) // 2
    assert abs(pe / spec.n - want) <= nterms * 2.0 ** -53 * abs(want)
    assert np.abs(o.force).max() <= 1e-12 * 24 * eps / sig


def test_perfect_fcc_lattice_forces_vanish_and_energy_is_lattice_sum():
    spec = W.argon_fcc_c1()
    o = po.OSystem.from_spec(spec)
    eps, sig, rc = spec.pairs[0][2:]
    check_lattice(o, spec, 5.26, eps, sig, rc)


def test_reference_argon4000_file_is_a_perfect_fcc():
    spec = W.read_lammps_data(os.path.join(GOLDEN, "argon4000.txt"))
    assert spec.n == 4000 and spec.box.lengths[0] == 54.1
    o = po.OSystem.from_spec(spec, velocities=False)
    eps, sig, rc = spec.pairs[0][2:]
    assert (eps, sig, rc) == (0.238, 3.405, 8.5)
    check_lattice(o, spec, 5.41, eps, sig, rc)


def test_data_file_cross_coefficient_misparse_is_reproduced():
    """`1 2 0.6 1.1 2.8` in PairCoeffs parses "2" as epsilon and overwrites (1,1)
    (commands.rs:271-278); the (1,2) pair stays undefined."""
    spec = W.read_lammps_data(os.path.join(GOLDEN, "data_run.txt"))
    table = {}
    for (ti, tj, e, s, c) in spec.pairs:
        table[(min(ti, tj), max(ti, tj))] = (e, s, c)
    assert table[(1, 1)] == (2.0, 0.6, 1.1)
    assert table[(2, 2)] == (0.8, 1.2, 3.0)
    assert (1, 2) not in table


def test_permutation_invariance():
    spec = W.ka_mixture(800, 1.2, 1.0, seed=11)
    o = po.OSystem.from_spec(spec)
    pe = o.compute(po.MGR_OFFSET)
    f = o.force.copy()
    perm = np.random.default_rng(12).permutation(spec.n)
    o2 = po.OSystem(spec.positions[perm], spec.type_ids[perm], spec.masses, spec.pairs, spec.box.h)
    pe2 = o2.compute(po.MGR_OFFSET)
    assert pe2 == pytest.approx(pe, rel=1e-12)
    np.testing.assert_allclose(o2.force, f[perm], rtol=0, atol=1e-11 * np.abs(f).max())


def test_translation_by_lattice_vector_and_shift_invariance():
    spec = W.lj_liquid(5, 0.8442, 0.9, seed=13)
    o = po.OSystem.from_spec(spec)
    pe = o.compute(po.MGR_BRUTE)
    f = o.force.copy()
    L = spec.box.lengths
    moved = W.wrap_positions(spec.positions + np.array([0.37, -1.21, 2.5]) + L, L)
    o2 = po.OSystem(moved, spec.type_ids, spec.masses, spec.pairs, spec.box.h)
    pe2 = o2.compute(po.MGR_BRUTE)
    assert pe2 == pytest.approx(pe, rel=1e-12)
    np.testing.assert_allclose(o2.force, f, atol=1e-10 * np.abs(f).max())


def test_full_list_builder_equals_brute_force_set():
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=14)
    o = po.OSystem.from_spec(spec)
    a = po.csr_to_sorted_pairs(*o.list(0.3))
    b = po.csr_to_sorted_pairs(*o.list(0.3, brute=True))
    assert np.array_equal(a, b)
    # symmetric, no self pairs
    assert np.all(a[:, 0] != a[:, 1])
    sw = a[:, ::-1]
    sw = sw[np.lexsort((sw[:, 1], sw[:, 0]))]
    assert np.array_equal(a, sw)


def test_full_list_builder_duplicates_with_fewer_than_three_cells():
    """build_neighbour_list clears its dedup set inside the dx loop (lennard_jones.rs:410),
    so with 2 cells in x a neighbour cell is visited twice and entries repeat."""
    spec = W.ka_mixture(300, 1.2, 1.0, seed=15)
    o = po.OSystem.from_spec(spec)
    st, ix = o.list(0.0)
    stb, ixb = o.list(0.0, brute=True)
    assert st[-1] > stb[-1]
    got = set(map(tuple, po.csr_to_sorted_pairs(st, ix).tolist()))
    assert got == set(map(tuple, po.csr_to_sorted_pairs(stb, ixb).tolist()))


# ---------------------------------------------------------------------- integrator
# potential.rs:15-33, properties.rs:9-65


def test_free_flight_is_wrapped_ballistic_motion():
    spec = W.lj_liquid(3, 0.2, 1.0, seed=16)
    spec.pairs = [(1, 1, 0.0, 1.0, 2.5)]  # eps = 0: no forces
    o = po.OSystem.from_spec(spec)
    x0 = o.pos.copy()
    v0 = o.vel.copy()
    dt, n = 0.01, 50
    o.compute(po.MGR_BRUTE)
    for _ in range(n):
        o.verlet_step_nve(dt, po.MGR_BRUTE)
    L = spec.box.lengths
    want = W.wrap_positions(x0 + v0 * dt * n, L)
    diff = (o.pos - want + L / 2) % L - L / 2
    assert np.abs(diff).max() <= 1e-11
    np.testing.assert_array_equal(o.vel, v0)


def test_time_reversal_returns_to_start():
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=17)
    o = po.OSystem.from_spec(spec)
    x0 = o.pos.copy()
    o.compute(po.MGR_OFFSET)
    for _ in range(100):
        o.verlet_step_nve(0.002, po.MGR_OFFSET)
    o.vel[:] = -o.vel
    for _ in range(100):
        o.verlet_step_nve(0.002, po.MGR_OFFSET)
    L = spec.box.lengths
    diff = (o.pos - x0 + L / 2) % L - L / 2
    assert np.abs(diff).max() <= 1e-9


def test_energy_error_scales_as_dt_squared():
    # at least 3 reference cells per side: below that the shipped manager's stencil
    # over-counts (see the quirk test above) and energy is not conserved at all
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=18)
    errs = []
    for dt in (0.004, 0.002):
        o = po.OSystem.from_spec(spec)
        th = o.run_nve(dt, int(round(0.4 / dt)), po.MGR_OFFSET, kb=1.0)
        errs.append(np.abs(th[:, 2] - th[0, 2]).max())
    assert 2.5 < errs[0] / errs[1] < 6.0


def test_reductions_closed_forms():
    spec = W.lj_liquid(2, 0.5, 1.0, seed=19)
    n = spec.n
    o = po.OSystem.from_spec(spec)
    o.vel[:] = [0.3, -0.4, 1.2]
    m = spec.masses[0]
    ke = o.kinetic_energy()
    assert ke == pytest.approx(0.5 * m * n * (0.09 + 0.16 + 1.44), rel=1e-14)
    kb = 0.0083144621
    assert o.temperature(ke, kb) == pytest.approx(2 * ke / (3 * n * kb), rel=1e-15)
    o.force[:] = 0.0
    # ideal gas: P = 2 KE / (3 V)
    assert o.pressure(ke) == pytest.approx(2 * ke / (3 * spec.box.volume()), rel=1e-14)


def test_run_matches_step_by_step_and_reports_thermo():
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=20)
    a = po.OSystem.from_spec(spec)
    th = a.run_nve(0.005, 20, po.MGR_OFFSET, kb=1.0)
    b = po.OSystem.from_spec(spec)
    b.compute(po.MGR_OFFSET)
    for t in range(20):
        pe = b.verlet_step_nve(0.005, po.MGR_OFFSET)
        ke = b.kinetic_energy()
        assert th[t, 0] == pe and th[t, 1] == ke and th[t, 2] == pe + ke
    np.testing.assert_array_equal(a.pos, b.pos)


# ---------------------------------------------------------------------- NVT chain
# ensemble/nvt.rs:4-119, potential.rs:35-58, simulation.rs:53-57,100-105


def _nhc(t_start, t_end, tau, chain=3, kb=1.0):
    import ctypes as C

    c = po.NHC()
    assert po.lib().orc_nhc_new(C.byref(c), t_start, t_end, tau, chain, kb) == 0
    return c


def test_chain_masses_follow_the_relaxation_time():
    # Q_1 = N_f-independent kB T tau^2 of the chain paper, later links 10x lighter each
    c = _nhc(1.3, 1.3, 0.4, chain=3, kb=0.5)
    q0 = 0.5 * 1.3 * 0.16
    assert list(c.q)[:3] == pytest.approx([q0, q0 / 10.0, q0 / 100.0], rel=1e-15)
    assert c.target == 1.3 and list(c.xi)[:3] == [0.0, 0.0, 0.0]


def test_chain_force_vanishes_at_the_target_kinetic_energy():
    import ctypes as C

    n, T = 500, 1.7
    c = _nhc(T, T, 0.5)
    po.lib().orc_nhc_compute_forces(C.byref(c), 1.5 * n * T, n)  # KE = (3N/2) kB T
    assert c.g[0] == pytest.approx(0.0, abs=1e-12)
    # idle links feel -kB T (their upstream velocity is zero)
    assert c.g[1] == -T and c.g[2] == -T


def test_chain_cools_a_hot_system_and_heats_a_cold_one():
    import ctypes as C

    n = 100
    for ke, sign in ((1.5 * n * 2.0, 1.0), (1.5 * n * 0.5, -1.0)):
        c = _nhc(1.0, 1.0, 0.5)
        po.lib().orc_nhc_compute_forces(C.byref(c), ke, n)
        po.lib().orc_nhc_propagate_half_step(C.byref(c), 0.005)
        # xi_0 > 0 scales velocities down (exp(-dt xi/2) < 1), xi_0 < 0 scales them up
        assert np.sign(c.xi[0]) == sign


def test_target_ramp_is_linear_in_the_finished_step():
    import ctypes as C

    c = _nhc(1.0, 2.0, 0.5)
    for i, want in ((0, 1.0), (5, 1.5), (9, 1.9)):
        po.lib().orc_nhc_target(C.byref(c), i, 10)
        assert c.target == pytest.approx(want, rel=1e-15)


@pytest.mark.parametrize("t_target", [1.4, 0.5])
def test_nvt_run_relaxes_to_the_target_temperature(t_target):
    spec = W.lj_liquid(6, 0.8442, 0.9, seed=31)
    o = po.OSystem.from_spec(spec)
    th = o.run_nvt(0.004, 1500, t_target, t_target, 0.2, 3, kb=1.0)
    t_late = th[-500:, 3].mean()
    assert abs(t_late - t_target) < 0.06 * t_target


def test_nvt_with_a_stiff_chain_reduces_to_nve():
    # tau -> infinity: Q huge, xi -> 0, both scalings -> 1
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=32)
    a = po.OSystem.from_spec(spec)
    b = po.OSystem.from_spec(spec)
    th_nvt = a.run_nvt(0.005, 50, 0.72, 0.72, 1e9, 3, kb=1.0)
    th_nve = b.run_nve(0.005, 50, po.MGR_OFFSET, kb=1.0)
    assert np.abs(th_nvt[:, 0] - th_nve[:, 0]).max() <= 1e-9 * np.abs(th_nve[:, 0]).max()
    assert np.abs(a.pos - b.pos).max() <= 1e-12


# ---- MTK barostat / NPT (ensemble/npt.rs, potential.rs:112-135) ---------------------


@pytest.mark.parametrize("scale", [1e-6, 0.3, 4.0])
def test_matrix_exponential_matches_scipy(scale):
    import scipy.linalg as sl

    rng = np.random.default_rng(7)
    for _ in range(20):
        a = rng.normal(size=(3, 3)) * scale
        want = sl.expm(a)
        # a large norm (scaling and squaring) loses digits to the problem's conditioning
        tol = 1e-14 if scale < 1.0 else 1e-12
        assert np.abs(po.mat3_exp(a) - want).max() <= tol * max(1.0, np.abs(want).max())
    assert np.array_equal(po.mat3_exp(np.zeros((3, 3))), np.eye(3))
    d = np.diag([0.1, -0.2, 0.3])
    assert np.allclose(po.mat3_exp(d), np.diag(np.exp([0.1, -0.2, 0.3])), rtol=1e-15, atol=0)


def test_pressure_tensors_against_numpy_sums():
    import ctypes as C

    spec = W.lj_box(200, 0.8, 1.0, seed=41)
    spec.masses = np.ones_like(spec.masses)
    o = po.OSystem.from_spec(spec)
    o.compute(po.MGR_OFFSET)
    kt, wt, pt = o.tensor("kinetic"), o.tensor("virial"), o.tensor("pressure")
    assert np.allclose(kt, o.vel.T @ o.vel, rtol=1e-12, atol=1e-12)     # no masses (quirk)
    assert np.allclose(wt, o.pos.T @ o.force, rtol=1e-12, atol=1e-10)
    vol = abs(np.linalg.det(o.box_h()))
    assert np.allclose(pt, (kt + wt) / vol, rtol=1e-12, atol=1e-14)
    # unit masses: tr(P) = 3 P with P = (2 KE + tr W) / (3 V)
    ke = o.kinetic_energy()
    assert np.trace(pt) == pytest.approx(3.0 * po.lib().orc_pressure(C.byref(o.s), ke), rel=1e-12)


def test_scale_box_keeps_fractional_coordinates():
    spec = W.lj_box(300, 0.8, 1.0, seed=42)
    o = po.OSystem.from_spec(spec)
    h0 = o.box_h()
    s0 = o.pos @ np.linalg.inv(h0).T
    scale = po.mat3_exp(np.array([[1e-3, 2e-4, 0.0], [2e-4, -5e-4, 1e-4], [0.0, 1e-4, 2e-3]]))
    o.scale_box(scale)
    h1 = o.box_h()
    assert np.allclose(h1, scale @ h0, rtol=1e-15, atol=1e-14)
    assert np.allclose(o.pos @ np.linalg.inv(h1).T, s0, rtol=0, atol=1e-13)
    assert np.allclose(np.array(o.s.box.hinv[:]).reshape(3, 3).T @ h1, np.eye(3), atol=1e-15)


def test_npt_with_a_heavy_barostat_reduces_to_nvt():
    # tau_p -> infinity: w huge, the barostat scalings are the identity to rounding
    spec = W.lj_liquid(5, 0.8442, 0.72, seed=43)
    a = po.OSystem.from_spec(spec)
    b = po.OSystem.from_spec(spec)
    th_npt, hb = a.run_npt(0.004, 40, 0.72, 0.72, 0.5, 0.0, 1e12, 3, kb=1.0)
    th_nvt = b.run_nvt(0.004, 40, 0.72, 0.72, 0.5, 3, kb=1.0)
    assert np.abs(th_npt[:, :2] - th_nvt[:, :2]).max() <= 1e-10 * np.abs(th_nvt[:, 0]).max()
    assert np.abs(hb[-1] - np.asarray(spec.box.h).ravel(order="F")).max() <= 1e-12


@pytest.mark.parametrize("p_target, grows", [(0.0, False), (6.0, True)])
def test_npt_volume_follows_the_pressure_difference(p_target, grows):
    # the instantaneous pressure of this liquid is ~1.4: a lower target lets the box expand
    # (P - P0 > 0 pushes the momentum up), a higher one compresses it.  The barostat mass
    # uses the constant KB = 0.0083 (npt.rs:32), so in these units it is light.
    spec = W.lj_liquid(5, 0.8442, 0.9, seed=44)
    o = po.OSystem.from_spec(spec)
    v0 = abs(np.linalg.det(o.box_h()))
    th, hb = o.run_npt(0.004, 100, 0.9, 0.9, 0.5, p_target, 10.0, 3, kb=1.0)
    v1 = abs(np.linalg.det(hb[-1].reshape(3, 3).T))
    assert (v1 > v0) == (not grows)
    assert np.isfinite(th).all()
    # the box picks up tilt from the symmetrised off-diagonal pressure (about 1 % of the
    # edge for 500 atoms in 100 steps): NPT needs a general box matrix
    h = hb[-1].reshape(3, 3).T
    assert 0.0 < np.abs(h - np.diag(np.diag(h))).max() < 5e-2 * np.abs(np.diag(h)).min()


# ------------------------------------------------------- list digest, subset forces
# (helpers for parity checks at sizes where whole lists or whole-system oracle runs are
# too large: each is pinned here against the stored list / full computation it replaces)


def test_mix64_is_splitmix64():
    # the first two outputs of SplitMix64 seeded with 0 (state advanced by the golden gamma
    # before mixing): published test vectors of the generator
    assert po.lib().orc_mix64(0) == 0xE220A8397B1DCDAF
    assert po.lib().orc_mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    xs = np.array([0, 1, 2, 12345, 2 ** 40 + 3], dtype=np.uint64)
    assert [int(v) for v in po.mix64(xs)] == [po.lib().orc_mix64(int(x)) for x in xs]


@pytest.mark.parametrize("name", ["lj_6", "ka_1500", "thin_400", "example"])
def test_list_digest_equals_the_stored_lists_digest(name):
    spec = {
        "lj_6": lambda: W.lj_liquid(6, 0.8442, 0.72, seed=3),
        "ka_1500": lambda: W.ka_mixture(1500, 1.2, 1.0, seed=2),
        # fewer than 3 cells per side: rows carry the reference's duplicate entries
        "thin_400": lambda: W.ka_mixture(400, 1.2, 1.0, seed=4),
        "example": lambda: _example_spec(),
    }[name]()
    o = po.OSystem.from_spec(spec)
    for add in (0.0, spec.skin):
        count, h = o.list_digest(add)
        st, ix = o.list(add)
        c2, h2 = po.csr_digest(st, ix)
        assert np.array_equal(count, c2)
        assert np.array_equal(h, h2)
    if name == "thin_400":
        assert po.divide_into_cells(o.s.box, max(p[4] for p in spec.pairs) + spec.skin)[0] < 3


def test_list_digest_detects_a_changed_row():
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=3)
    o = po.OSystem.from_spec(spec)
    st, ix = o.list(spec.skin)
    c0, h0 = po.csr_digest(st, ix)
    ix2 = ix.copy()
    ix2[st[5]] = (ix2[st[5]] + 1) % spec.n  # one partner swapped for another atom
    c1, h1 = po.csr_digest(st, ix2)
    assert np.array_equal(c0, c1)
    assert h1[5] != h0[5] and np.array_equal(np.delete(h1, 5), np.delete(h0, 5))


def _example_spec():
    # the reference's example/example.txt: 3 cells per side, the smallest grid without the
    # stencil's duplicates
    return W.read_lammps_data(os.path.join(GOLDEN, "example.txt"))


@pytest.mark.parametrize("seed", [1, 2])
def test_forces_subset_equals_the_full_list_gather(seed):
    spec = W.lj_liquid(6, 0.8442, 0.72, seed=seed)
    o = po.OSystem.from_spec(spec)
    pe_full, pe_atom = o.compute_fulllist(0.0, False)
    f_full = o.force.copy()
    idx = np.random.default_rng(seed).choice(spec.n, 97, replace=False)
    f, pe = o.forces_subset(idx)
    assert np.array_equal(f, f_full[idx])  # same rows, same order of summation
    assert np.array_equal(pe, pe_atom[idx])
    # and the shipped half-list manager agrees to rounding
    o.compute(po.MGR_OFFSET)
    scale = np.sqrt((o.force ** 2).sum(1).mean())
    assert np.abs(f - o.force[idx]).max() <= 1e-12 * scale


def test_forces_subset_on_a_mixture_matches_brute_force():
    spec = W.ka_mixture(1500, 1.2, 1.0, seed=6)
    o = po.OSystem.from_spec(spec)
    o.compute(po.MGR_BRUTE)
    idx = np.arange(0, spec.n, 7)
    f, _ = o.forces_subset(idx)
    scale = np.sqrt((o.force ** 2).sum(1).mean())
    assert np.abs(f - o.force[idx]).max() <= 1e-12 * scale


def test_pair_virial_against_numpy_brute_force():
    spec = W.ka_mixture(1500, 1.2, 1.0, seed=6)
    o = po.OSystem.from_spec(spec)
    L = np.asarray(spec.box.lengths)
    pos = spec.positions
    types = np.asarray(spec.type_ids)
    eps = {(1, 1): 1.0, (1, 2): 1.5, (2, 2): 0.5}
    sig = {(1, 1): 1.0, (1, 2): 0.8, (2, 2): 0.88}
    tab = {(min(a, b), max(a, b)): (e, s, rc) for (a, b, e, s, rc) in spec.pairs}
    w = 0.0
    for i in range(spec.n - 1):
        d = pos[i + 1:] - pos[i]
        d -= L * np.round(d / L)
        r2 = (d ** 2).sum(1)
        for j_off in range(d.shape[0]):
            j = i + 1 + j_off
            key = (min(types[i], types[j]), max(types[i], types[j]))
            e, s, rc = tab[key]
            if r2[j_off] <= rc * rc:
                sr6 = (s * s / r2[j_off]) ** 3
                w += 24 * e * (2 * sr6 * sr6 - sr6)  # r . F for the LJ pair
    assert eps and sig  # the KA table above documents the mixture
    assert o.pair_virial() == pytest.approx(w, rel=1e-11)
