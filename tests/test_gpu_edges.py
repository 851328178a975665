"""GPU edge cases through the C ABI, each checked against the oracle and against a closed
form that does not come from the oracle:

* exact ties.  A simple cubic lattice of integer coordinates (spacing 1 sigma, box 12 sigma)
  with one vacancy: every pair distance is the square root of an integer, so many pairs sit
  exactly on the force cutoff (r = 3: (3,0,0), (2,2,1), ...) and on the list radius
  (rc + skin = 4: (4,0,0)).  The reference keeps a pair unless |r_ij| > rc
  (lennard_jones.rs:404), so a pair at exactly rc counts -- when its vector comes out of the
  minimum image exactly.  The image goes through fractional coordinates
  (simulation_box.rs:17-27), and a component of 10 or 11 sigma in this box comes back as
  -2.000000000000001 or -1.0000000000000004: some ties across the periodic boundary land
  just past rc and drop out, so atoms near the box faces feel a net force.  The FP64 path
  must reproduce that bit for bit.  In exact arithmetic every force on the full lattice
  vanishes by inversion symmetry, so an atom's force is minus the pair force the missing
  atom would have exerted (nonzero at exactly r = rc), zero beyond rc: the oracle and the
  FP64 path meet that closed form wherever no boundary tie is involved, and the mixed path,
  which decides the cutoff in FP32 on tile-relative differences (exact here), meets it on
  every atom.  A list row holds every lattice vector of length <= 4 (the count of integer
  points in a ball of radius 4, by enumeration), one fewer next to the vacancy.  The
  integer coordinates put many atoms exactly on cell faces, where binning and the x windows
  of the list build meet their edge cases.
* one atom (no pairs: zero force and PE, free flight) and two atoms (one pair: the LJ force
  and the shifted energy U(r) - U(rc) in closed form).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import Engine, workloads as W
from tests.conftest import force_err

pytestmark = pytest.mark.gpu

L_LATTICE, VACANCY, RC, SKIN = 12, (5, 5, 5), 3.0, 1.0


def _vacancy_lattice():
    g = np.arange(L_LATTICE, dtype=np.float64)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    keep = ~np.all(pos == np.array(VACANCY, dtype=np.float64), axis=1)
    pos = np.ascontiguousarray(pos[keep])
    n = pos.shape[0]
    rng = np.random.default_rng(5)
    return W.SystemSpec(positions=pos, type_ids=np.ones(n, dtype=np.uint32), masses=np.array([1.0]),
                        pairs=[(1, 1, 1.0, 1.0, RC)], box=W.SimulationBox.cubic(float(L_LATTICE)),
                        velocities=W.start_velocities(np.ones(n), 0.5, 1.0, rng), skin=SKIN, kb=1.0,
                        dt=0.005, steps=10, name="sc_vacancy")


def _lj_force_on_i(d):
    """Force on atom i from atom j at d = r_i - r_j (eps = sigma = 1): -dU/dr along d."""
    r2 = float(np.dot(d, d))
    s6 = r2 ** -3
    return 24.0 * (2.0 * s6 * s6 - s6) / r2 * np.asarray(d, dtype=np.float64)


def _index_of(pos, p):
    hit = np.flatnonzero(np.all(pos == np.asarray(p, dtype=np.float64), axis=1))
    assert hit.size == 1
    return int(hit[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_exactly_on_the_cutoff_and_the_list_radius(precision):
    spec = _vacancy_lattice()
    print(f"sc_vacancy {precision}: N={spec.n}, rc {RC}, skin {SKIN}", flush=True)
    with Engine(spec, precision) as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
        off, idx = eng.neighbor_list()
    o = po.OSystem.from_spec(spec)
    pe_ref = o.compute(po.MGR_OFFSET)
    f_ref = o.force.copy()
    st, ix = o.list(rcut_add=SKIN)
    # the pair set at rc + skin, ties included, bit for bit
    assert np.array_equal(po.csr_to_sorted_pairs(off, idx), po.csr_to_sorted_pairs(st, ix))
    # row lengths: integer vectors of length <= 4, one fewer for the vacancy's neighbours
    g = np.arange(-4, 5)
    v = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    full_row = int(np.count_nonzero((v * v).sum(1) <= 16)) - 1
    d_vac = spec.positions - np.array(VACANCY, dtype=np.float64)
    near = (d_vac * d_vac).sum(1) <= (RC + SKIN) ** 2
    rows = np.diff(np.asarray(off, dtype=np.int64))
    assert np.all(rows[~near] == full_row) and np.all(rows[near] == full_row - 1)
    # forces: minus the missing atom's pair force within rc (ties count), zero beyond it.
    # FP32 pair sums cancel the lattice's symmetric pairs only to the rounding of the
    # largest pair force (24 at r = sigma), hence the mixed bar of 1e-5 of that force.
    f_nn = np.linalg.norm(_lj_force_on_i(np.array([1.0, 0.0, 0.0])))
    f_tie = np.linalg.norm(_lj_force_on_i(np.array([3.0, 0.0, 0.0])))
    tol = 1e-12 if precision == "fp64" else 1e-5 * f_nn
    assert f_tie > 10 * tol  # counting the tie pair at exactly r = rc is visible
    d_img = d_vac - L_LATTICE * np.round(d_vac / L_LATTICE)  # exact: integers
    within = (d_img * d_img).sum(1) <= RC * RC
    want = np.zeros_like(f_ref)
    for i in np.flatnonzero(within):
        want[i] = -_lj_force_on_i(d_img[i])
    for p in [(8, 5, 5), (7, 7, 6), (5, 3, 7), (6, 5, 5), (9, 5, 5), (7, 7, 7)]:
        i = _index_of(spec.positions, p)
        assert np.linalg.norm(f_ref[i] - want[i]) <= 1e-12
        assert np.linalg.norm(f[i] - want[i]) <= tol, (p, f[i], want[i])
    # the reference drops some boundary ties: the corner atom feels a net force
    corner = _index_of(spec.positions, (0, 0, 0))
    assert np.linalg.norm(f_ref[corner]) > 10 * tol
    if precision == "fp64":
        assert np.linalg.norm(f[corner] - f_ref[corner]) <= 1e-12
        assert force_err(f, f_ref) <= 1e-10
        assert abs(pe - pe_ref) <= 1e-12 * abs(pe_ref)
    else:
        # every atom on the closed form; U(rc) - U(rc) = 0, so the PE does not see ties
        assert np.abs(f - want).max() <= tol
        assert abs(pe - pe_ref) <= 2e-6 * abs(pe_ref)


def _tiny_box(pos, vel):
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
    n = pos.shape[0]
    return W.SystemSpec(positions=pos, type_ids=np.ones(n, dtype=np.uint32), masses=np.array([1.0]),
                        pairs=[(1, 1, 1.0, 1.0, 2.5)], box=W.SimulationBox.cubic(10.0),
                        velocities=np.ascontiguousarray(np.asarray(vel, dtype=np.float64)), skin=0.3,
                        kb=1.0, dt=0.005, steps=10, name=f"tiny_{n}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_atom_has_no_pairs_and_flies_free(precision):
    spec = _tiny_box([[9.9, 2.0, 0.05]], [[0.3, -0.2, -0.1]])
    dt, nsteps = 0.005, 200
    with Engine(spec, precision) as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        _, _, f0 = eng.get_state(0)
        th = eng.run_nve(dt, nsteps)
        pos, vel, f = eng.get_state(0)
        off, idx = eng.neighbor_list()
    assert pe == 0.0 and np.all(f0 == 0.0) and np.all(f == 0.0)
    assert int(off[-1]) == 0
    assert np.all(vel == spec.velocities)
    # x(t) = wrap(x0 + v t): it crosses the x face and the z face
    want = np.mod(spec.positions + spec.velocities * (dt * nsteps), 10.0)
    assert np.abs(pos - want).max() <= 1e-12
    ke = 0.5 * float((spec.velocities ** 2).sum())
    assert np.allclose(th[:, 0, 1], ke, rtol=1e-15, atol=0.0)
    assert np.all(th[:, 0, 0] == 0.0)
    o = po.OSystem.from_spec(spec)
    o.compute(po.MGR_OFFSET)
    o.run_nve(dt, nsteps)
    assert np.abs(pos - o.pos).max() <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_atoms_one_pair_in_closed_form(precision):
    r = 1.5
    spec = _tiny_box([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    with Engine(spec, precision) as eng:
        eng.load()
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
    u = lambda x: 4.0 * (x ** -12 - x ** -6)  # noqa: E731
    pe_want = u(r) - u(2.5)
    f0 = _lj_force_on_i(np.array([-r, 0.0, 0.0]))
    assert f0[0] > 0.0  # r > 2^(1/6): the pair attracts
    tol = 1e-14 if precision == "fp64" else 1e-6
    assert abs(pe - pe_want) <= tol * abs(pe_want)
    assert np.abs(f[0] - f0).max() <= tol * abs(f0[0])
    assert np.abs(f[1] + f0).max() <= tol * abs(f0[0])
    o = po.OSystem.from_spec(spec)
    pe_ref = o.compute(po.MGR_OFFSET)
    assert abs(pe_ref - pe_want) <= 1e-14 * abs(pe_want)
    assert np.abs(o.force[0] - f0).max() <= 1e-14 * abs(f0[0])
