"""Synthetic inputs for the benchmark configurations (numpy; host side only).

FCC sites follow the reference's generator (src/bin/sapphire/main.rs:17-24, 60-78):
basis (0,0,0), (0,.5,.5), (.5,0,.5), (.5,.5,0) inside loops ix -> iy -> iz, position
a*(cell + basis).  Velocities follow Atoms::start_velocities (src/atoms/velocities.rs:10-59):
a per-atom Gaussian of standard deviation sqrt(kB T / m), centre-of-mass drift removed,
then a rescale to T with 3N degrees of freedom.  The Gaussian stream is numpy's PCG64,
not the reference's rand SmallRng, so velocities are seeded but not the reference's bits.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .box import SimulationBox

KB_REFERENCE = 0.0083144621  # src/constants.rs:3 (kJ/(mol K), the molar gas constant)
EV_TO_KJMOL = 96.4853


@dataclass
class SystemSpec:
    """Everything one simulated system needs (mirrors what System::contextualize
    assembles: Atoms + the pair_coeff table, src/system.rs:100-175)."""

    positions: np.ndarray  # (N, 3) float64
    type_ids: np.ndarray  # (N,) 1-based
    masses: np.ndarray  # (ntypes,)
    pairs: list  # [(ti, tj, eps, sigma, rcut)]
    box: SimulationBox
    velocities: np.ndarray | None = None
    skin: float = 0.3
    kb: float = 1.0
    shift: bool = True
    dt: float = 0.005
    steps: int = 1000
    name: str = "system"
    meta: dict = field(default_factory=dict)
    tile: tuple | None = None  # cells per tile from an autotuner; None: the planner's model

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def fcc_sites(nx: int, ny: int, nz: int, a: float) -> np.ndarray:
    basis = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    return a * frac


def box_lattice(n_atoms: int, L: float) -> np.ndarray:
    """FCC sites (at least n_atoms of them) filling a cubic box of side L.  The cell counts
    per axis may differ by a little: of the shapes holding enough sites, the one with the
    largest nearest-site distance is taken, so that the surplus sites a cubic 4m^3 lattice
    would need (up to ~2x the atoms just above 4m^3) do not squeeze atoms together."""
    m0 = max(1, int(round((n_atoms / 4.0) ** (1.0 / 3.0))))
    best = None
    for mx in range(max(1, m0 - 2), m0 + 3):
        for my in range(max(1, m0 - 2), mx + 1):
            for mz in range(max(1, m0 - 2), my + 1):
                if 4 * mx * my * mz < n_atoms:
                    continue
                a = (L / mx, L / my, L / mz)
                d = min(np.hypot(a[0], a[1]), np.hypot(a[0], a[2]), np.hypot(a[1], a[2])) / 2.0
                key = (d, -4 * mx * my * mz)
                if best is None or key > best[0]:
                    best = (key, (mx, my, mz))
    mx, my, mz = best[1]
    return fcc_sites(mx, my, mz, 1.0) * np.array([L / mx, L / my, L / mz])


def wrap_positions(pos: np.ndarray, lengths) -> np.ndarray:
    L = np.asarray(lengths, dtype=np.float64)
    return pos - L * np.floor(pos / L)


def start_velocities(masses_per_atom: np.ndarray, temperature: float, kb: float, rng) -> np.ndarray:
    """Gaussian velocities, drift removed, rescaled to `temperature` (velocities.rs:10-59)."""
    n = masses_per_atom.shape[0]
    sigma = np.sqrt(kb * temperature / masses_per_atom)
    v = rng.standard_normal((n, 3)) * sigma[:, None]
    total_mass = masses_per_atom.sum()
    vcm = (v * masses_per_atom[:, None]).sum(axis=0) / total_mass
    v = v - vcm
    ke = 0.5 * np.sum(masses_per_atom * np.sum(v * v, axis=1))
    t_cur = 2.0 * ke / (3 * n * kb)
    return v * np.sqrt(temperature / t_cur)


def lj_liquid(
    ncells: int,
    density: float,
    temperature: float,
    seed: int = 12345,
    jitter: float = 0.05,
    rcut: float = 2.5,
    skin: float = 0.3,
    dt: float = 0.005,
    steps: int = 10000,
    name: str = "lj",
) -> SystemSpec:
    """Single-species LJ in reduced units (eps = sigma = m = kB = 1): FCC ncells^3 at the
    given density, each component jittered by U(-jitter, jitter)."""
    rng = np.random.default_rng(seed)
    a = (4.0 / density) ** (1.0 / 3.0)
    L = ncells * a
    pos = fcc_sites(ncells, ncells, ncells, a)
    if jitter:
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    pos = wrap_positions(pos, [L, L, L])
    n = pos.shape[0]
    types = np.ones(n, dtype=np.uint32)
    masses = np.array([1.0])
    vel = start_velocities(np.ones(n), temperature, 1.0, rng)
    return SystemSpec(
        positions=pos,
        type_ids=types,
        masses=masses,
        pairs=[(1, 1, 1.0, 1.0, rcut)],
        box=SimulationBox.cubic(L),
        velocities=vel,
        skin=skin,
        kb=1.0,
        dt=dt,
        steps=steps,
        name=name,
        meta={"density": density, "temperature": temperature, "ncells": ncells},
    )


def argon_fcc_c1(seed: int = 2024) -> SystemSpec:
    """Argon FCC crystal, 4000 atoms (10^3 cells, a = 5.26 A), eps = 0.0104 eV, sigma =
    3.405 A, rcut = 2.5 sigma, skin 0.3 sigma, 100 K, dt = 1 fs, in the reference's own
    units (kJ/mol, g/mol, A; time unit 100 fs, so dt = 0.01)."""
    rng = np.random.default_rng(seed)
    a = 5.26
    sigma = 3.405
    eps = 0.0104 * EV_TO_KJMOL
    pos = fcc_sites(10, 10, 10, a)
    n = pos.shape[0]
    mass = 39.948
    vel = start_velocities(np.full(n, mass), 100.0, KB_REFERENCE, rng)
    return SystemSpec(
        positions=pos,
        type_ids=np.ones(n, dtype=np.uint32),
        masses=np.array([mass]),
        pairs=[(1, 1, eps, sigma, 2.5 * sigma)],
        box=SimulationBox.cubic(10 * a),
        velocities=vel,
        skin=0.3 * sigma,
        kb=KB_REFERENCE,
        dt=0.01,
        steps=1000,
        name="argon_fcc_4000",
        meta={"temperature": 100.0},
    )


KA_PAIRS = [  # Kob-Andersen 80:20 (eps, sigma); rcut = 2.5 sigma_ab
    (1, 1, 1.0, 1.0, 2.5),
    (1, 2, 1.5, 0.8, 2.0),
    (2, 2, 0.5, 0.88, 2.2),
]


def ka_mixture(n_atoms: int, density: float, temperature: float, seed: int, skin: float = 0.3,
               dt: float = 0.005, steps: int = 5000) -> SystemSpec:
    """Kob-Andersen binary mixture: FCC sites of box_lattice with the surplus sites removed
    at random, 80 % type 1 and 20 % type 2 assigned at random."""
    rng = np.random.default_rng(seed)
    L = (n_atoms / density) ** (1.0 / 3.0)
    sites = box_lattice(n_atoms, L)
    keep = np.sort(rng.choice(sites.shape[0], size=n_atoms, replace=False))
    pos = wrap_positions(sites[keep], [L, L, L])
    nb = int(round(0.2 * n_atoms))
    types = np.ones(n_atoms, dtype=np.uint32)
    types[rng.choice(n_atoms, size=nb, replace=False)] = 2
    vel = start_velocities(np.ones(n_atoms), temperature, 1.0, rng)
    return SystemSpec(
        positions=pos,
        type_ids=types,
        masses=np.array([1.0, 1.0]),
        pairs=list(KA_PAIRS),
        box=SimulationBox.cubic(L),
        velocities=vel,
        skin=skin,
        kb=1.0,
        dt=dt,
        steps=steps,
        name=f"ka_{n_atoms}",
        meta={"density": density, "temperature": temperature},
    )


def lj_box(n_atoms: int, density: float, temperature: float, seed: int, rcut: float = 2.5,
           skin: float = 0.3, dt: float = 0.005, steps: int = 1000) -> SystemSpec:
    """Single-species LJ box of any atom count: box_lattice sites, surplus removed at random."""
    rng = np.random.default_rng(seed)
    L = (n_atoms / density) ** (1.0 / 3.0)
    sites = box_lattice(n_atoms, L)
    keep = np.sort(rng.choice(sites.shape[0], size=n_atoms, replace=False))
    pos = wrap_positions(sites[keep], [L, L, L])
    vel = start_velocities(np.ones(n_atoms), temperature, 1.0, rng)
    return SystemSpec(
        positions=pos,
        type_ids=np.ones(n_atoms, dtype=np.uint32),
        masses=np.array([1.0]),
        pairs=[(1, 1, 1.0, 1.0, rcut)],
        box=SimulationBox.cubic(L),
        velocities=vel,
        skin=skin,
        kb=1.0,
        dt=dt,
        steps=steps,
        name=f"lj_{n_atoms}",
        meta={"density": density, "temperature": temperature},
    )


def ka_ensemble(n_systems: int = 512, seed: int = 7, nmin: int = 256, nmax: int = 32768,
                steps: int = 5000):
    """Batched ensemble: n_systems Kob-Andersen systems, atom counts log-uniform on
    [nmin, nmax], density 1.2, T* uniform on [0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_systems)).astype(int)
    temps = rng.uniform(0.5, 1.5, size=n_systems)
    return [ka_mixture(int(n), 1.2, float(t), seed * 1000 + k, steps=steps)
            for k, (n, t) in enumerate(zip(ns, temps))]


def read_lammps_data(path: str) -> SystemSpec:
    """LAMMPS-style data file as the reference reads it (run_read_data,
    src/readers/input_file/commands.rs:147-384; restated in pis_amd.script.read_data):
    header counts, box bounds, Masses, PairCoeffs (a line `i eps sigma [rcut]`, rcut
    defaulting to 2.5 sigma), Atoms and an optional Velocities section.  Zero tilts,
    periodic everywhere.  Velocities are None when the file has no Velocities section."""
    from . import script

    ctx = script.Context()
    script.read_data(path, ctx)
    a = ctx.atoms
    pairs = [(i, j, e, sg, r) for (i, j), (e, sg, r) in (ctx.table or {}).items() if i <= j]
    vel = None if ctx.starting_velocity.start_velocity else a.velocities
    return SystemSpec(positions=a.positions, type_ids=a.type_ids.astype(np.uint32), masses=a.masses, pairs=pairs,
                      box=a.sim_box, velocities=vel, kb=KB_REFERENCE, name=path)
