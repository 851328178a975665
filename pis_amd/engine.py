"""Python face of the C ABI: a device-resident engine over one or many systems, and a
mirror of the reference's manager interface for drop-in use.

Reference interface mirrored (src/potentials/potential.rs:12-33, lennard_jones.rs:182-244):
    mgr = LJGpuManager(); mgr.insert((1, 1), LennardJones(eps, sigma, rcut, True))
    pe = mgr.compute_potential(atoms)     # accumulates into atoms.forces, returns PE
    pe = mgr.verlet_step_nve(atoms, dt)   # one step, returns PE(t+dt)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import PRECISION_FP64, PRECISION_MIXED, PisError, check, lib
from .box import SimulationBox
from .workloads import SystemSpec

_PRECISIONS = {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}


def _dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class _DescHolder:
    """A pis_system_desc plus the arrays it points into (kept alive together)."""

    def __init__(self, spec: SystemSpec, precision: str, neighbor_capacity: int = 0):
        self.types = np.ascontiguousarray(spec.type_ids, dtype=np.uint32)
        self.masses = np.ascontiguousarray(spec.masses, dtype=np.float64)
        npairs = len(spec.pairs)
        self.pairs = (_capi.PairCoeff * max(npairs, 1))()
        for k, (ti, tj, eps, sig, rc) in enumerate(spec.pairs):
            self.pairs[k] = _capi.PairCoeff(int(ti), int(tj), float(eps), float(sig), float(rc))
        d = _capi.SystemDesc()
        d.n_atoms = spec.n
        d.n_types = int(self.masses.shape[0])
        d.n_pairs = npairs
        d.type_ids = self.types.ctypes.data_as(C.POINTER(C.c_uint32))
        d.masses = self.masses.ctypes.data_as(C.POINTER(C.c_double))
        d.pairs = C.cast(self.pairs, C.POINTER(_capi.PairCoeff)) if npairs else None
        for k in range(9):
            d.h[k] = spec.box.h[k]
            d.h_inv[k] = spec.box.h_inv[k]
        for k in range(3):
            d.pbc[k] = 1 if spec.box.pbc[k] else 0
        d.shift = 1 if spec.shift else 0
        d.precision = _PRECISIONS[precision]
        d.skin = float(spec.skin)
        d.kb = float(spec.kb)
        d.neighbor_capacity = int(neighbor_capacity)
        d.flags = 0
        if spec.tile is not None:
            for k in range(3):
                d.tile[k] = int(spec.tile[k])
        self.desc = d


def plan(spec: SystemSpec, precision: str = "fp64") -> dict:
    """Host-only plan of a system (grid, tiles, mode, capacities) -- needs no GPU."""
    h = _DescHolder(spec, precision)
    info = _capi.PlanInfo()
    check(lib().pis_plan_system(C.byref(h.desc), C.byref(info)))
    return {
        "ncell": tuple(info.ncell),
        "tile": tuple(info.tile),
        "ntile": tuple(info.ntile),
        "mode": info.mode,
        "neighbor_capacity": info.neighbor_capacity,
        "stage_capacity": info.stage_capacity,
        "r2_list_max": info.r2_list_max,
        "density": info.density,
    }


def thermo_to_array(recs) -> np.ndarray:
    """pis_thermo records -> float64 array [..., 7] (pe, ke, etot, T, P, W_abs, W_pair)."""
    return np.ctypeslib.as_array(recs).view(np.float64).reshape(-1, 7).copy()


class PinnedArray:
    """float64 array in pinned host memory (pis_host_alloc), the destination that makes a
    position snapshot's transfer asynchronous.  Freed by free() or when collected."""

    def __init__(self, shape):
        count = int(np.prod(shape))
        self._p = C.c_void_p()
        check(lib().pis_host_alloc(8 * count, C.byref(self._p)))
        self.array = np.ctypeslib.as_array((C.c_double * count).from_address(self._p.value)).reshape(shape)

    def free(self):
        if self._p:
            self.array = None
            lib().pis_host_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """Device-resident state of one or many systems (pis_ctx)."""

    def __init__(self, specs, precision: str = "fp64", neighbor_capacity: int = 0):
        if isinstance(specs, SystemSpec):
            specs = [specs]
        self.specs = list(specs)
        self.precision = precision
        self._holders = [_DescHolder(s, precision, neighbor_capacity) for s in self.specs]
        arr = (_capi.SystemDesc * len(self._holders))(*[h.desc for h in self._holders])
        self._ptr = C.c_void_p()
        check(lib().pis_batch_create(arr, len(self._holders), C.byref(self._ptr)))

    @property
    def nsys(self) -> int:
        return len(self.specs)

    def close(self):
        if self._ptr:
            lib().pis_ctx_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, k: int, pos, vel=None, force=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        force = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        n = self.specs[k].n
        for a in (pos, vel, force):
            if a is not None and a.size != 3 * n:
                raise ValueError(f"expected {3 * n} values, got {a.size}")
        check(lib().pis_sys_set_state(self._ptr, k, _dptr(pos), _dptr(vel), _dptr(force)))

    def load(self):
        """Upload every spec's positions and velocities."""
        for k, s in enumerate(self.specs):
            self.set_state(k, s.positions, s.velocities)

    def get_state(self, k: int = 0):
        n = self.specs[k].n
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        force = np.empty((n, 3))
        check(lib().pis_sys_get_state(self._ptr, k, _dptr(pos), _dptr(vel), _dptr(force)))
        return pos, vel, force

    def get_positions(self, k: int = 0) -> np.ndarray:
        """Positions of system k alone (original order), for trajectory frames."""
        pos = np.empty((self.specs[k].n, 3))
        check(lib().pis_sys_get_state(self._ptr, k, _dptr(pos), None, None))
        return pos

    def snapshot_positions(self, k: int, dest: PinnedArray):
        """Start copying system k's positions into dest behind the queued steps; returns at
        once.  dest is valid after snapshot_wait()."""
        if dest.array is None or dest.array.size != 3 * self.specs[k].n:
            raise ValueError("snapshot destination has the wrong size")
        check(lib().pis_snapshot_positions(self._ptr, k, _dptr(dest.array)))

    def snapshot_wait(self):
        check(lib().pis_snapshot_wait(self._ptr))

    def compute_potential(self) -> np.ndarray:
        pe = np.empty(self.nsys)
        check(lib().pis_batch_compute_potential(self._ptr, _dptr(pe)))
        return pe

    def run_nve(self, dt, nsteps: int, record: bool = True):
        """nsteps of verlet_step_nve for every system; returns thermo [nsteps, nsys, 7]."""
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nsys,)))
        recs = (_capi.Thermo * (nsteps * self.nsys))() if record else None
        check(lib().pis_batch_verlet_step_nve(self._ptr, _dptr(dts), nsteps, recs))
        if not record:
            return None
        return thermo_to_array(recs).reshape(nsteps, self.nsys, 7)

    def nvt_init(self, t_start: float, t_end: float, tau: float, chain: int = 3):
        """Nose-Hoover chain for every system (NHThermostatChain::new_from_args)."""
        prm = _capi.NHCParams(float(t_start), float(t_end), float(tau), int(chain), 0)
        check(lib().pis_nvt_init(self._ptr, C.byref(prm)))

    def run_nvt(self, dt, nsteps: int, record: bool = True, ramp=None):
        """nsteps of verlet_step_nvt_nhc for every system; thermo [nsteps, nsys, 7] with
        etot = PE + KE + chain KE.  ramp=(first step, total steps) continues the target
        temperature ramp of a longer run (default: this call is the whole run)."""
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nsys,)))
        recs = (_capi.Thermo * (nsteps * self.nsys))() if record else None
        if ramp is None:
            check(lib().pis_batch_verlet_step_nvt(self._ptr, _dptr(dts), nsteps, recs))
        else:
            check(lib().pis_batch_verlet_step_nvt_ramp(self._ptr, _dptr(dts), nsteps, int(ramp[0]), int(ramp[1]),
                                                       recs))
        if not record:
            return None
        return thermo_to_array(recs).reshape(nsteps, self.nsys, 7)

    def npt_init(self, t_start: float, t_end: float, tau_t: float, p_start: float, tau_p: float, chain: int = 3):
        """MTK barostat + Nose-Hoover chain for every system (MTKBarostat::new_from_args,
        target pressure p_start * I)."""
        prm = _capi.NHCParams(float(t_start), float(t_end), float(tau_t), int(chain), 0)
        mtk = _capi.MTKParams(float(p_start), float(tau_p))
        check(lib().pis_npt_init(self._ptr, C.byref(prm), C.byref(mtk)))

    def run_npt(self, dt, nsteps: int, record: bool = True, ramp=None):
        """nsteps of verlet_step_npt_mtk for every system; thermo [nsteps, nsys, 7] with
        etot the NPT Hamiltonian.  ramp as in run_nvt."""
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nsys,)))
        recs = (_capi.Thermo * (nsteps * self.nsys))() if record else None
        r0, rt = (0, 0) if ramp is None else (int(ramp[0]), int(ramp[1]))
        check(lib().pis_batch_verlet_step_npt(self._ptr, _dptr(dts), nsteps, r0, rt, recs))
        if not record:
            return None
        return thermo_to_array(recs).reshape(nsteps, self.nsys, 7)

    def npt_box(self, k: int = 0):
        """(h, momentum) of system k as 3x3 arrays (h's columns are the box vectors)."""
        h = np.zeros(9)
        m = np.zeros(9)
        check(lib().pis_npt_get_box(self._ptr, k, _dptr(h), _dptr(m)))
        return h.reshape(3, 3).T.copy(), m.reshape(3, 3).T.copy()

    def nvt_chain(self, k: int = 0, chain: int = 3):
        """(xi[chain], target temperature) of system k's thermostat chain."""
        xi = np.zeros(8)
        t = C.c_double(0.0)
        check(lib().pis_nvt_get_chain(self._ptr, k, _dptr(xi), C.byref(t)))
        return xi[:chain].copy(), t.value

    def neighbor_list(self):
        """(offsets, idx) CSR of the full list at the current positions, original indices."""
        n = self.specs[0].n
        offsets = np.zeros(n + 1, dtype=np.uint64)
        total = C.c_uint64(0)
        check(lib().pis_get_neighbor_list(self._ptr, offsets.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          None, 0, C.byref(total)))
        idx = np.zeros(max(int(total.value), 1), dtype=np.uint32)
        check(lib().pis_get_neighbor_list(self._ptr, offsets.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size,
                                          C.byref(total)))
        return offsets.astype(np.int64), idx[: int(total.value)].astype(np.int64)

    def list_digest(self):
        """(counts, hashes) per atom (original order) of the full list at the current
        positions: row length and sum of mix64(j) mod 2^64 over the row."""
        n = self.specs[0].n
        counts = np.zeros(n, dtype=np.uint32)
        hashes = np.zeros(n, dtype=np.uint64)
        check(lib().pis_get_list_digest(self._ptr, counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        hashes.ctypes.data_as(C.POINTER(C.c_uint64))))
        return counts, hashes

    def stats(self) -> dict:
        s = _capi.Stats()
        check(lib().pis_get_stats(self._ptr, C.byref(s)))
        return {f[0]: getattr(s, f[0]) for f in _capi.Stats._fields_}

    def set_graph_mode(self, enable: bool):
        check(lib().pis_set_graph_mode(self._ptr, 1 if enable else 0))

    def set_kernel_timing(self, enable: bool):
        check(lib().pis_set_kernel_timing(self._ptr, 1 if enable else 0))


# ---------------------------------------------------------------------------------
# Mirror of the reference's types (drop-in shape of the Rust API)


class LennardJones:
    """LennardJones::new(epsilon, sigma, rcut, shift) (lennard_jones.rs:14-30)."""

    def __init__(self, epsilon: float, sigma: float, rcut: float, shift: bool = True):
        self.epsilon, self.sigma, self.rcut, self.shift = float(epsilon), float(sigma), float(rcut), bool(shift)

    def get_rcut(self) -> float:
        return self.rcut


class Atoms:
    """Atoms (src/atoms/new.rs:9-17): positions / velocities / forces are (N, 3) float64
    arrays, i.e. the reference's column-major 3xN memory."""

    def __init__(self, type_ids, masses, positions, velocities=None, sim_box: SimulationBox | None = None):
        # Atoms::new takes ownership of its arrays: copies here, so stepping these atoms
        # never writes into the caller's arrays
        self.type_ids = np.array(type_ids, dtype=np.uint32)
        self.masses = np.array(masses, dtype=np.float64)
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.n_atoms = self.positions.shape[0]
        self.velocities = (np.zeros_like(self.positions) if velocities is None
                           else np.array(velocities, dtype=np.float64).reshape(-1, 3))
        self.forces = np.zeros_like(self.positions)
        self.sim_box = sim_box


class LJGpuManager:
    """GPU PotentialManager: the table API of PairPotentialManager (potential.rs:148-193) and
    compute_potential / verlet_step_nve through the C ABI.  Each call ships the state across
    PCIe like the reference's by-reference call would; Engine.run_nve keeps it resident."""

    def __init__(self, skin: float = 0.0, kb: float = 0.0083144621, precision: str = "fp64"):
        self.table = {}
        self.skin = skin
        self.kb = kb
        self.precision = precision
        self._engine = None
        self._key = None

    def insert(self, key, potential: LennardJones):
        self.table[(int(key[0]), int(key[1]))] = potential
        self._engine = None

    def get(self, key):
        return self.table.get(key)

    def is_empty(self) -> bool:
        return not self.table

    def max_rcut(self) -> float:
        return max((p.rcut for p in self.table.values()), default=0.0)

    def _spec(self, atoms: Atoms) -> SystemSpec:
        shift = all(p.shift for p in self.table.values()) if self.table else True
        pairs = [(i, j, p.epsilon, p.sigma, p.rcut) for (i, j), p in self.table.items()]
        return SystemSpec(positions=atoms.positions, type_ids=atoms.type_ids, masses=atoms.masses,
                          pairs=pairs, box=atoms.sim_box, skin=self.skin, kb=self.kb, shift=shift)

    def _engine_for(self, atoms: Atoms) -> Engine:
        key = (atoms.n_atoms, tuple(atoms.sim_box.h), atoms.type_ids.tobytes(), atoms.masses.tobytes())
        if self._engine is None or self._key != key:
            self._engine = Engine(self._spec(atoms), self.precision)
            self._key = key
        return self._engine

    def compute_potential(self, atoms: Atoms) -> float:
        eng = self._engine_for(atoms)
        eng.set_state(0, atoms.positions, atoms.velocities, None)
        pe = float(eng.compute_potential()[0])
        _, _, f = eng.get_state(0)
        atoms.forces += f
        return pe

    def verlet_step_nve(self, atoms: Atoms, dt: float) -> float:
        eng = self._engine_for(atoms)
        eng.set_state(0, atoms.positions, atoms.velocities, atoms.forces)
        th = eng.run_nve(dt, 1)
        atoms.positions[:], atoms.velocities[:], atoms.forces[:] = eng.get_state(0)
        return float(th[0, 0, 0])


__all__ = ["Engine", "LJGpuManager", "LennardJones", "Atoms", "PisError", "plan", "thermo_to_array"]
