"""ctypes binding of liboracle.so, the CPU FP64 restatement of the reference's LJ path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, always as the checker.  The product (pis_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

MGR_OFFSET, MGR_BRUTE, MGR_VERLET27, MGR_FULLLIST = 0, 1, 2, 3


class Box(C.Structure):
    _fields_ = [("h", C.c_double * 9), ("hinv", C.c_double * 9), ("pbc", C.c_int32 * 3)]


class Table(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int32),
        ("present", C.POINTER(C.c_int32)),
        ("eps", C.POINTER(C.c_double)),
        ("sigma", C.POINTER(C.c_double)),
        ("rcut", C.POINTER(C.c_double)),
        ("shift", C.c_int32),
    ]


class Sys(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pos", C.POINTER(C.c_double)),
        ("vel", C.POINTER(C.c_double)),
        ("force", C.POINTER(C.c_double)),
        ("type", C.POINTER(C.c_int32)),
        ("mass", C.POINTER(C.c_double)),
        ("box", Box),
        ("table", Table),
    ]


class NHC(C.Structure):
    """orc_nhc: Nose-Hoover chain state (ensemble/nvt.rs:4-16)."""
    _fields_ = [
        ("chain", C.c_int32),
        ("t_start", C.c_double),
        ("t_end", C.c_double),
        ("target", C.c_double),
        ("kb", C.c_double),
        ("xi", C.c_double * 8),
        ("eta", C.c_double * 8),
        ("g", C.c_double * 8),
        ("q", C.c_double * 8),
    ]


class MTK(C.Structure):
    """orc_mtk: MTK barostat state (ensemble/npt.rs:10-20)."""
    _fields_ = [("target", C.c_double * 9), ("momentum", C.c_double * 9), ("w", C.c_double)]


class List(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("start", C.POINTER(C.c_int64)),
        ("idx", C.POINTER(C.c_int64)),
        ("cap", C.c_int64),
    ]


_lib = None
_D = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_box_inverse": (C.c_int, [_D, _D]),
            "orc_min_image": (None, [C.POINTER(Box), _D]),
            "orc_wrap": (None, [C.POINTER(Box), _D]),
            "orc_divide_into_cells": (None, [C.POINTER(Box), C.c_double, C.POINTER(C.c_int64)]),
            "orc_cell_of": (C.c_int64, [C.POINTER(Box), C.POINTER(C.c_int64), _D]),
            "orc_norm": (C.c_double, [_D]),
            "orc_lj_pair": (None, [C.c_double, C.c_double, C.c_double, C.c_int32, _D, _D, _D]),
            "orc_compute": (C.c_double, [C.POINTER(Sys), C.c_int32]),
            "orc_build_list": (C.c_int, [C.POINTER(Sys), C.c_double, C.POINTER(List)]),
            "orc_brute_list": (C.c_int, [C.POINTER(Sys), C.c_double, C.POINTER(List)]),
            "orc_list_free": (None, [C.POINTER(List)]),
            "orc_compute_fulllist": (C.c_double, [C.POINTER(Sys), C.POINTER(List), C.c_int32, _D]),
            "orc_verlet_step_nve": (C.c_double, [C.POINTER(Sys), C.c_double, C.c_int32]),
            "orc_mix64": (C.c_uint64, [C.c_uint64]),
            "orc_pair_virial": (C.c_double, [C.POINTER(Sys)]),
            "orc_list_digest": (C.c_int, [C.POINTER(Sys), C.c_double, C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint64)]),
            "orc_forces_subset": (C.c_int, [C.POINTER(Sys), C.POINTER(C.c_int64), C.c_int64, _D, _D]),
            "orc_kinetic_energy": (C.c_double, [C.POINTER(Sys)]),
            "orc_temperature": (C.c_double, [C.POINTER(Sys), C.c_double, C.c_double]),
            "orc_virial_trace": (C.c_double, [C.POINTER(Sys)]),
            "orc_pressure": (C.c_double, [C.POINTER(Sys), C.c_double]),
            "orc_run_nve": (C.c_int, [C.POINTER(Sys), C.c_double, C.c_int64, C.c_int32, C.c_double, _D]),
            "orc_ideal_run_nve": (C.c_int, [C.POINTER(Sys), C.c_double, C.c_int64, C.c_double, C.c_int32, _D]),
            "orc_run_nvt": (C.c_int, [C.POINTER(Sys), C.c_double, C.c_int64, C.c_int32, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int32, _D]),
            "orc_nhc_new": (C.c_int, [C.POINTER(NHC), C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double]),
            "orc_nhc_compute_forces": (None, [C.POINTER(NHC), C.c_double, C.c_int64]),
            "orc_nhc_propagate_half_step": (None, [C.POINTER(NHC), C.c_double]),
            "orc_nhc_kinetic_energy": (C.c_double, [C.POINTER(NHC)]),
            "orc_nhc_potential_energy": (C.c_double, [C.POINTER(NHC), C.c_int64]),
            "orc_nhc_target": (None, [C.POINTER(NHC), C.c_int64, C.c_int64]),
            "orc_verlet_step_nvt_nhc": (C.c_double, [C.POINTER(Sys), C.c_double, C.POINTER(NHC), C.c_int32]),
            "orc_mtk_new": (None, [C.POINTER(MTK), C.c_double, C.c_double, C.c_int64, C.c_double]),
            "orc_kinetic_tensor": (None, [C.POINTER(Sys), _D]),
            "orc_virial_tensor": (None, [C.POINTER(Sys), _D]),
            "orc_pressure_tensor": (None, [C.POINTER(Sys), _D]),
            "orc_mtk_delta_momentum": (None, [C.POINTER(MTK), C.POINTER(Sys), C.c_double, _D]),
            "orc_mtk_scale": (None, [C.POINTER(MTK), C.c_double, C.c_int32, _D]),
            "orc_mtk_kinetic_energy": (C.c_double, [C.POINTER(MTK)]),
            "orc_mtk_potential_energy": (C.c_double, [C.POINTER(MTK), _D]),
            "orc_mat3_mul": (None, [_D, _D, _D]),
            "orc_mat3_exp": (None, [_D, _D]),
            "orc_scale_box": (None, [C.POINTER(Sys), _D]),
            "orc_verlet_step_npt_mtk": (C.c_double, [C.POINTER(Sys), C.c_double, C.POINTER(MTK), C.POINTER(NHC),
                                                     C.c_int32]),
            "orc_run_npt": (C.c_int, [C.POINTER(Sys), C.c_double, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def box_inverse(h):
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(9)
    if lib().orc_box_inverse(h.ctypes.data_as(_D), out.ctypes.data_as(_D)) != 0:
        raise ValueError("singular box")
    return out


def make_box(h, hinv=None, pbc=(1, 1, 1)) -> Box:
    b = Box()
    hinv = box_inverse(h) if hinv is None else hinv
    for k in range(9):
        b.h[k] = float(h[k])
        b.hinv[k] = float(hinv[k])
    for k in range(3):
        b.pbc[k] = int(pbc[k])
    return b


class OSystem:
    """Oracle system: owns contiguous AoS arrays and the dense pair table.  Pairs later in
    `pairs` overwrite earlier ones with the same (min, max) key, like HashMap::insert."""

    def __init__(self, positions, type_ids, masses, pairs, h, hinv=None, velocities=None,
                 forces=None, shift=True):
        self.pos = np.ascontiguousarray(np.array(positions, dtype=np.float64).reshape(-1, 3))
        n = self.pos.shape[0]
        self.vel = (np.zeros((n, 3)) if velocities is None
                    else np.ascontiguousarray(np.array(velocities, dtype=np.float64).reshape(-1, 3)))
        self.force = (np.zeros((n, 3)) if forces is None
                      else np.ascontiguousarray(np.array(forces, dtype=np.float64).reshape(-1, 3)))
        self.type = np.ascontiguousarray(np.asarray(type_ids, dtype=np.int32))
        self.mass = np.ascontiguousarray(np.asarray(masses, dtype=np.float64))
        nt = self.mass.shape[0]
        self.present = np.zeros(nt * nt, dtype=np.int32)
        self.eps = np.zeros(nt * nt)
        self.sigma = np.ones(nt * nt)
        self.rcut = np.zeros(nt * nt)
        for (ti, tj, e, s, rc) in pairs:
            a, b = min(ti, tj) - 1, max(ti, tj) - 1
            for k in (a * nt + b, b * nt + a):
                self.present[k] = 1
                self.eps[k] = e
                self.sigma[k] = s
                self.rcut[k] = rc
        s = Sys()
        s.n = n
        s.pos = self.pos.ctypes.data_as(_D)
        s.vel = self.vel.ctypes.data_as(_D)
        s.force = self.force.ctypes.data_as(_D)
        s.type = self.type.ctypes.data_as(C.POINTER(C.c_int32))
        s.mass = self.mass.ctypes.data_as(_D)
        s.box = make_box(h, hinv)
        t = Table()
        t.ntypes = nt
        t.present = self.present.ctypes.data_as(C.POINTER(C.c_int32))
        t.eps = self.eps.ctypes.data_as(_D)
        t.sigma = self.sigma.ctypes.data_as(_D)
        t.rcut = self.rcut.ctypes.data_as(_D)
        t.shift = 1 if shift else 0
        s.table = t
        self.s = s

    @classmethod
    def from_spec(cls, spec, velocities=True):
        return cls(spec.positions, spec.type_ids, spec.masses, spec.pairs, spec.box.h, spec.box.h_inv,
                   velocities=spec.velocities if velocities else None, shift=spec.shift)

    @property
    def n(self):
        return int(self.s.n)

    def compute(self, manager=MGR_OFFSET, zero=True) -> float:
        if zero:
            self.force[:] = 0.0
        return lib().orc_compute(C.byref(self.s), manager)

    def list(self, rcut_add=0.0, brute=False):
        L = List()
        f = lib().orc_brute_list if brute else lib().orc_build_list
        if f(C.byref(self.s), rcut_add, C.byref(L)) != 0:
            raise RuntimeError("oracle list build failed")
        n = self.n
        start = np.ctypeslib.as_array(L.start, shape=(n + 1,)).copy()
        total = int(start[-1])
        idx = np.ctypeslib.as_array(L.idx, shape=(max(total, 1),))[:total].copy() if total else np.zeros(0, np.int64)
        lib().orc_list_free(C.byref(L))
        return start, idx

    def compute_fulllist(self, rcut_add=0.0, check_rcut=True, brute=False):
        """Per-atom gather over a full list (the reference's cell-built one, or the
        brute-force LJManager-semantics one, correct for any grid); returns (PE, pe_i)."""
        L = List()
        build = lib().orc_brute_list if brute else lib().orc_build_list
        if build(C.byref(self.s), rcut_add, C.byref(L)) != 0:
            raise RuntimeError("oracle list build failed")
        self.force[:] = 0.0
        pe_atom = np.zeros(self.n)
        pe = lib().orc_compute_fulllist(C.byref(self.s), C.byref(L), 1 if check_rcut else 0,
                                        pe_atom.ctypes.data_as(_D))
        lib().orc_list_free(C.byref(L))
        return pe, pe_atom

    def list_digest(self, rcut_add=0.0):
        """(count, hash) per atom of the full list at rcut + rcut_add: row length and the
        order-independent row hash sum(mix64(j)) mod 2^64 (orc_list_digest)."""
        count = np.zeros(self.n, dtype=np.uint32)
        h = np.zeros(self.n, dtype=np.uint64)
        rc = lib().orc_list_digest(C.byref(self.s), rcut_add, count.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   h.ctypes.data_as(C.POINTER(C.c_uint64)))
        if rc != 0:
            raise RuntimeError("oracle list digest failed")
        return count, h

    def pair_virial(self) -> float:
        """sum_{i<j} r_ij . F_ij over the pairs within the cutoff (orc_pair_virial)."""
        return lib().orc_pair_virial(C.byref(self.s))

    def forces_subset(self, idx):
        """(forces [m, 3], unhalved pair energies [m]) of the atoms idx, as the full-list
        gather gives them (orc_forces_subset)."""
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
        f = np.zeros((idx.size, 3))
        pe = np.zeros(idx.size)
        rc = lib().orc_forces_subset(C.byref(self.s), idx.ctypes.data_as(C.POINTER(C.c_int64)), idx.size,
                                     f.ctypes.data_as(_D), pe.ctypes.data_as(_D))
        if rc != 0:
            raise RuntimeError("oracle subset forces failed")
        return f, pe

    def verlet_step_nve(self, dt, manager=MGR_OFFSET) -> float:
        return lib().orc_verlet_step_nve(C.byref(self.s), dt, manager)

    def kinetic_energy(self) -> float:
        return lib().orc_kinetic_energy(C.byref(self.s))

    def virial_trace(self) -> float:
        return lib().orc_virial_trace(C.byref(self.s))

    def pressure(self, ke) -> float:
        return lib().orc_pressure(C.byref(self.s), ke)

    def temperature(self, ke, kb) -> float:
        return lib().orc_temperature(C.byref(self.s), ke, kb)

    def run_nve(self, dt, nsteps, manager=MGR_OFFSET, kb=1.0):
        th = np.zeros((nsteps, 5))
        if lib().orc_run_nve(C.byref(self.s), dt, nsteps, manager, kb, th.ctypes.data_as(_D)) != 0:
            raise RuntimeError("oracle run failed")
        return th

    def run_nvt(self, dt, nsteps, t_start, t_end, tau, chain=3, manager=MGR_OFFSET, kb=1.0):
        th = np.zeros((nsteps, 5))
        rc = lib().orc_run_nvt(C.byref(self.s), dt, nsteps, manager, kb, t_start, t_end, tau, chain,
                               th.ctypes.data_as(_D))
        if rc != 0:
            raise RuntimeError("oracle NVT run failed")
        return th

    def run_npt(self, dt, nsteps, t_start, t_end, tau_t, p0, tau_p, chain=3, manager=MGR_OFFSET, kb=1.0):
        """Simulation::run with Ensemble::NPT; returns thermo [nsteps, 5] and the box matrix
        after every step [nsteps, 9] (column-major)."""
        th = np.zeros((nsteps, 5))
        hb = np.zeros((nsteps, 9))
        rc = lib().orc_run_npt(C.byref(self.s), dt, nsteps, manager, kb, t_start, t_end, tau_t, chain, p0, tau_p,
                               th.ctypes.data_as(_D), hb.ctypes.data_as(_D))
        if rc != 0:
            raise RuntimeError("oracle NPT run failed")
        return th, hb

    def tensor(self, which):
        """'kinetic', 'virial' or 'pressure' tensor of the current state (3x3, row-major view)."""
        out = np.zeros(9)
        f = {"kinetic": lib().orc_kinetic_tensor, "virial": lib().orc_virial_tensor,
             "pressure": lib().orc_pressure_tensor}[which]
        f(C.byref(self.s), out.ctypes.data_as(_D))
        return out.reshape(3, 3).T.copy()

    def box_h(self):
        return np.array(self.s.box.h[:]).reshape(3, 3).T.copy()

    def scale_box(self, scale):
        m = np.ascontiguousarray(np.asarray(scale, dtype=np.float64).T.reshape(9))
        lib().orc_scale_box(C.byref(self.s), m.ctypes.data_as(_D))

    def ideal_run_nve(self, dt, nsteps, skin, nthreads=0):
        pe = C.c_double(0.0)
        if lib().orc_ideal_run_nve(C.byref(self.s), dt, nsteps, skin, nthreads, C.byref(pe)) != 0:
            raise RuntimeError("ideal CPU run failed")
        return pe.value


def min_image(box: Box, d):
    d = np.ascontiguousarray(np.array(d, dtype=np.float64))
    lib().orc_min_image(C.byref(box), d.ctypes.data_as(_D))
    return d


def orc_norm_py(v):
    """nalgebra norm() as the oracle computes it (sqrt of (x^2+y^2)+z^2)."""
    v = np.ascontiguousarray(np.array(v, dtype=np.float64))
    return lib().orc_norm(v.ctypes.data_as(_D))


def wrap(box: Box, r):
    r = np.ascontiguousarray(np.array(r, dtype=np.float64))
    lib().orc_wrap(C.byref(box), r.ctypes.data_as(_D))
    return r


def lj_pair(eps, sigma, rcut, shift, rij):
    rij = np.ascontiguousarray(np.array(rij, dtype=np.float64))
    u = np.zeros(1)
    f = np.zeros(3)
    lib().orc_lj_pair(eps, sigma, rcut, 1 if shift else 0, rij.ctypes.data_as(_D), u.ctypes.data_as(_D),
                      f.ctypes.data_as(_D))
    return float(u[0]), f


def divide_into_cells(box: Box, rcut):
    n = np.zeros(3, dtype=np.int64)
    lib().orc_divide_into_cells(C.byref(box), rcut, n.ctypes.data_as(C.POINTER(C.c_int64)))
    return tuple(int(x) for x in n)


def mix64(x):
    """splitmix64's output function on a uint64 array (numpy wraps on overflow) -- the
    same entry hash as orc_mix64, restated independently for the digest tests."""
    x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def csr_digest(start, idx):
    """(count, hash) per row of a CSR list: row length and sum of mix64(j) mod 2^64."""
    start = np.asarray(start, dtype=np.int64)
    counts = np.diff(start).astype(np.uint32)
    h = mix64(np.asarray(idx, dtype=np.uint64))
    cs = np.concatenate([[np.uint64(0)], np.cumsum(h, dtype=np.uint64)])
    return counts, (cs[start[1:]] - cs[start[:-1]]).astype(np.uint64)


def pair_set(start, idx):
    """Set of (i, j) from a CSR list."""
    out = set()
    for i in range(len(start) - 1):
        for j in idx[start[i]:start[i + 1]]:
            out.add((i, int(j)))
    return out


def csr_to_sorted_pairs(start, idx) -> np.ndarray:
    """(M, 2) int64 array of (i, j) sorted lexicographically -- fast set comparison."""
    start = np.asarray(start, dtype=np.int64)
    counts = np.diff(start)
    i = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    pairs = np.stack([i, np.asarray(idx, dtype=np.int64)], axis=1)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order]


def mat3_exp(a):
    """The oracle's 3x3 matrix exponential (row-major in, row-major out)."""
    m = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T.reshape(9))
    out = np.zeros(9)
    lib().orc_mat3_exp(m.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return out.reshape(3, 3).T.copy()
