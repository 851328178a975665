"""Continuous-batching job server over the C ABI (pis_server_*).

    with Server(hbm_budget_bytes=64 << 30, max_systems=64, round_steps=50) as srv:
        jid = srv.submit(spec, nsteps=1000)          # returns at once
        res = srv.wait(jid)                          # final state + last-step thermo

A native worker thread owns the device: it keeps every admitted job's state resident in
one block of a device memory pool and advances all of them together in rounds of
`round_steps` batched NVE steps.  While a round runs, on streams of their own, jobs that
finished in the previous round leave the device and queued jobs (arrival order, while the
HBM budget and the slot limit allow) are placed, listed and given their initial forces;
between rounds they swap in and out of the running batch.  Each job's result equals the
same job run alone.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import PisError, check, lib
from .engine import _PRECISIONS, _DescHolder, _dptr
from .workloads import SystemSpec


@dataclass
class JobResult:
    job_id: int
    status: int
    error: str
    n_atoms: int
    steps: int
    submit_s: float
    admit_s: float
    done_s: float
    thermo: np.ndarray  # pe, ke, etot, T, P, W_abs, W_pair of the final step
    positions: np.ndarray | None
    velocities: np.ndarray | None
    forces: np.ndarray | None

    @property
    def latency_s(self) -> float:
        return self.done_s - self.submit_s


class Server:
    def __init__(self, hbm_budget_bytes: int, max_systems: int = 64, round_steps: int = 50,
                 precision: str = "mixed"):
        self.precision = precision
        self._ptr = C.c_void_p()
        check(lib().pis_server_create(int(hbm_budget_bytes), int(max_systems), int(round_steps),
                                      _PRECISIONS[precision], C.byref(self._ptr)))
        self._n = {}

    def submit(self, spec: SystemSpec, nsteps: int | None = None, dt: float | None = None) -> int:
        h = _DescHolder(spec, self.precision)
        pos = np.ascontiguousarray(spec.positions, dtype=np.float64)
        vel = None if spec.velocities is None else np.ascontiguousarray(spec.velocities, dtype=np.float64)
        jid = C.c_uint64(0)
        check(lib().pis_server_submit(self._ptr, C.byref(h.desc), _dptr(pos), _dptr(vel),
                                      float(spec.dt if dt is None else dt),
                                      int(spec.steps if nsteps is None else nsteps), C.byref(jid)))
        self._n[jid.value] = spec.n
        return jid.value

    def wait(self, job_id: int, timeout_s: float = -1.0, state: bool = True) -> JobResult:
        n = self._n.get(job_id, 0)
        r = _capi.JobResult()
        arrs = [np.empty((n, 3)) for _ in range(3)] if state else [None, None, None]
        check(lib().pis_server_wait(self._ptr, int(job_id), float(timeout_s), C.byref(r),
                                    *[_dptr(a) for a in arrs]))
        self._n.pop(job_id, None)
        err = "" if r.status == 0 else lib().pis_last_error().decode(errors="replace")
        ok = r.status == 0
        th = np.array([r.last.pe, r.last.ke, r.last.etot, r.last.temperature, r.last.pressure,
                       r.last.virial_abs, r.last.virial_pair])
        return JobResult(r.job_id, r.status, err, r.n_atoms, r.steps, r.submit_s, r.admit_s, r.done_s, th,
                         arrs[0] if ok else None, arrs[1] if ok else None, arrs[2] if ok else None)

    def stats(self) -> dict:
        s = _capi.ServerStats()
        check(lib().pis_server_get_stats(self._ptr, C.byref(s)))
        return {f[0]: getattr(s, f[0]) for f in _capi.ServerStats._fields_ if f[0] != "reserved"}

    def close(self):
        if self._ptr:
            lib().pis_server_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "JobResult", "PisError"]
