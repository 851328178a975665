"""The reference's input script and data file as a job spec, and its run loop on the GPU.

Restates, with the same accepted syntax and the same errors:
  * the script reader `System::read` (src/system.rs:49-94) and its commands (timestep,
    run, velocity, read_data, fix, pair_style, pair_coeff, dump;
    src/readers/input_file/commands.rs:25-61, 64-146, 386-505),
  * the data-file reader `run_read_data` (commands.rs:147-384),
  * `System::contextualize` (system.rs:100-175): starting velocities and the lj/cut table,
  * `Simulation::run` (src/simulation.rs:8-116): the per-step thermo line and the dump
    cadence, with the steps run by the device-resident engine (NVE, NVT and NPT),
  * the trajectory writer `DumpTraj` (src/writers/dump_traj.rs:12-75), numbers printed the
    way Rust's `{}` prints them.

Reference quirks kept: `pair_coeff i j` is stored under (i, j) but looked up under
(min, max), so an entry with i > j never applies; a data-file PairCoeffs line `1 2 ...`
reads "2" as epsilon; `velocity ... dist` is parsed and then ignored; the initial velocities
come from this package's own seeded generator (the reference's SmallRng stream is not
reproduced).
"""
from __future__ import annotations

import os
import re
import threading
from dataclasses import dataclass, field
from decimal import Decimal

import numpy as np

from .box import SimulationBox
from .workloads import KB_REFERENCE, SystemSpec, start_velocities

# ---------------------------------------------------------------------------------
# errors (src/errors.rs) and Rust-style scalar parsing (src/extensions.rs)


class PisScriptError(Exception):
    """A PisError variant: `kind` is its name, `fields` its payload."""

    def __init__(self, kind: str, **fields):
        self.kind = kind
        self.fields = fields
        super().__init__(f"{kind} {fields}" if fields else kind)


_FLOAT = re.compile(r"^[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf|infinity|nan)$", re.IGNORECASE)
_INT = re.compile(r"^[+-]?\d+$")


def _get(args, i, line):
    if i >= len(args):
        raise PisScriptError("MissingArgument", line=line)
    return args[i]


def _float(args, i, line) -> float:
    s = _get(args, i, line)
    if not _FLOAT.match(s):  # str::parse::<f64>
        raise PisScriptError("FloatParseError", string=s)
    return float(s)


def _int(args, i, line) -> int:
    s = _get(args, i, line)
    if not _INT.match(s) or not (-(2 ** 31) <= int(s) < 2 ** 31):  # str::parse::<i32>
        raise PisScriptError("IntParseError", string=s)
    return int(s)


def _usize(value: int, line) -> int:
    if value < 0:
        raise PisScriptError("NegativeValue", value=value, line=line)
    return value


# ---------------------------------------------------------------------------------
# context (src/readers/simulation_context.rs)


@dataclass
class StartVelocity:
    group: str = "all"
    start_velocity: bool = True
    start_temperature: float | None = None
    seed: int | None = None
    dist: str | None = None


@dataclass
class PotentialArgs:
    pair_style_args: list = field(default_factory=list)
    pair_style_line: int = 0
    pair_coeff_args: list = field(default_factory=list)
    pair_coeff_lines: list = field(default_factory=list)


@dataclass
class DumpArgs:
    name: str = "default_dump"
    group: str = "all"
    style: str = "atoms"
    dump_step: int = 1
    file_name: str = "dump.lammpstrj"


@dataclass
class Atoms:
    """Atoms as run_read_data builds them (types 1-based, per-type masses)."""
    n_atoms: int
    type_ids: np.ndarray
    masses: np.ndarray
    positions: np.ndarray
    velocities: np.ndarray
    sim_box: SimulationBox


@dataclass
class Context:
    atoms: Atoms | None = None
    timestep: float = 0.01
    table: dict | None = None  # (i, j) -> (eps, sigma, rcut): the manager's pair table
    starting_velocity: StartVelocity | None = None
    steps: int = 100
    potential_args: PotentialArgs | None = None
    nh_chain_args: dict | None = None
    mtk_barostat_args: dict | None = None
    dump_args: DumpArgs = field(default_factory=DumpArgs)
    base_dir: str = "."

    @property
    def ensemble(self) -> str:
        """Ensemble::from_ctx (simulation.rs:126-132)."""
        if self.nh_chain_args is not None and self.mtk_barostat_args is not None:
            return "npt"
        if self.nh_chain_args is not None:
            return "nvt"
        return "nve"


# ---------------------------------------------------------------------------------
# commands (commands.rs)

COMMANDS = ("timestep", "run", "velocity", "read_data", "fix", "pair_style", "pair_coeff", "dump")


def command_from_str(s: str):
    return s if s in COMMANDS else None


def run_command(cmd: str, args, line: int, ctx: Context) -> None:
    {
        "timestep": _timestep,
        "run": _run,
        "velocity": _velocity,
        "read_data": _read_data_cmd,
        "fix": _fix,
        "pair_style": _pair_style,
        "pair_coeff": _pair_coeff,
        "dump": _dump,
    }[cmd](list(args), line, ctx)


def _timestep(args, line, ctx):
    ctx.timestep = _float(args, 0, line)


def _run(args, line, ctx):
    ctx.steps = _usize(_int(args, 0, line), line)


def _velocity(args, line, ctx):
    sv = StartVelocity()
    k = 0
    sv.group = _get(args, k, line)
    k += 1
    style = _get(args, k, line)
    k += 1
    if style != "create":
        raise PisScriptError("InvalidArgument", string=style, line=line)
    sv.start_temperature = _float(args, k, line)
    k += 1
    try:
        sv.seed = _usize(_int(args, k, line), line)
        k += 1
    except PisScriptError as e:
        if e.kind == "NegativeValue":
            raise
        sv.seed = 0
    while k < len(args):
        kw = args[k]
        k += 1
        if kw == "dist":
            val = _get(args, k, line)
            k += 1
            if val not in ("uniform", "gaussian"):
                raise PisScriptError("InvalidArgument", string=val, line=line)
            sv.dist = val
        else:
            raise PisScriptError("InvalidArgument", string=kw, line=line)
    ctx.starting_velocity = sv


def _fix(args, line, ctx):
    k = 0
    name = _get(args, k, line)
    k += 1
    group = _get(args, k, line)
    k += 1
    style = _get(args, k, line)
    k += 1
    while k < len(args):
        kw = args[k]
        k += 1
        if kw == "temp":
            t0 = _float(args, k, line)
            t1 = _float(args, k + 1, line)
            tau = _float(args, k + 2, line)
            k += 3
            if style in ("npt", "nvt"):
                ctx.nh_chain_args = {"name": name, "group": group, "start_temperature": t0,
                                     "end_temperature": t1, "tau": tau}
        elif kw == "iso":
            p0 = _float(args, k, line)
            _p1 = _float(args, k + 1, line)
            tau = _float(args, k + 2, line)
            k += 3
            if style == "npt":
                ctx.mtk_barostat_args = {"name": name, "group": group, "start_pressure": p0,
                                         "end_pressure": p0, "tau": tau}
        else:
            raise PisScriptError("InvalidArgument", string=kw, line=line)


def _pair_style(args, line, ctx):
    ctx.potential_args = PotentialArgs(pair_style_args=list(args), pair_style_line=line)


def _pair_coeff(args, line, ctx):
    if ctx.potential_args is None:
        raise PisScriptError("PotentialNotInitialized")
    ctx.potential_args.pair_coeff_args.append(list(args))
    ctx.potential_args.pair_coeff_lines.append(line)


def _dump(args, line, ctx):
    d = ctx.dump_args
    d.name = _get(args, 0, line)
    d.group = _get(args, 1, line)
    d.style = _get(args, 2, line)
    d.dump_step = _usize(_int(args, 3, line), line)
    d.file_name = _get(args, 4, line)


def _read_data_cmd(args, line, ctx):
    path = _get(args, 0, line)
    full = path if os.path.isabs(path) else os.path.join(ctx.base_dir, path)
    read_data(full, ctx, shown_path=path)


def read_data(path: str, ctx: Context, shown_path: str | None = None) -> None:
    """run_read_data (commands.rs:147-384)."""
    try:
        fh = open(path)
    except OSError:
        raise PisScriptError("InputFileError", path=shown_path or path)
    section = ""
    n_atoms = 0
    lo = {"x": 0.0, "y": 0.0, "z": 0.0}
    hi = {"x": 1.0, "y": 1.0, "z": 1.0}
    masses = []
    types = np.zeros(0, dtype=np.int64)
    pos = np.zeros((0, 3))
    vel = np.zeros((0, 3))
    start = True
    table = {}
    with fh:
        for ln, raw in enumerate(fh):
            s = raw.strip()
            if not s or s.startswith("#"):
                continue
            w = s.split()
            if w[0] in ("Masses", "Atoms", "PairCoeffs"):
                section = w[0]
                continue
            if w[0] == "Velocities":
                start = False
                section = w[0]
                continue
            if len(w) > 1 and w[1] == "atoms":
                n_atoms = _usize(_int(w, 0, ln), ln)
                types = np.zeros(n_atoms, dtype=np.int64)
                pos = np.zeros((n_atoms, 3))
                vel = np.zeros((n_atoms, 3))
                continue
            if len(w) > 1 and w[1] == "atom":
                nt = _usize(_int(w, 0, ln), ln)
                masses = (masses + [0.0] * nt)[:nt]
                continue
            if len(w) > 2 and w[2] in ("xlo", "ylo", "zlo"):
                d = w[2][0]
                lo[d] = _float(w, 0, ln)
                hi[d] = _float(w, 1, ln)
                continue
            if section == "Masses":
                t = _usize(_int(w, 0, ln), ln)
                m = _float(w, 1, ln)
                if t < 1:
                    raise PisScriptError("InvalidAtomType", type_id=t)
                masses[t - 1] = m
            elif section == "PairCoeffs":
                i = _usize(_int(w, 0, ln), ln)
                try:
                    eps = _float(w, 1, ln)  # "1 2 eps ..." reads 2 as epsilon (kept)
                    sig = _float(w, 2, ln)
                    try:
                        rc = _float(w, 3, ln)
                    except PisScriptError:
                        rc = 2.5 * sig
                    table[(i, i)] = (eps, sig, rc)
                except PisScriptError as e:
                    if e.kind != "FloatParseError" or e.fields.get("string") != w[1]:
                        raise
                    j = _usize(_int(w, 1, ln), ln)
                    eps = _float(w, 2, ln)
                    sig = _float(w, 3, ln)
                    try:
                        rc = _float(w, 4, ln)
                    except PisScriptError:
                        rc = 2.5 * sig
                    table[(i, j)] = (eps, sig, rc)
            elif section == "Atoms":
                aid = _usize(_int(w, 0, ln), ln)
                if aid == 0 or aid > n_atoms:
                    raise PisScriptError("AtomCountMismatch", expected=n_atoms, found=aid)
                types[aid - 1] = _usize(_int(w, 1, ln), ln)
                pos[aid - 1] = [_float(w, 2, ln), _float(w, 3, ln), _float(w, 4, ln)]
            elif section == "Velocities":
                aid = _usize(_int(w, 0, ln), ln)
                if aid == 0 or aid > n_atoms:
                    raise PisScriptError("AtomCountMismatch", expected=n_atoms, found=aid)
                vel[aid - 1] = [_float(w, 1, ln), _float(w, 2, ln), _float(w, 3, ln)]
    box = SimulationBox.from_lammps_data(lo["x"], hi["x"], lo["y"], hi["y"], lo["z"], hi["z"])
    ctx.atoms = Atoms(n_atoms, types, np.asarray(masses, dtype=np.float64), pos, vel, box)
    if table:
        ctx.table = table
    if ctx.starting_velocity is not None:
        ctx.starting_velocity.start_velocity = start
    else:
        ctx.starting_velocity = StartVelocity(start_velocity=start)


# ---------------------------------------------------------------------------------
# System::read / contextualize (system.rs:49-175)


def read_script(path: str, base_dir: str | None = None) -> Context:
    """System::read: one command per line, '#' starts a comment, lines with fewer than two
    words are skipped, an unknown command is an error.  Relative data-file paths resolve
    against `base_dir` (default: the working directory, as in the reference)."""
    ctx = Context(base_dir=base_dir if base_dir is not None else os.getcwd())
    try:
        fh = open(path)
    except OSError:
        raise PisScriptError("InputFileError", path=path)
    with fh:
        for n, raw in enumerate(fh, start=1):
            s = raw.strip()
            if not s or s.startswith("#"):
                continue
            s = s.split("#", 1)[0].strip()
            w = s.split()
            if len(w) < 2:
                continue
            cmd = command_from_str(w[0])
            if cmd is None:
                raise PisScriptError("UnknownCommand", command=w[0], line=n)
            run_command(cmd, w[1:], n, ctx)
    return ctx


def contextualize(ctx: Context) -> Context:
    """Starting velocities (this package's generator; 300 K and seed 0 by default) and the
    lj/cut table from pair_style / pair_coeff."""
    a = ctx.atoms
    if a is not None:
        if a.n_atoms == 0:
            raise PisScriptError("NoAtomsDefined")
        sv = ctx.starting_velocity
        if sv is not None and sv.start_velocity:
            t = 300.0 if sv.start_temperature is None else sv.start_temperature
            seed = 0 if sv.seed is None else sv.seed
            m = a.masses[a.type_ids - 1]
            a.velocities = start_velocities(m, t, KB_REFERENCE, np.random.default_rng(seed))
    pa = ctx.potential_args
    if pa is not None:
        style = _get(pa.pair_style_args, 0, pa.pair_style_line)
        if style != "lj/cut":
            raise PisScriptError("UnknownPairStyle", style=style)
        rc_global = _float(pa.pair_style_args, 1, pa.pair_style_line)
        table = {}
        for args, ln in zip(pa.pair_coeff_args, pa.pair_coeff_lines):
            i = _usize(_int(args, 0, ln), ln)
            j = _usize(_int(args, 1, ln), ln)
            eps = _float(args, 2, ln)
            sig = _float(args, 3, ln)
            try:
                rc = _float(args, 4, ln)
            except PisScriptError:
                rc = rc_global
            table[(i, j)] = (eps, sig, rc)
        ctx.table = table
    return ctx


def to_spec(ctx: Context, skin: float = 0.3) -> SystemSpec:
    """The engine's system description of a contextualized run.  Table entries stored
    under (i, j) with i > j are never looked up by the reference (potential.rs:181-192)
    and are left out."""
    a = ctx.atoms
    if a is None:
        raise PisScriptError("NoAtomsDefined")
    if not ctx.table:
        raise PisScriptError("PotentialNotInitialized")
    pairs = [(i, j, e, s, r) for (i, j), (e, s, r) in ctx.table.items() if i <= j]
    return SystemSpec(positions=a.positions, type_ids=a.type_ids.astype(np.uint32), masses=a.masses, pairs=pairs,
                      box=a.sim_box, velocities=a.velocities, skin=skin, kb=KB_REFERENCE, dt=ctx.timestep,
                      steps=ctx.steps, name="script")


# ---------------------------------------------------------------------------------
# DumpTraj (src/writers/dump_traj.rs) and Simulation::run (src/simulation.rs)


def rust_float(x: float) -> str:
    """f64 as Rust's `{}` prints it: the shortest digits that read back to the same value,
    never in exponent form ("0", "-0", "52.6", "0.00001", "inf", "NaN")."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    s = format(Decimal(repr(x)), "f")
    if "." in s:
        s = s.rstrip("0").rstrip(".")
    return s


class DumpTraj:
    """Text trajectory, one block per written step (dump_traj.rs:33-74).  The block is
    formatted and written by libpis_hip's host writer (pis_dump_write_block), which prints
    numbers exactly as `rust_float` does, on several threads."""

    def __init__(self, path: str, threads: int | None = None):
        self.path = path
        # File::create (dump_traj.rs:19-21): created or truncated
        self.fd = os.open(path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        self.threads = threads if threads is not None else min(16, os.cpu_count() or 1)

    def write_step(self, positions, type_ids, box: SimulationBox, step: int) -> None:
        import ctypes as C
        from ._capi import check, lib

        typ = np.ascontiguousarray(type_ids, dtype=np.uint32)
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(len(typ), 3)
        h = np.ascontiguousarray(box.h, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        check(lib().pis_dump_write_block(self.fd, int(step), len(typ), pos.ctypes.data_as(dp),
                                         typ.ctypes.data_as(C.POINTER(C.c_uint32)), h.ctypes.data_as(dp),
                                         self.threads))

    def close(self):
        if self.fd >= 0:
            os.close(self.fd)
            self.fd = -1


class AsyncFrames:
    """Trajectory frames of a running engine with asynchronous D2H: each frame's positions
    are snapshotted into one of two pinned buffers and travel while the next steps run; a
    frame is formatted (on a helper thread) once the next frame is requested or at close()."""

    def __init__(self, eng, dumper: DumpTraj, type_ids, k: int = 0):
        from .engine import PinnedArray

        self.eng, self.dumper, self.type_ids, self.k = eng, dumper, type_ids, k
        n = eng.specs[k].n
        self.buf = [PinnedArray((n, 3)), PinnedArray((n, 3))]
        self.slot = 0
        self.pending = None  # (step, box, slot) whose transfer is in flight
        self.writer = None

    def _start_pending(self):
        if self.pending is None:
            return
        step, box, slot = self.pending
        self.pending = None
        if self.writer is not None:  # the frame before used the other buffer
            self.writer.join()
        self.eng.snapshot_wait()
        self.writer = threading.Thread(target=self.dumper.write_step,
                                       args=(self.buf[slot].array, self.type_ids, box, step))
        self.writer.start()

    def frame(self, step: int, box: SimulationBox) -> None:
        # the writer of two frames back, which read this frame's buffer, was joined here
        self._start_pending()
        self.eng.snapshot_positions(self.k, self.buf[self.slot])
        self.pending = (step, box, self.slot)
        self.slot ^= 1

    def close(self) -> None:
        self._start_pending()
        if self.writer is not None:
            self.writer.join()
            self.writer = None
        for b in self.buf:
            b.free()

    def abort(self) -> None:
        """After an error: no further frame is written; the writer and the transfer in
        flight finish before the pinned buffers are freed."""
        if self.writer is not None:
            self.writer.join()
            self.writer = None
        self.pending = None
        try:
            self.eng.snapshot_wait()
        except Exception:
            pass
        for b in self.buf:
            b.free()


def thermo_line(step: int, th) -> str:
    """`{} {:.3} {:.3} {:.3} {:.3} {:.3}` of step, PE, KE, H, T, P (simulation.rs:83-86)."""
    return f"{step} {th[0]:.3f} {th[1]:.3f} {th[2]:.3f} {th[3]:.3f} {th[4]:.3f}"


def run(ctx: Context, precision: str = "fp64", skin: float = 0.3, print_fn=print, dump: bool = True):
    """Simulation::run on the GPU: initial compute, then the steps in chunks of the dump
    cadence, a thermo line per step and a trajectory block every dump_step steps.  A
    block's positions reach the host while the device runs the next chunk, and its text is
    written by a helper thread meanwhile.  Returns the thermo records [steps, 7]."""
    from .engine import Engine

    ens = ctx.ensemble
    spec = to_spec(ctx, skin)
    nd = max(1, ctx.dump_args.dump_step)
    dumper = DumpTraj(os.path.join(ctx.base_dir, ctx.dump_args.file_name)) if dump else None
    out = []
    try:
        with Engine(spec, precision) as eng:
            eng.set_state(0, spec.positions, spec.velocities)
            if dumper is not None:
                dumper.write_step(spec.positions, spec.type_ids, spec.box, 0)
            pe0 = float(eng.compute_potential()[0])
            print_fn(f"0 {rust_float(pe0)}")
            nh = ctx.nh_chain_args
            if ens == "nvt":
                eng.nvt_init(nh["start_temperature"], nh["end_temperature"], nh["tau"], 3)
            elif ens == "npt":
                # MTKBarostat::new_from_args (npt.rs:66-87): the chain's start temperature
                mtk = ctx.mtk_barostat_args
                eng.npt_init(nh["start_temperature"], nh["end_temperature"], nh["tau"], mtk["start_pressure"],
                             mtk["tau"], 3)
            _run_chunks(eng, ctx, ens, spec, nd, dumper, print_fn, out)
    finally:
        if dumper is not None:
            dumper.close()
    return np.concatenate(out) if out else np.zeros((0, 7))


def _run_chunks(eng, ctx, ens, spec, nd, dumper, print_fn, out):
    """The steps in chunks ending at every dump_step: thermo lines, trajectory frames."""
    frames = None
    done = 0
    try:
        while done < ctx.steps:
            k = min(nd - done % nd, ctx.steps - done)
            if ens == "nvt":
                th = eng.run_nvt(ctx.timestep, k, ramp=(done, ctx.steps))[:, 0, :]
            elif ens == "npt":
                th = eng.run_npt(ctx.timestep, k, ramp=(done, ctx.steps))[:, 0, :]
            else:
                th = eng.run_nve(ctx.timestep, k)[:, 0, :]
            for j in range(k):
                print_fn(thermo_line(done + j + 1, th[j]))
            out.append(th)
            done += k
            if dumper is not None and done % nd == 0:
                box = spec.box
                if ens == "npt":  # the barostat moved the box
                    box = SimulationBox(eng.npt_box(0)[0].T.reshape(9))
                if frames is None:
                    frames = AsyncFrames(eng, dumper, spec.type_ids)
                frames.frame(done, box)
    except BaseException:
        if frames is not None:
            frames.abort()
        raise
    if frames is not None:
        frames.close()


__all__ = ["PisScriptError", "Context", "read_script", "read_data", "contextualize", "to_spec", "run",
           "DumpTraj", "rust_float", "thermo_line", "command_from_str", "run_command"]
