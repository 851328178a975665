"""GPU: an input script run end to end (System::read -> contextualize -> Simulation::run)
with the device engine, against the oracle running the same system step by step: thermo
lines and trajectory blocks."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from pis_amd import script as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


FIX = {"nve": "", "nvt": "fix t all nvt temp 100.0 120.0 1.0\n",
       "npt": "fix t all npt temp 100.0 120.0 1.0 iso 0.02 0.02 5.0\n"}


@pytest.mark.parametrize("ensemble", ["nve", "nvt", "npt"])
def test_script_run_matches_oracle(tmp_path, ensemble):
    fix = FIX[ensemble]
    inp = tmp_path / "in.pis"
    inp.write_text(f"timestep 0.01\nread_data {os.path.join(GOLDEN, 'example.txt')}\n"
                   f"velocity all create 100.0 7\n{fix}dump d all atom 10 traj.lammpstrj\nrun 40\n")
    ctx = S.contextualize(S.read_script(str(inp), base_dir=str(tmp_path)))
    spec = S.to_spec(ctx)
    lines = []
    th = S.run(ctx, "fp64", print_fn=lines.append)
    assert len(lines) == 41 and lines[0].startswith("0 ")
    # the oracle, step by step, with the trajectory at every 10th step
    o = po.OSystem.from_spec(spec)
    hb = None
    if ensemble == "nvt":
        ref = o.run_nvt(0.01, 40, 100.0, 120.0, 1.0, 3, po.MGR_OFFSET, kb=spec.kb)
        frames = None
    elif ensemble == "npt":
        ref, hb = o.run_npt(0.01, 40, 100.0, 120.0, 1.0, 0.02, 5.0, 3, po.MGR_OFFSET, kb=spec.kb)
        frames = None
    else:
        o.compute(po.MGR_OFFSET)
        ref, frames = [], {}
        for t in range(1, 41):
            pe = o.verlet_step_nve(0.01, po.MGR_OFFSET)
            ke = o.kinetic_energy()
            ref.append([pe, ke, pe + ke, o.temperature(ke, spec.kb), o.pressure(ke)])
            if t % 10 == 0:
                frames[t] = o.pos.copy()
        ref = np.array(ref)
    scale = np.abs(ref[:, 0]).max()
    assert np.abs(th[:, :3] - ref[:, :3]).max() <= 1e-9 * scale
    for t in range(40):
        words = lines[t + 1].split()
        assert int(words[0]) == t + 1
        assert np.allclose([float(w) for w in words[1:]], ref[t], rtol=0, atol=1.5e-3)
    # trajectory: blocks at 0, 10, 20, 30, 40 in the reference's layout
    text = (tmp_path / "traj.lammpstrj").read_text().splitlines()
    n = spec.n
    block = 9 + n
    assert len(text) == 5 * block
    for b, step in enumerate((0, 10, 20, 30, 40)):
        head = text[b * block: b * block + 9]
        assert head[0] == "ITEM: TIMESTEP" and head[1] == str(step) and head[3] == str(n)
        if hb is not None and step:
            # the barostat's box: BOX BOUNDS print the diagonal of h (dump_traj.rs:44-50)
            want = hb[step - 1].reshape(3, 3).T
            for k in range(3):
                lo, hi = head[5 + k].split()
                assert lo == "0" and abs(float(hi) - want[k, k]) <= 1e-9 * want[k, k]
        rows = np.array([[float(v) for v in r.split()[2:]] for r in text[b * block + 9:(b + 1) * block]])
        if frames is not None and step in frames:
            L = spec.box.lengths
            d = (rows - frames[step] + L / 2) % L - L / 2
            assert np.abs(d).max() <= 1e-9


def test_position_snapshots_are_taken_behind_queued_steps():
    """pis_snapshot_positions: the frame holds the positions after the steps queued before
    it, arrives while later steps run, a second call waits for the first, and a batched
    context snapshots the system asked for."""
    from pis_amd import Engine, workloads as W
    from pis_amd.engine import PinnedArray

    a = W.lj_liquid(8, 0.8442, 0.9, seed=71, rcut=2.5, skin=0.3, dt=0.004)
    b = W.lj_liquid(6, 0.8442, 1.1, seed=72, rcut=2.5, skin=0.3, dt=0.004)
    with Engine([a, b], "fp64") as eng:
        eng.load()
        eng.compute_potential()
        eng.run_nve(a.dt, 10)
        ref = eng.get_positions(0)
        buf = [PinnedArray((a.n, 3)), PinnedArray((a.n, 3))]
        eng.snapshot_positions(0, buf[0])
        eng.run_nve(a.dt, 5)  # runs while the frame travels
        eng.snapshot_wait()
        assert np.array_equal(buf[0].array, ref)
        later = eng.get_positions(0)
        assert not np.array_equal(later, ref)
        eng.snapshot_positions(0, buf[0])
        eng.snapshot_positions(0, buf[1])
        eng.snapshot_wait()
        assert np.array_equal(buf[0].array, later) and np.array_equal(buf[1].array, later)
        other = PinnedArray((b.n, 3))
        eng.snapshot_positions(1, other)
        eng.snapshot_wait()
        assert np.array_equal(other.array, eng.get_positions(1))
        with pytest.raises(ValueError):
            eng.snapshot_positions(1, buf[0])  # wrong size
        for p in (*buf, other):
            p.free()


def test_script_run_stopped_mid_run_leaves_whole_frames(tmp_path):
    """An error inside the step loop (here the thermo printer raising at step 23) ends the
    run: the frame whose positions are still in flight is dropped, the writer of the
    previous one finishes, and the trajectory holds only complete blocks."""
    inp = tmp_path / "in.pis"
    inp.write_text(f"timestep 0.01\nread_data {os.path.join(GOLDEN, 'example.txt')}\n"
                   "velocity all create 100.0 7\ndump d all atom 5 traj.lammpstrj\nrun 40\n")
    ctx = S.contextualize(S.read_script(str(inp), base_dir=str(tmp_path)))
    lines = []

    def printer(line):
        lines.append(line)
        if len(lines) == 24:  # "0 PE", then steps 1..23
            raise RuntimeError("stop")

    with pytest.raises(RuntimeError):
        S.run(ctx, "fp64", print_fn=printer)
    blocks = (tmp_path / "traj.lammpstrj").read_text().split("ITEM: TIMESTEP\n")[1:]
    steps = [int(b.splitlines()[0]) for b in blocks]
    assert steps[:4] == [0, 5, 10, 15] and len(steps) <= 5
    for b in blocks:
        assert len(b.splitlines()) == 1 + 2 + 1 + 3 + 1 + 500
