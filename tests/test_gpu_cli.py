"""The drop-in CLI end to end on the GPU: same directory contract as
`ActFlow_V3-2 <dir>` (snapshots, energy.txt, messages, exit codes), and the
resulting state agrees with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "2d_active_matter_simulation_amd", "actflow")

SPEC = """T:
{T}
dt:
0.001
Lx:
6.283185307
Ly:
6.283185307
Nx:
64
Ny:
64
alpha:
-0.8
beta:
0.01
lambda:
3.5
dt_snap:
0.01
initial_condition:
{ic}
Gamma0:
-0.045
Gamma2:
2.5e-5
precision:
64
snapshot_precision:
17
"""


def test_cli_run_and_continue(af, orc, tmp_path):
    d = tmp_path / "sim"
    d.mkdir()
    (d / "specifications.txt").write_text(SPEC.format(T="0.05", ic="bandlimited 3 0.5 8"))
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "program finished successfully" in r.stdout
    snaps = sorted(os.listdir(d))
    for j in range(6):
        assert f"ActTurb_snap{j}.bin" in snaps
    lines = (d / "energy.txt").read_text().splitlines()
    assert len(lines) == 50
    t0, e0 = map(float, lines[0].split())
    assert np.isclose(t0, 0.001)          # F7: step m labelled m*dt
    # oracle from snapshot 0 for 50 steps == snapshot 5
    w0, M0 = af.snapshot_read(d / "ActTurb_snap0.bin", 64, 64)
    s = af.spec_read(str(d))
    o = orc.Oracle(64, lx=s.Lx, ly=s.Ly, dt=s.dt, lam=s.lambda_, beta=s.beta, c0=-s.alpha, c2=-s.gamma0,
                   c4=-s.gamma2)
    o.set_state(w0, M0)
    e = o.step(50, energy=True)
    w5, M5 = af.snapshot_read(d / "ActTurb_snap5.bin", 64, 64)
    wo, mo = o.get_state()
    assert rel_l2(w5, wo) < 1e-12
    assert np.isclose(e[0], e0, rtol=1e-5)   # energy.txt has 6 digits
    # continue: one more unit of time from the last snapshot (index 5)
    (d / "specifications.txt").write_text(SPEC.format(T="0.05", ic="continue 0.02"))
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "continue execution from snapshot 5" in r.stdout
    assert (d / "ActTurb_snap7.bin").exists()
    assert len((d / "energy.txt").read_text().splitlines()) == 50 + 21   # re-runs step 50 (.cu:612)
    # resume parity against an uninterrupted run (.cu:612-613, 718-824): the
    # continued loop starts at m = 5 * snap and steps once before it writes
    # snapshot 5 again, so the continued snapshots 5 and 7 hold the
    # uninterrupted states after 51 and 71 steps (17-digit snapshots: lossless)
    o.step(1)
    w5b, _ = af.snapshot_read(d / "ActTurb_snap5.bin", 64, 64)
    assert rel_l2(w5b, o.get_state()[0]) < 1e-12
    o.step(20)
    w7, M7 = af.snapshot_read(d / "ActTurb_snap7.bin", 64, 64)
    wo7, mo7 = o.get_state()
    assert rel_l2(w7, wo7) < 1e-12
    assert np.allclose(M7, mo7, rtol=1e-11, atol=1e-11)


def test_cli_snapshot_write_failure_mid_run(tmp_path):
    """Snapshot 2 cannot be created (a directory of that name is in the way):
    the run stops with the reference's message and exit code 1, and the files
    are those of a run that stopped there: energy lines up to step 2*snap,
    snapshots 0 and 1, nothing after (the writer thread drops later jobs)."""
    d = tmp_path / "sim"
    d.mkdir()
    (d / "specifications.txt").write_text(SPEC.format(T="0.05", ic="bandlimited 3 0.5 8"))
    (d / "ActTurb_snap2.bin").mkdir()
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Error: unable to create snapshot file" in r.stderr
    assert (d / "ActTurb_snap0.bin").is_file() and (d / "ActTurb_snap1.bin").is_file()
    for j in (3, 4, 5):
        assert not (d / f"ActTurb_snap{j}.bin").exists()
    # jobs after the failed one are never written, whenever the main loop
    # notices the failure: exactly the 20 lines before snapshot 2
    lines = (d / "energy.txt").read_text().splitlines()
    assert len(lines) == 20 and np.isclose(float(lines[-1].split()[0]), 0.02)


def test_cli_nan_exit(tmp_path):
    d = tmp_path / "sim"
    d.mkdir()
    spec = SPEC.format(T="50", ic="bandlimited 1 100 8").replace("-0.8\nbeta", "-50\nbeta").replace(
        "dt:\n0.001", "dt:\n0.5").replace("dt_snap:\n0.01", "dt_snap:\n5")
    (d / "specifications.txt").write_text(spec)
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "array contains nan values after" in r.stderr
