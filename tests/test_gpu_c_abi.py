"""The reference's time loop driven from plain C99 through include/actflow.h
(tests/c_abi/step_loop.c): the per-step energies and the final state equal
those of the same run made through the Python binding, bitwise, and in FP64
the oracle's to 1e-12."""
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from test_abi import build_c_example

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
def test_c_step_loop_matches_binding(af, orc, tmp_path, prec):
    n, steps, chunk = 64, 50, 16
    exe = build_c_example(tmp_path)
    snap = tmp_path / "final.bin"
    r = subprocess.run([exe, str(n), str(prec), str(steps), str(chunk), str(snap)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    e_c = np.array([float(l.split()[1]) for l in lines if l and l[0].isdigit()])
    m_c = np.array([float(x) for x in next(l for l in lines if l.startswith("M ")).split()[1:]])
    assert len(e_c) == steps
    w_c, _ = af.snapshot_read(snap, n, n)

    p = af.params(n, dt=2e-4, lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5, precision=prec)
    w0 = af.ic_bandlimited(n, seed=11, amp=1.0, kmax=8.0)
    M0 = (0.05 * n * n, -0.02 * n * n)
    run = af.Run(p)
    run.set_state(w0, M0)
    e_py = np.concatenate([run.step(min(chunk, steps - m), energy=True) for m in range(0, steps, chunk)])
    w_py, m_py = run.get_state()
    run.close()
    assert np.array_equal(e_c, e_py) and np.array_equal(m_c, m_py)
    assert np.array_equal(w_c, w_py)   # 17-digit snapshot: lossless
    if prec == 64:
        o = orc.Oracle.from_params(p)
        o.set_state(w0, M0)
        e_o = o.step(steps, energy=True)
        w_o, _ = o.get_state()
        assert rel_l2(w_c, w_o) <= 1e-12
        assert np.allclose(e_c, e_o, rtol=1e-12)
