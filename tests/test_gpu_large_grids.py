"""2048^2 and 4096^2 single runs with the lambda0 values of the scaling
workload (lambda0 in {2, 3.5, 5}, dt = 5e-5, band-limited IC): the FP32 run
tracks the FP64 run of the same size and parameters over the first steps
(relative L2 of omega and the per-step energies), and stays finite."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
BASE = dict(alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def _run(af, n, prec, lam, steps):
    run = af.Run(af.params(n, dt=5e-5, lam=lam, precision=prec, **BASE))
    run.set_state(af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0))
    e = run.step(steps, energy=True)
    w, m = run.get_state()
    run.close()
    return w, m, e


@pytest.mark.parametrize("lam", [2.0, 3.5, 5.0])
def test_2048_fp32_tracks_fp64(af, lam):
    w32, m32, e32 = _run(af, 2048, 32, lam, 200)
    w64, m64, e64 = _run(af, 2048, 64, lam, 200)
    assert np.all(np.isfinite(e32)) and np.all(np.isfinite(w32.view(np.float64)))
    assert rel_l2(w32, w64) <= 1e-4, rel_l2(w32, w64)
    assert np.max(np.abs(e32 - e64) / np.abs(e64)) <= 1e-5


def test_4096_fp32_tracks_fp64(af):
    w32, m32, e32 = _run(af, 4096, 32, 5.0, 40)
    w64, m64, e64 = _run(af, 4096, 64, 5.0, 40)
    assert np.all(np.isfinite(e32))
    assert rel_l2(w32, w64) <= 1e-4, rel_l2(w32, w64)
    assert np.max(np.abs(e32 - e64) / np.abs(e64)) <= 1e-5
