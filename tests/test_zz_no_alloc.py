"""af_run_diagnostics_noalloc (static per-device scratch, no allocation)
against af_run_diagnostics, through the C ABI: two calls in a row, the second
on a smaller grid with fewer shells (the shared scratch keeps the first
call's contents, so stale bins or partials would show), then the size limit.
Energy and enstrophy must be bitwise equal (same block partials, summed in
the same order); spectrum bins are FP64 atomic sums, equal to 1e-13."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
JW = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def _check(run, nshell):
    e0, z0, s0 = run.diagnostics(nshell=nshell)
    e1, z1, s1 = run.diagnostics_noalloc(nshell=nshell)
    assert e1 == e0 and z1 == z0, (e0, e1, z0, z1)
    if nshell:
        assert np.allclose(s1, s0, rtol=1e-13, atol=1e-13 * s0.max())
        assert s1[nshell - 1] == pytest.approx(s0[nshell - 1], rel=1e-13, abs=1e-300)
    e2, z2, _ = run.diagnostics_noalloc(nshell=0)
    assert e2 == e0 and z2 == z0


def test_noalloc_diagnostics_match(af):
    big = af.Run(af.params(512, dt=2e-4, precision=64, **JW))
    big.set_state(af.ic_bandlimited(512, seed=3, amp=2.0, kmax=40.0), (0.1 * 512 * 512, -0.2 * 512 * 512))
    big.step(5)
    _check(big, 362)
    small = af.Run(af.params(256, dt=2e-4, precision=32, **JW))
    small.set_state(af.ic_bandlimited(256, seed=4, amp=1.0, kmax=20.0), (0.05 * 256 * 256, 0.0))
    small.step(3)
    _check(small, 40)
    with pytest.raises(af.ActFlowError) as ei:
        small.diagnostics_noalloc(nshell=4097)
    assert ei.value.status == af.AF_EINVAL
    big.close()
    small.close()
