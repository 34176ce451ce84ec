"""The Hermitian projection of the kx = 0 column (reality_func,
ActFlow_V3-2.cu:1228-1243, applied before every accumulate at .cu:745, 766,
788, 809) and the mean-velocity ODE (.cu:729-732, 756-821) on the GPU, from
inputs where they act: a kx = 0 column whose rows y and Ny - y are
independent, and a mean flow with sigma(0) = 0.

Tolerances: FP64 <= 1e-12 relative L2 on omega against the oracle, FP32
(FP32 fields and FFTs, FP64 factors, reductions and ODE) <= 2e-6 against the
FP64 oracle, checked after each of steps 1-4 (stage 1 projects omega_old,
whose integrating factor is even in ky; stages 2-4 project acc)."""
import numpy as np
import pytest

from conftest import rel_l2
from test_oracle_physics import broken_column_state, projection_closed_form

pytestmark = pytest.mark.gpu
JW = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def test_projection_value_closed_form_gpu(af):
    """lambda = beta = 0 from a broken kx = 0 column: one step is exactly
    P(E w0) (the oracle's closed-form test, on the HIP path)."""
    for nx, ny in ((256, 256), (128, 256)):
        dt, alpha = 1e-2, -0.8
        p = af.params(nx, ny, dt=dt, lam=0.0, alpha=alpha, beta=0.0, precision=64)
        w0 = broken_column_state(nx, ny)
        run = af.Run(p)
        run.set_state(w0, (0.0, 0.0))
        run.step(1)
        w, _ = run.get_state()
        run.close()
        kx = np.arange(nx // 2 + 1)[None, :]
        ky = np.fft.fftfreq(ny, 1.0 / ny)[:, None]
        k2 = kx ** 2 + ky ** 2
        E = np.exp((-(1 - k2) ** 2 - alpha) * dt)
        ref = projection_closed_form(w0, E)
        assert np.abs(w - ref).max() <= 1e-14 * np.abs(ref).max(), (nx, ny)
        assert np.abs(E * w0 - ref)[:, 0].max() > 1e-3 * np.abs(ref).max()


def _four_steps(af, orc, p, w0, M0, tol, nthreads=8):
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    o = orc.Oracle.from_params(q, nthreads=nthreads)
    o.set_state(w0, M0)
    run = af.Run(p)
    run.set_state(w0, M0)
    errs = []
    try:
        for k in range(4):
            e = run.step(1, energy=True)
            eo = o.step(1, energy=True)
            w, m = run.get_state()
            wo, mo = o.get_state()
            errs.append(rel_l2(w, wo))
            assert errs[-1] <= tol, (k + 1, errs)
            assert abs(e[0] - eo[0]) <= tol * abs(eo[0]), (k + 1, e, eo)
            assert np.allclose(m, mo, rtol=10 * tol, atol=10 * tol * np.abs(mo).max()), (k + 1, m, mo)
            # the projection really acted: the kx = 0 column is Hermitian now
            col = w[:, 0]
            n = len(col)
            assert np.abs(col[1:n // 2] - np.conj(col[n - 1:n // 2:-1])).max() <= 1e-5 * np.abs(col).max()
    finally:
        run.close()
    return errs


@pytest.mark.parametrize("prec", [64, 32])
def test_broken_symmetry_every_schedule_256(af, orc, prec):
    """256^2, every instantiated schedule (width-1 and width-2 column blocks,
    radix 4/8/16, one or several fields per LDS exchange, every option bit)."""
    n = 256
    p = af.params(n, dt=2e-4, precision=prec, **JW)
    w0 = broken_column_state(n, n, seed=5, amp=0.5, kmax=12)
    M0 = (0.2 * n * n, -0.1 * n * n)
    tol = 1e-12 if prec == 64 else 2e-6
    widths = set()
    try:
        for i, sc in enumerate(af.schedules(prec, n)):
            af.select_schedule(prec, n, n, i)
            widths.add(sc[7])
            errs = _four_steps(af, orc, p, w0, M0, tol, nthreads=4)
            assert max(errs) <= tol, (i, sc, errs)
    finally:
        af.select_schedule(prec, n, n, -1)
    if prec == 64:
        assert 1 in widths, "no width-1 column-block schedule exercised"


@pytest.mark.parametrize("n,prec", [(1024, 32), (1024, 64), (2048, 32), (2048, 64)])
def test_broken_symmetry_tuned_large(af, orc, n, prec):
    """The tuned schedules: 1024^2 (the headline FP32 schedule) and 2048^2
    (radix-16 column passes)."""
    p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, precision=prec, **JW)
    w0 = broken_column_state(n, n, seed=7, amp=0.5, kmax=12)
    M0 = (0.05 * n * n, 0.03 * n * n)
    errs = _four_steps(af, orc, p, w0, M0, 1e-12 if prec == 64 else 2e-6, nthreads=16)
    print(f"\n{n}^2 FP{prec} broken kx=0 column, steps 1-4: rel L2 vs oracle {['%.1e' % e for e in errs]}")


def test_mean_ode_closed_form_gpu(af):
    """sigma(0) = 0, omega = 0: the GPU's FP64 mean update converges at order
    4 to |M(t)|^2 = |M0|^2 / (1 + 2 beta |M0|^2 t / N^4) (classical RK4 once
    e = e2 = 1), both mean schemes."""
    N, beta, T = 64, 0.5, 2.0
    M0 = np.array([0.9, -0.3]) * N * N
    ref = M0 * np.sqrt(1.0 / (1.0 + 2.0 * beta * np.dot(M0, M0) * T / (N * N) ** 2))
    for scheme in (0, 1):
        errs = []
        for dt in (0.05, 0.025, 0.0125):
            p = af.params(N, dt=dt, lam=2.0, alpha=-1.0, beta=beta, precision=64, mean_scheme=scheme)
            run = af.Run(p)
            run.set_state(np.zeros((N, N // 2 + 1), complex), M0)
            run.step(int(round(T / dt)))
            w, M = run.get_state()
            run.close()
            assert np.abs(w).max() < 1e-9 * N * N
            errs.append(np.linalg.norm(M - ref) / np.linalg.norm(ref))
        r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
        assert errs[0] < 1e-8 and 13.0 < r1 < 19.0 and 13.0 < r2 < 19.0, (scheme, errs, r1, r2)
