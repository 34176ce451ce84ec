"""GPU parity: the HIP step (through the C ABI) against the CPU oracle on the
same seeded inputs, plus size-independent properties at full sizes.

Tolerances: FP64 <= 1e-12 relative L2 on omega (real-space norm via Parseval
weights) after 100 steps; FP32 (FP32 fields and FFTs, FP64 integrating
factors and reductions) against the FP64 oracle <= 2e-5 after 100 steps
(measured ~5e-6)."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

JW = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def oracle_for(orc, p, nthreads=4):
    return orc.Oracle.from_params(p, nthreads=nthreads)


def compare(af, orc, p, w0, M0, chunks, tol_w, tol_e, nthreads=4):
    run = af.Run(p)
    run.set_state(w0, M0)
    o = oracle_for(orc, p if p.precision == 64 else _as64(af, p), nthreads)
    o.set_state(w0, M0)
    worst = 0.0
    for n in chunks:
        eg = run.step(n, energy=True)
        eo = o.step(n, energy=True)
        wg, mg = run.get_state()
        wo, mo = o.get_state()
        err = rel_l2(wg, wo)
        worst = max(worst, err)
        assert err <= tol_w, (n, err)
        assert np.abs(eg - eo).max() <= tol_e * np.abs(eo).max()
        assert np.allclose(mg, mo, rtol=tol_w * 10, atol=tol_w * 10 * max(1.0, np.abs(mo).max()))
    run.close()
    return worst


def _as64(af, p):
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    return q


@pytest.mark.parametrize("seed", [1, 2])
def test_parity_fp64_256(af, orc, seed):
    """256^2, dt = 2e-4, band-limited IC (amp 1e-3, |k| <= 10), 100 steps."""
    p = af.params(256, dt=2e-4, precision=64, **JW)
    w0 = af.ic_bandlimited(256, seed=seed, amp=1e-3, kmax=10.0)
    compare(af, orc, p, w0, (0.0, 0.0), (1, 9, 90), 1e-12, 1e-12)


def test_parity_fp64_256_energy_series(af, orc):
    """256^2 FP64 at its full length: the per-step energy series of 2000 steps and
    the final state match the oracle (the flow is in its linear growth phase,
    where rounding differences do not grow relative to the field)."""
    p = af.params(256, dt=2e-4, precision=64, **JW)
    w0 = af.ic_bandlimited(256, seed=1, amp=1e-3, kmax=10.0)
    worst = compare(af, orc, p, w0, (0.0, 0.0), (2000,), 1e-11, 1e-11, nthreads=8)
    print(f"\n256^2 FP64, 2000 steps: rel L2 vs oracle {worst:.2e}")


def test_parity_fp64_developed(af, orc):
    """O(1) amplitude so the nonlinear and cubic terms are large, with a mean flow."""
    N = 128
    p = af.params(N, dt=2e-4, precision=64, **JW)
    w0 = af.ic_bandlimited(N, seed=3, amp=2.0, kmax=12.0)
    compare(af, orc, p, w0, (0.4 * N * N, -0.25 * N * N), (1, 99), 1e-12, 1e-12)


@pytest.mark.parametrize("nx,ny,dealias,mean_scheme", [(64, 128, 0, 0), (128, 64, 1, 1), (32, 32, 0, 1),
                                                       (256, 32, 1, 0)])
def test_parity_fp64_reference_operator(af, orc, nx, ny, dealias, mean_scheme):
    """The reference's own operator -(1+lap)^2 - alpha (.cu:1030) in a
    non-2pi box, rectangular grids, both masks and both mean schemes."""
    p = af.params(nx, ny, lx=2 * np.pi * 4, ly=2 * np.pi * 4 * ny / nx, dt=5e-3, lam=7.0, alpha=-0.8, beta=0.01,
                  precision=64, dealias=dealias, mean_scheme=mean_scheme)
    w0 = af.ic_bandlimited(nx, ny, seed=11, amp=0.5, kmax=8.0)
    compare(af, orc, p, w0, (0.3 * nx * ny, 0.1 * nx * ny), (3, 47), 1e-12, 1e-12)


def test_every_grid_shape_one_step(af, orc):
    """Every supported grid, Nx and Ny powers of two from 32 to 4096 (64
    shapes, each its own pair of row and column kernels with the default or
    tuned schedule), one step in FP64 and in FP32 from the same O(1) state with
    a mean flow, against one FP64 oracle step: FP64 to 1e-12, FP32 to 2e-6."""
    sizes = [32, 64, 128, 256, 512, 1024, 2048, 4096]
    nthreads = min(16, os.cpu_count() or 1)
    for nx in sizes:
        for ny in sizes:
            dt = 1e-4 if max(nx, ny) <= 1024 else 5e-5
            w0 = af.ic_bandlimited(nx, ny, seed=13, amp=2.0, kmax=10.0)
            M0 = (0.05 * nx * ny, -0.02 * nx * ny)
            o = oracle_for(orc, af.params(nx, ny, dt=dt, precision=64, **JW), nthreads)
            o.set_state(w0, M0)
            eo = o.step(1, energy=True)
            wo, mo = o.get_state()
            for prec, tol in ((64, 1e-12), (32, 2e-6)):
                run = af.Run(af.params(nx, ny, dt=dt, precision=prec, **JW))
                run.set_state(w0, M0)
                e = run.step(1, energy=True)
                w, m = run.get_state()
                run.close()
                assert rel_l2(w, wo) <= tol, (nx, ny, prec, rel_l2(w, wo))
                assert abs(e[0] - eo[0]) <= tol * abs(eo[0]), (nx, ny, prec)
                assert np.allclose(m, mo, rtol=10 * tol), (nx, ny, prec)


def test_parity_fp32_error_stated(af, orc):
    p = af.params(256, dt=2e-4, precision=32, **JW)
    w0 = af.ic_bandlimited(256, seed=1, amp=2.0, kmax=10.0)
    worst = compare(af, orc, p, w0, (0.0, 0.0), (1, 9, 90), 2e-5, 2e-6)
    print(f"FP32 vs FP64 oracle, 256^2, 100 steps: rel L2 {worst:.2e}")


def test_linear_exactness_gpu(af):
    """lambda = beta = 0: omega_n = E^n omega_0 (T4)."""
    N, dt = 64, 1e-2
    p = af.params(N, dt=dt, lam=0.0, alpha=-0.8, beta=0.0, precision=64)
    w0 = af.ic_bandlimited(N, seed=5, amp=1.0, kmax=6.0)
    run = af.Run(p)
    run.set_state(w0)
    run.step(7)
    w, M = run.get_state()
    kx = np.arange(N // 2 + 1)[None, :]
    ky = np.fft.fftfreq(N, 1.0 / N)[:, None]
    k2 = kx ** 2 + ky ** 2
    E = np.exp((-(1 - k2) ** 2 + 0.8) * dt)
    assert np.abs(w - w0 * E ** 7).max() <= 1e-13 * np.abs(w0).max()


def test_mean_ode_gpu(af, orc):
    """omega = 0, M != 0: the 2-D machinery reduces to the scalar recurrence."""
    from test_oracle_physics import mean_recurrence
    N, beta, dt = 64, 0.05, 1e-2
    for scheme in (0, 1):
        p = af.params(N, dt=dt, lam=2.0, alpha=-0.8, beta=beta, precision=64, mean_scheme=scheme)
        run = af.Run(p)
        M0 = np.array([0.9, -0.3]) * N * N
        run.set_state(np.zeros((N, N // 2 + 1), complex), M0)
        run.step(20)
        w, M = run.get_state()
        ref = mean_recurrence(M0, beta, N, dt, np.exp(-0.2 * dt), np.exp(-0.1 * dt), 20, scheme == 1)   # c0 = -1 - alpha
        assert np.allclose(M, ref, rtol=1e-13)
        assert np.abs(w).max() < 1e-9 * N * N


def test_graph_and_direct_launch_identical(af):
    p = af.params(128, dt=2e-4, precision=32, **JW)
    w0 = af.ic_bandlimited(128, seed=4, amp=1.0)
    out = []
    for g in (True, False):
        run = af.Run(p)
        run.use_graph(g)
        run.set_state(w0, (10.0, -5.0))
        e = run.step(25, energy=True)
        out.append((run.get_state(), e))
        run.close()
    (wa, ma), ea = out[0]
    (wb, mb), eb = out[1]
    assert np.array_equal(wa, wb) and np.array_equal(ma, mb) and np.array_equal(ea, eb)


def test_profiled_steps_are_ordinary_steps(af):
    """The per-kernel timing hook (each event-timed step behind a 4-step
    graph) advances the run by 5 steps per timed step, with the results of
    ordinary steps, bitwise; the energy ring and the step count follow."""
    p = af.params(128, dt=2e-4, precision=32, **JW)
    w0 = af.ic_bandlimited(128, seed=5, amp=1.0)
    a, b = af.Run(p), af.Run(p)
    for r in (a, b):
        r.set_state(w0, (3.0, 1.0))
    row_ms, col_ms, tot_ms = a.profile(3)
    assert a.steps_done == 15 and row_ms > 0 and col_ms > 0 and tot_ms > 0
    b.step(15)
    ea, eb = a.step(7, energy=True), b.step(7, energy=True)
    (wa, ma), (wb, mb) = a.get_state(), b.get_state()
    a.close()
    b.close()
    assert np.array_equal(wa, wb) and np.array_equal(ma, mb) and np.array_equal(ea, eb)


def test_state_roundtrip_and_nyquist(af):
    """set/get round trip; a state with a nonzero kx = Nx/2 column (not
    represented on this backend) is refused, not silently changed."""
    N = 64
    rng = np.random.default_rng(0)
    w = rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))
    for prec in (64, 32):
        run = af.Run(af.params(N, precision=prec))
        with pytest.raises(af.ActFlowError) as ei:
            run.set_state(w, (1.25, -3.5))
        assert ei.value.status == af.AF_EUNSUPPORTED
        with pytest.raises(af.ActFlowError):
            run.step(1)               # no state loaded
        ref = w.copy()
        ref[:, -1] = 0
        run.set_state(ref, (1.25, -3.5))
        back, M = run.get_state()
        if prec == 64:
            assert np.array_equal(back, ref)
        else:
            assert np.array_equal(back, ref.real.astype(np.float32) + 1j * ref.imag.astype(np.float32))
        assert np.array_equal(M, [1.25, -3.5])
        eng = af.Engine(1 << 30)
        with pytest.raises(af.ActFlowError) as ei:
            eng.add(af.params(N, precision=prec), w)
        assert ei.value.status == af.AF_EUNSUPPORTED and eng.count == 0 and eng.bytes_in_use == 0
        eng.close()


def test_nan_watchdog(af):
    """A blow-up is reported as AF_ENAN with the step count (check_func,
    .cu:884-897)."""
    N = 64
    p = af.params(N, dt=0.5, lam=50.0, alpha=-50.0, beta=-5.0, precision=32)
    run = af.Run(p)
    run.set_state(af.ic_bandlimited(N, seed=1, amp=100.0, kmax=8.0))
    with pytest.raises(af.ActFlowError) as ei:
        run.step(200)
    assert ei.value.status == af.AF_ENAN


def _energy_spectral(w, M, N, lx=2 * np.pi):
    """1/2 <|u|^2> from omega^ by Biot-Savart in spectral space + Parseval."""
    kx = 2 * np.pi / lx * np.arange(N // 2 + 1)[None, :]
    ky = 2 * np.pi / lx * np.fft.fftfreq(N, 1.0 / N)[:, None]
    k2 = kx ** 2 + ky ** 2
    k2[0, 0] = 1.0
    u2 = np.abs(w) ** 2 / k2
    u2[0, 0] = 0.0
    wt = np.full(w.shape, 2.0)
    wt[:, 0] = 1
    wt[:, -1] = 1
    return 0.5 * np.sum(wt * u2) / (N * N) ** 2 + 0.5 * np.dot(M, M) / (N * N) ** 2


@pytest.mark.parametrize("N,prec", [(512, 64), (1024, 32), (2048, 32), (2048, 64), (4096, 32), (4096, 64)])
def test_full_size_properties(af, N, prec):
    """At full sizes: w^[0] stays 0, masked modes stay 0, the reported stage-1
    energy equals 1/2<|u|^2> of the state (Parseval), finite evolution."""
    p = af.params(N, dt=5e-5, precision=prec, **JW)
    run = af.Run(p)
    run.set_state(af.ic_bandlimited(N, seed=1, amp=0.5, kmax=10.0), (0.1 * N * N, 0.0))
    run.step(5)
    w, M = run.get_state()
    e = run.step(1, energy=True)[0]
    assert w[0, 0] == 0
    x = np.arange(N // 2 + 1)[None, :]
    y = np.minimum(np.arange(N), N - np.arange(N))[:, None]
    assert np.all(w[9 * (x ** 2 + y ** 2) > N * N] == 0)
    ref = _energy_spectral(w, M, N)
    assert abs(e - ref) <= (1e-12 if prec == 64 else 2e-6) * ref
    run.close()


def test_parity_fp64_1024_two_steps(af, orc):
    N = 1024
    p = af.params(N, dt=1e-4, precision=64, **JW)
    w0 = af.ic_bandlimited(N, seed=1, amp=2.0, kmax=12.0)
    compare(af, orc, p, w0, (0.0, 0.0), (2,), 1e-12, 1e-12, nthreads=8)


@pytest.mark.parametrize("N", [1024, 2048, 4096])
@pytest.mark.parametrize("prec", [64, 32])
def test_parity_large_one_step(af, orc, prec, N):
    """The tuned 1024^2 (the headline FP32 schedule), 2048^2 and 4096^2
    schedules (radix 8/16 rows and columns, cache and table options) against
    the oracle directly, from a developed state with a mean flow: FP64 to
    1e-12, FP32 to its rounding, 2e-6 (the FP64 oracle as reference)."""
    p = af.params(N, dt=1e-4 if N <= 1024 else 5e-5, precision=prec, **JW)
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    w0 = af.ic_bandlimited(N, seed=3, amp=2.0, kmax=12.0)
    run = af.Run(p)
    run.set_state(w0, (0.05 * N * N, -0.02 * N * N))
    e = run.step(1, energy=True)
    w, m = run.get_state()
    run.close()
    o = orc.Oracle.from_params(q, nthreads=16)
    o.set_state(w0, (0.05 * N * N, -0.02 * N * N))
    eo = o.step(1, energy=True)
    wo, mo = o.get_state()
    tol = 1e-12 if prec == 64 else 2e-6
    assert rel_l2(w, wo) <= tol, rel_l2(w, wo)
    assert abs(e[0] - eo[0]) <= tol * abs(eo[0])
    assert np.allclose(m, mo, rtol=10 * tol)
