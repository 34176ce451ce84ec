"""Every instantiated schedule (brick shape, radix, LDS grouping, options) of a
size computes the same step: at 256^2 and 512^2 FP64 against the oracle to
1e-12 after 20 steps and FP32 against the FP64 oracle to 2e-6; at 1024^2,
2048^2 and 4096^2 the same bounds after one step."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
JW = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n", [256, 512])
def test_every_schedule_matches_oracle(af, orc, prec, n):
    p = af.params(n, dt=2e-4, precision=prec, **JW)
    w0 = af.ic_bandlimited(n, seed=9, amp=2.0, kmax=12.0)
    M0 = (0.2 * n * n, -0.1 * n * n)
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    o = orc.Oracle.from_params(q, nthreads=4)
    o.set_state(w0, M0)
    eo = o.step(20, energy=True)
    wo, mo = o.get_state()
    tol = 1e-12 if prec == 64 else 2e-6
    try:
        for i, sc in enumerate(af.schedules(prec, n)):
            af.select_schedule(prec, n, n, i)
            run = af.Run(p)
            run.set_state(w0, M0)
            e = run.step(20, energy=True)
            w, m = run.get_state()
            run.close()
            assert rel_l2(w, wo) <= tol, (i, sc, rel_l2(w, wo))
            assert np.abs(e - eo).max() <= tol * np.abs(eo).max(), (i, sc)
            assert np.allclose(m, mo, rtol=10 * tol), (i, sc)
    finally:
        af.select_schedule(prec, n, n, -1)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_every_schedule_one_step_at_large_sizes(af, orc, prec, n):
    """Every instantiated schedule of the sizes the tuner chooses among for
    the benchmarks (radix 4/8/16, grouped or single-field exchanges, brick
    shapes, cache options, paired rows, half-brick row tiles): one step from
    an O(1) band-limited state with a mean flow against one FP64 oracle step,
    FP64 to 1e-12 and FP32 to 2e-6 (relative L2 on omega); energy and mean
    velocity alike."""
    p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, precision=prec, **JW)
    w0 = af.ic_bandlimited(n, seed=11, amp=2.0, kmax=12.0)
    M0 = (0.05 * n * n, -0.02 * n * n)
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    o = orc.Oracle.from_params(q, nthreads=min(16, os.cpu_count() or 1))
    o.set_state(w0, M0)
    eo = o.step(1, energy=True)
    wo, mo = o.get_state()
    tol = 1e-12 if prec == 64 else 2e-6
    scheds = af.schedules(prec, n)
    try:
        for i, sc in enumerate(scheds):
            af.select_schedule(prec, n, n, i)
            run = af.Run(p)
            run.set_state(w0, M0)
            e = run.step(1, energy=True)
            w, m = run.get_state()
            run.close()
            assert rel_l2(w, wo) <= tol, (i, sc, rel_l2(w, wo))
            assert abs(e[0] - eo[0]) <= tol * abs(eo[0]), (i, sc)
            assert np.allclose(m, mo, rtol=10 * tol), (i, sc)
    finally:
        af.select_schedule(prec, n, n, -1)
    print(f"\n{len(scheds)} schedules of {n}^2 FP{prec} match the oracle")


def test_narrow_blocks_reduce_every_partial(af, orc):
    """A column block narrower than a wave (Ny = 256, W = 1, radix 8: 32
    lanes) must still fold every row-block partial of the mean-velocity RHS."""
    n = 256
    idx = [i for i, sc in enumerate(af.schedules(64, n)) if sc[7] == 1]
    if not idx:
        pytest.skip("no width-1 column-block schedule at 256^2")
    p = af.params(n, dt=2e-4, precision=64, **JW)
    w0 = af.ic_bandlimited(n, seed=2, amp=2.0, kmax=12.0)
    M0 = (0.3 * n * n, 0.2 * n * n)
    try:
        af.select_schedule(64, n, n, idx[0])
        run = af.Run(p)
        run.set_state(w0, M0)
        run.step(10)
        w, m = run.get_state()
    finally:
        af.select_schedule(64, n, n, -1)
    o = orc.Oracle.from_params(p, nthreads=4)
    o.set_state(w0, M0)
    o.step(10)
    wo, mo = o.get_state()
    assert np.allclose(m, mo, rtol=1e-12)
    assert rel_l2(w, wo) <= 1e-12


def test_factor_table_option_at_radix16(af):
    """Option 16 (integrating factors read from the per-run FP32 table by a
    radix-16 column pass) computes the step of the same schedule evaluating
    them in the kernel: the table holds the same FP64-evaluated factors
    rounded once, so the states agree to FP32 rounding."""
    n = 2048
    scheds = af.schedules(32, n)
    pairs = [(i, scheds.index(sc[:8] + (sc[8] | 16,))) for i, sc in enumerate(scheds)
             if sc[4] == 16 and not sc[8] & 16 and sc[:8] + (sc[8] | 16,) in scheds]
    if not pairs:
        pytest.skip("no radix-16 column schedule with the table option at 2048^2")
    p = af.params(n, dt=5e-5, precision=32, **JW)
    w0 = af.ic_bandlimited(n, seed=4, amp=2.0, kmax=30.0)
    states = {}
    try:
        for i in {k for pr in pairs[:2] for k in pr}:
            af.select_schedule(32, n, n, i)
            run = af.Run(p)
            run.set_state(w0, (0.1 * n * n, 0.0))
            run.step(3)
            states[i] = run.get_state()[0]
            run.close()
    finally:
        af.select_schedule(32, n, n, -1)
    for a, b in pairs[:2]:
        assert rel_l2(states[b], states[a]) <= 1e-6, (scheds[a], scheds[b], rel_l2(states[b], states[a]))
