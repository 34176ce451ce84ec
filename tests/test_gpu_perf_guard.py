"""Guards against silent slow paths: the tuned tables are picked up, and the
1024^2 FP32 single run stays below 0.2 ms/step (about twice its measured
0.092 ms/step on MI355X)."""
import os
import time

import pytest

pytestmark = pytest.mark.gpu
JW = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def _table_line(af, name, prec, n):
    path = os.path.join(os.path.dirname(af.LIB_PATH), name)
    for line in open(path):
        v = line.split("#")[0].split()
        if len(v) >= 12 and (int(v[0]), int(v[1]), int(v[2])) == (prec, n, n):
            return tuple(int(x) for x in v[3:12])
    return None


def test_tuned_schedule_is_active(af):
    for prec in (32, 64):
        want = _table_line(af, "schedules.txt", prec, 1024)
        if want is None:
            pytest.skip("no tuned line for 1024^2")
        assert af.active_schedule(prec, 1024) == want


def test_1024_fp32_step_time(af):
    n = 1024
    run = af.Run(af.params(n, dt=1e-4, precision=32, **JW))
    run.set_state(af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0))
    run.step(50, energy=True)
    run.get_state()
    t0 = time.perf_counter()
    run.step(500, energy=True)
    run.get_state()
    ms = (time.perf_counter() - t0) / 500 * 1e3
    run.close()
    print(f"\n1024^2 FP32: {ms:.4f} ms/step")
    assert ms < 0.2, ms
