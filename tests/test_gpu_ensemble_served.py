"""The ensemble (64 instances of three sizes) and the served workload (a
request mix from 128^2 to 2048^2) checked against the oracle, not only
against themselves.

Ensemble: 32 x 256^2 + 24 x 512^2 + 8 x 1024^2 FP32 instances with seeded
alpha in [-1.2, -0.4], beta in [0.005, 0.02], lambda0 in [1, 5], plus one
FP64 copy of the first instance of each size, all in one engine.  After one
step the first FP32 instance of each size is within 2e-6 (relative L2 on
omega) of the FP64 oracle; after 20 steps the FP64 copies equal the oracle
to 1e-12 and sampled FP32 instances equal the same instance stepped alone in
an engine, bitwise.  The band-limited IC has amplitude 0.5 (not the 1e-3 of
the benchmark), so the nonlinear terms are O(1) in these comparisons.

Served: requests of every size 128^2 ... 2048^2 with snapshots; one FP64
request (128^2, 200 steps) whose final state equals the oracle to 1e-12,
and one FP32 request (256^2, 100 steps) within 2e-5 of the FP64 oracle; the
2048^2 request served alone gives bitwise its result in the mix."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
GAM = dict(gamma0=-0.045, gamma2=2.5e-5)


def _oracle_run(orc, p, w0, steps, M0=(0.0, 0.0), nthreads=16):
    q = orc.Oracle.from_params(p, nthreads=nthreads)
    if p.precision != 64:
        raise ValueError("oracle reference runs are FP64")
    q.set_state(w0, M0)
    q.step(steps)
    return q.get_state()


def _as64(af, p):
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    return q


def test_mixed_ensemble_64_runs_against_oracle(af, orc):
    rng = np.random.default_rng(7)
    mix = [(256, 32), (512, 24), (1024, 8)]
    eng = af.Engine(200 << 30)
    inst = {}     # n -> list of (slot, params, ic)
    for n, cnt in mix:
        for _ in range(cnt):
            p = af.params(n, dt=2e-4 if n <= 256 else 1e-4, lam=float(rng.uniform(1, 5)),
                          alpha=float(rng.uniform(-1.2, -0.4)), beta=float(rng.uniform(0.005, 0.02)),
                          precision=32, **GAM)
            ic = af.ic_bandlimited(n, seed=int(rng.integers(1, 1 << 30)), amp=0.5, kmax=10.0)
            inst.setdefault(n, []).append((eng.add(p, ic), p, ic))
    copies = {n: eng.add(_as64(af, lst[0][1]), lst[0][2]) for n, lst in inst.items()}
    assert eng.count == 64 + 3

    eng.step(1)
    for n, lst in inst.items():
        slot, p, ic = lst[0]
        wo, _ = _oracle_run(orc, _as64(af, p), ic, 1)
        err = rel_l2(eng.get_state(slot)[0], wo)
        assert err <= 2e-6, (n, err)

    eng.step(19)
    for n, lst in inst.items():
        slot, p, ic = lst[0]
        wo, mo = _oracle_run(orc, _as64(af, p), ic, 20)
        w64, m64 = eng.get_state(copies[n])
        assert rel_l2(w64, wo) <= 1e-12, (n, rel_l2(w64, wo))
        assert np.allclose(m64, mo, rtol=1e-11, atol=1e-11 * max(1.0, np.abs(mo).max()))
        for k in (0, len(lst) // 2, len(lst) - 1):
            slot, p, ic = lst[k]
            solo = af.Engine(8 << 30)
            s1 = solo.add(p, ic)
            solo.step(20)
            wa, ma = solo.get_state(s1)
            wm, mm = eng.get_state(slot)
            assert np.array_equal(wa, wm) and np.array_equal(ma, mm), (n, k)
            solo.close()
    eng.close()


def test_served_mix_128_to_2048_against_oracle(af, orc):
    rng = np.random.default_rng(5)
    reqs = []
    t = 0.0
    for i, n in enumerate([128, 256, 512, 1024, 2048, 128, 256, 512]):
        t += rng.exponential(0.01)
        lam, alpha, beta = float(rng.uniform(1, 5)), float(rng.uniform(-1.2, -0.4)), float(rng.uniform(0.005, 0.02))
        p = af.params(n, dt=af.served_dt(n, lam, alpha, beta), lam=lam, alpha=alpha, beta=beta, precision=32, **GAM)
        reqs.append(af.Request(arrival_s=t, params=p, horizon=int(rng.integers(40, 120)), snapshot_every=25,
                               seed=200 + i, ic_amp=0.5, ic_kmax=10.0))
    p64 = af.params(128, dt=1e-4, lam=3.5, alpha=-0.8, beta=0.01, precision=64, **GAM)
    reqs.append(af.Request(arrival_s=t, params=p64, horizon=200, snapshot_every=50, seed=301, ic_amp=0.5,
                           ic_kmax=10.0))
    p32 = af.params(256, dt=1e-4, lam=3.5, alpha=-0.8, beta=0.01, precision=32, **GAM)
    reqs.append(af.Request(arrival_s=t, params=p32, horizon=100, snapshot_every=0, seed=302, ic_amp=0.5,
                           ic_kmax=10.0))
    snaps = {}
    res, st, finals = af.serve(reqs, hbm_budget=200 << 30, max_tick_steps=16,
                               on_snapshot=lambda q, s, w, M: snaps.setdefault(q, []).append(s), keep_final=True)
    assert st["completed"] == len(reqs), [r["status"] for r in res]
    for i, r in enumerate(reqs):
        assert res[i]["status"] == 0
        if r.snapshot_every:
            assert snaps.get(i, []) == list(range(r.snapshot_every, r.horizon + 1, r.snapshot_every)), i
    # FP64 request: the oracle to 1e-12; FP32 request: the FP64 oracle to 2e-5
    w0 = af.ic_bandlimited(128, seed=301, amp=0.5, kmax=10.0)
    wo, mo = _oracle_run(orc, p64, w0, 200)
    assert rel_l2(finals[-2], wo) <= 1e-12, rel_l2(finals[-2], wo)
    assert np.allclose(res[-2]["M"], mo, rtol=1e-11, atol=1e-9)
    w0 = af.ic_bandlimited(256, seed=302, amp=0.5, kmax=10.0)
    wo, _ = _oracle_run(orc, _as64(af, p32), w0, 100)
    err = rel_l2(finals[-1], wo)
    print(f"\nserved FP32 256^2, 100 steps: rel L2 vs FP64 oracle {err:.2e}")
    assert err <= 2e-5, err
    # the 2048^2 request alone: bitwise its result in the mix
    k = 4
    alone = af.Request(arrival_s=0.0, params=reqs[k].params, horizon=reqs[k].horizon, snapshot_every=0,
                       seed=reqs[k].seed, ic_amp=0.5, ic_kmax=10.0)
    ra, _, fa = af.serve([alone], hbm_budget=200 << 30, max_tick_steps=7, keep_final=True)
    assert np.array_equal(fa[0], finals[k])
