"""Serving loop: continuous batching with HBM-budget admission.  Each
request's result is bitwise the same served alone and in the mix; the
budget is never exceeded; snapshots arrive; latency percentiles exist."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reqs(af):
    rng = np.random.default_rng(3)
    out = []
    t = 0.0
    for i in range(10):
        n = [64, 128, 256][i % 3]
        t += rng.exponential(0.005)
        p = af.params(n, dt=2e-4, lam=float(rng.uniform(1, 5)), alpha=float(rng.uniform(-1.2, -0.4)),
                      beta=float(rng.uniform(0.005, 0.02)), gamma0=-0.045, gamma2=2.5e-5, precision=32)
        out.append(af.Request(arrival_s=t, params=p, horizon=int(rng.integers(20, 120)), snapshot_every=25,
                              seed=100 + i, ic_amp=0.5, ic_kmax=10.0))
    return out


def test_serve_alone_equals_mix(af):
    reqs = _reqs(af)
    snaps = {}
    res, st, finals = af.serve(reqs, hbm_budget=1 << 33, max_tick_steps=16,
                               on_snapshot=lambda q, s, w, M: snaps.setdefault(q, []).append(s), keep_final=True)
    assert st["completed"] == len(reqs)
    for i, r in enumerate(reqs):
        assert res[i]["status"] == 0 and res[i]["finish_s"] >= res[i]["admit_s"] >= r.arrival_s - 1e-9
        assert res[i]["snapshots"] == r.horizon // 25
        assert snaps.get(i, []) == [25 * (k + 1) for k in range(r.horizon // 25)]
    for i in (0, 4, 8):
        r = reqs[i]
        alone_req = af.Request(arrival_s=0.0, params=r.params, horizon=r.horizon, snapshot_every=0, seed=r.seed,
                               ic_amp=r.ic_amp, ic_kmax=r.ic_kmax)
        ra, sa, fa = af.serve([alone_req], hbm_budget=1 << 33, max_tick_steps=7, keep_final=True)
        assert np.array_equal(fa[0], finals[i]), i
        assert ra[0]["M"] == res[i]["M"]
    assert st["step_p99_s"] >= st["step_p50_s"] > 0
    assert st["ttc_p99_s"] >= st["ttc_p50_s"] > 0


def test_serve_respects_budget(af):
    reqs = _reqs(af)
    p256 = af.params(256, precision=32)
    budget = 2 * af.device_bytes(p256) + 4 * af.device_bytes(af.params(128, precision=32))
    res, st, _ = af.serve(reqs, hbm_budget=budget, max_tick_steps=16)
    assert st["peak_bytes"] <= budget
    assert st["completed"] == len(reqs)
    waited = [r["admit_s"] - q.arrival_s for r, q in zip(res, reqs)]
    assert max(waited) >= 0


def test_serve_rejects_request_larger_than_budget(af):
    p = af.params(512, precision=32)
    reqs = [af.Request(arrival_s=0.0, params=p, horizon=5, seed=1)]
    res, st, _ = af.serve(reqs, hbm_budget=af.device_bytes(p) // 2, max_tick_steps=4)
    assert res[0]["status"] == af.AF_ENOMEM and st["completed"] == 0


def test_serve_nan_request_fails_alone(af):
    """A request whose state blows up (the parameters of test_nan_watchdog)
    ends with AF_ENAN and no finish time; the other requests complete with
    the same results as in a mix without it."""
    reqs = _reqs(af)[:6]
    bad = af.Request(arrival_s=reqs[1].arrival_s,
                     params=af.params(64, dt=0.5, lam=50.0, alpha=-50.0, beta=-5.0, precision=32),
                     horizon=200, snapshot_every=0, seed=1, ic_amp=100.0, ic_kmax=8.0)
    res, st, finals = af.serve(reqs + [bad], hbm_budget=1 << 33, max_tick_steps=16, keep_final=True)
    assert res[-1]["status"] == af.AF_ENAN and res[-1]["finish_s"] < 0 <= res[-1]["admit_s"]
    assert st["completed"] == len(reqs)
    ref, _, ref_finals = af.serve(reqs, hbm_budget=1 << 33, max_tick_steps=16, keep_final=True)
    for i in range(len(reqs)):
        assert res[i]["status"] == 0
        assert np.array_equal(finals[i], ref_finals[i]), i
        assert res[i]["M"] == ref[i]["M"]
