"""Ensemble engine: runs of mixed size and parameters stepped in shared
launches give bitwise the same results as each run stepped alone; admission
respects the HBM budget."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _cases(af):
    specs = [
        (64, 32, dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5), 1),
        (128, 32, dict(lam=1.5, alpha=-1.1, beta=0.02, gamma0=-0.045, gamma2=2.5e-5), 2),
        (64, 32, dict(lam=4.5, alpha=-0.5, beta=0.006, gamma0=-0.045, gamma2=2.5e-5), 3),
        (128, 64, dict(lam=3.0, alpha=-0.9, beta=0.01, gamma0=-0.045, gamma2=2.5e-5), 4),
        (256, 32, dict(lam=2.0, alpha=-0.7, beta=0.015, gamma0=-0.045, gamma2=2.5e-5), 5),
    ]
    out = []
    for n, prec, kw, seed in specs:
        p = af.params(n, dt=2e-4, precision=prec, **kw)
        out.append((p, af.ic_bandlimited(n, seed=seed, amp=1.5), (0.05 * n * n, -0.02 * n * n)))
    return out


def _engine_table_schedule(af, prec, n):
    """The schedule the engine's tuned table (schedules_batched.txt next to
    the library) names for an n x n grid, or None (the engine then takes the
    single-run choice)."""
    path = os.path.join(os.path.dirname(af.LIB_PATH), "schedules_batched.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        v = line.split("#")[0].split()
        if len(v) >= 10 and (int(v[0]), int(v[1]), int(v[2])) == (prec, n, n):
            sc = tuple(int(x) for x in v[3:])
            if len(sc) == 7:
                sc += (sc[1],)
            return sc + (0,) if len(sc) == 8 else sc
    return None


def test_alone_equals_mix_bitwise(af):
    """Each run stepped in a mix equals the same run stepped alone in an
    engine, bitwise; and equals the single-run API, bitwise where the engine
    and single runs use the same schedule, else to FP rounding."""
    cases = _cases(af)
    eng = af.Engine(1 << 34)
    slots = [eng.add(p, w, M) for p, w, M in cases]
    eng.step(17)
    eng.step(13)
    for (p, w, M), s in zip(cases, slots):
        solo = af.Engine(1 << 34)
        s1 = solo.add(p, w, M)
        solo.step(30)
        wa, ma = solo.get_state(s1)
        wm, mm = eng.get_state(s)
        assert np.array_equal(wa, wm) and np.array_equal(ma, mm), s
        assert np.array_equal(eng.energy(s, 30), solo.energy(s1, 30))
        assert eng.steps_done(s) == 30
        solo.close()
        run = af.Run(p)
        run.set_state(w, M)
        e_run = run.step(30, energy=True)
        wr, mr = run.get_state()
        esched = _engine_table_schedule(af, p.precision, p.nx)
        if esched is None or esched == af.active_schedule(p.precision, p.nx):
            assert np.array_equal(wr, wm) and np.array_equal(mr, mm), s
            assert np.array_equal(eng.energy(s, 30), e_run)
        else:
            tol = 1e-5 if p.precision == 32 else 1e-12
            assert np.linalg.norm(wr - wm) <= tol * np.linalg.norm(wr), s
            assert np.allclose(mr, mm, rtol=tol) and np.allclose(eng.energy(s, 30), e_run, rtol=tol), s
        run.close()
    eng.close()


def test_subset_stepping_and_membership_changes(af):
    cases = _cases(af)
    eng = af.Engine(1 << 34)
    slots = [eng.add(p, w, M) for p, w, M in cases]
    eng.step(5, slots=[slots[0], slots[2]])
    assert eng.steps_done(slots[0]) == 5 and eng.steps_done(slots[1]) == 0
    eng.remove(slots[1])
    eng.step(5)
    p, w, M = cases[0]
    run = af.Run(p)
    run.set_state(w, M)
    run.step(10)
    assert np.array_equal(run.get_state()[0], eng.get_state(slots[0])[0])
    # a slot freed by remove is reused by the next admission
    s_new = eng.add(*cases[1])
    assert s_new == slots[1]
    eng.close()


def test_hbm_budget_admission(af):
    p = af.params(128, precision=32)
    w = af.ic_bandlimited(128, seed=1)
    b = af.device_bytes(p)
    eng = af.Engine(2 * b + b // 2)
    s0 = eng.add(p, w)
    s1 = eng.add(p, w)
    with pytest.raises(af.ActFlowError) as ei:
        eng.add(p, w)
    assert ei.value.status == af.AF_ENOMEM
    assert eng.bytes_in_use == 2 * b and eng.count == 2
    eng.remove(s0)
    eng.add(p, w)
    assert eng.count == 2
    eng.close()


def test_engine_schedule_matches_single_run_to_rounding(af):
    """At 1024^2 FP32 the engine's tuned schedule differs from the single-run
    one (radix, fields per exchange); both integrate the same equation, so
    the states agree to FP32 rounding after a few steps (relative L2 with
    the half-spectrum weights, tolerance 1e-5)."""

    n = 1024
    p = af.params(n, dt=1e-4, precision=32, lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)
    w0 = af.ic_bandlimited(n, seed=5, amp=1.0, kmax=10.0)
    eng = af.Engine(1 << 34)
    s = eng.add(p, w0, (0.0, 0.0))
    eng.step(10)
    we, me = eng.get_state(s)
    eng.close()
    run = af.Run(p)
    run.set_state(w0)
    run.step(10)
    wr, mr = run.get_state()
    run.close()
    assert rel_l2(we, wr) <= 1e-5
    assert np.allclose(me, mr, rtol=1e-5, atol=1e-6 * n * n)


def test_engine_batched_schedule_2048_against_oracle(af, orc):
    """Two 2048^2 FP32 runs in one engine (the tuned batched schedule: 8x4
    bricks with half-brick row tiles, radix-16 columns, 4-column blocks,
    streaming options) against the FP64 oracle after one step, each to FP32
    rounding."""
    N = 2048
    kw = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)
    p = af.params(N, dt=5e-5, precision=32, **kw)
    q = af.AfParams.from_buffer_copy(p)
    q.precision = 64
    eng = af.Engine(1 << 34)
    ws = [af.ic_bandlimited(N, seed=s, amp=2.0, kmax=12.0) for s in (5, 6)]
    slots = [eng.add(p, w, (0.05 * N * N, 0.0)) for w in ws]
    eng.step(1)
    for w0, s in zip(ws, slots):
        w, m = eng.get_state(s)
        o = orc.Oracle.from_params(q, nthreads=16)
        o.set_state(w0, (0.05 * N * N, 0.0))
        o.step(1)
        wo, mo = o.get_state()
        err = np.linalg.norm(w - wo) / np.linalg.norm(wo)
        assert err <= 2e-6, err
        assert np.allclose(m, mo, rtol=2e-5)
    eng.close()


# The mix instantiated for single-launch mode: 256-lane blocks in both passes.
MIX = {1024: (4, 2, 8, 1, 8, 1, 1, 2, 0), 512: (4, 4, 4, 1, 8, 1, 1, 4, 0), 256: (8, 4, 4, 1, 4, 1, 1, 4, 0)}


def test_single_launch_mix_bitwise_and_oracle(af, orc):
    """Single-launch mode (one launch per pass over the 256^2, 512^2 and 1024^2
    classes) against per-class launches with the same schedules: bitwise equal
    states and energies after 13 steps, with an empty descriptor entry in a
    class (a removed run); after one step each size's first run is within
    2e-6 of the FP64 oracle."""
    rng = np.random.default_rng(11)
    specs = []
    for n, cnt in ((256, 5), (512, 3), (1024, 2)):
        for i in range(cnt):
            p = af.params(n, dt=2e-4 if n <= 256 else 1e-4, lam=float(rng.uniform(1, 5)),
                          alpha=float(rng.uniform(-1.2, -0.4)), beta=float(rng.uniform(0.005, 0.02)),
                          gamma0=-0.045, gamma2=2.5e-5, precision=32)
            specs.append((n, p, af.ic_bandlimited(n, seed=int(rng.integers(1, 1 << 30)), amp=0.5, kmax=10.0)))
    try:
        for n, sc in MIX.items():
            af.select_schedule(32, n, n, af.schedules(32, n).index(sc))
        engines = []
        for single in (True, False):
            eng = af.Engine(8 << 30)
            eng.set_single_launch(single)
            slots = [eng.add(p, w) for _, p, w in specs]
            eng.remove(slots[1])            # a hole in the 256^2 class
            engines.append((eng, slots))
        (es, ss), (ef, sf) = engines
        assert es.single_launch_active and not ef.single_launch_active
        for eng, _ in engines:
            eng.step(1)
        firsts = {}
        for k, (n, p, w0) in enumerate(specs):
            firsts.setdefault(n, k)
        for n, k in firsts.items():
            q = af.AfParams.from_buffer_copy(specs[k][1])
            q.precision = 64
            o = orc.Oracle.from_params(q, nthreads=16)
            o.set_state(specs[k][2])
            o.step(1)
            err = rel_l2(es.get_state(ss[k])[0], o.get_state()[0])
            assert err <= 2e-6, (n, err)
        for eng, _ in engines:
            eng.step(12)
        for k in range(len(specs)):
            if k == 1:
                continue
            wa, ma = es.get_state(ss[k])
            wb, mb = ef.get_state(sf[k])
            assert np.array_equal(wa, wb) and np.array_equal(ma, mb), k
            assert np.array_equal(es.energy(ss[k], 13), ef.energy(sf[k], 13)), k
        for eng, _ in engines:
            eng.close()
    finally:
        for n in MIX:
            af.select_schedule(32, n, n, -1)


def test_descriptor_array_growth_holes_and_compaction(af):
    """Membership changes edit a class's descriptor array in place: 20 runs of
    one size (the array grows past its first 16 entries), then 7 removed (holes
    beyond a quarter of the entries: the next step compacts, moving runs to
    other entries) and 2 added (filling holes, reusing retired blocks).  Every
    surviving run equals the same run stepped alone, bitwise, and the bytes
    held follow the membership."""
    n = 64
    rng = np.random.default_rng(21)
    specs = []
    for i in range(22):
        p = af.params(n, dt=2e-4, lam=float(rng.uniform(1, 5)), alpha=float(rng.uniform(-1.2, -0.4)),
                      beta=float(rng.uniform(0.005, 0.02)), gamma0=-0.045, gamma2=2.5e-5, precision=32)
        specs.append((p, af.ic_bandlimited(n, seed=100 + i, amp=1.0, kmax=8.0)))
    per_run = af.device_bytes(specs[0][0])
    eng = af.Engine(1 << 32)
    slots = {i: eng.add(*specs[i]) for i in range(20)}
    eng.step(3)
    steps = {i: 3 for i in slots}
    for i in (0, 3, 4, 9, 12, 17, 19):
        eng.remove(slots.pop(i))
        steps.pop(i)
    assert eng.count == 13 and eng.bytes_in_use == 13 * per_run
    eng.step(4)                      # compaction happens here
    for i in steps:
        steps[i] += 4
    for i in (20, 21):
        slots[i] = eng.add(*specs[i])
        steps[i] = 0
    assert eng.count == 15 and eng.bytes_in_use == 15 * per_run
    eng.step(5)
    for i in steps:
        steps[i] += 5
    for i, s in slots.items():
        solo = af.Engine(1 << 30)
        s1 = solo.add(*specs[i])
        solo.step(steps[i])
        wa, ma = solo.get_state(s1)
        wm, mm = eng.get_state(s)
        assert eng.steps_done(s) == steps[i]
        assert np.array_equal(wa, wm) and np.array_equal(ma, mm), i
        solo.close()
    eng.close()


def test_rectangular_grids_in_one_engine(af, orc):
    """Nx != Ny runs (both orientations, a non-2pi box, both precisions)
    share an engine with a square run: each equals the same run stepped
    through the single-run API, bitwise (rectangular sizes take the built-in
    schedule in both tables), and the FP64 runs equal the oracle to 1e-12."""
    kw = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)
    specs = [(128, 64, 64), (64, 256, 64), (256, 32, 32), (32, 128, 32), (128, 128, 64)]
    cases = []
    for i, (nx, ny, prec) in enumerate(specs):
        p = af.params(nx, ny, lx=2 * np.pi * 2, ly=2 * np.pi * 2 * ny / nx, dt=2e-4, precision=prec, **kw)
        w = af.ic_bandlimited(nx, ny, seed=20 + i, amp=1.0, kmax=6.0)
        cases.append((p, w, (0.03 * nx * ny, -0.01 * nx * ny)))
    eng = af.Engine(1 << 34)
    slots = [eng.add(p, w, M) for p, w, M in cases]
    eng.step(9)
    eng.step(11)
    for (p, w, M), s in zip(cases, slots):
        we, me = eng.get_state(s)
        run = af.Run(p)
        run.set_state(w, M)
        run.step(20)
        wr, mr = run.get_state()
        run.close()
        assert np.array_equal(we, wr) and np.array_equal(me, mr), (p.nx, p.ny, p.precision)
        if p.precision == 64:
            o = orc.Oracle.from_params(p)
            o.set_state(w, M)
            o.step(20)
            wo, mo = o.get_state()
            assert rel_l2(we, wo) <= 1e-12, (p.nx, p.ny)
            assert np.allclose(me, mo, rtol=1e-11, atol=1e-11)
    eng.close()


def test_engine_4096_runs_equal_single_runs_bitwise(af):
    """Two 4096^2 FP32 runs and a 256^2 run in one engine: at 4096^2 the
    engine takes the single-run table's schedule (no batched line), whose row
    pass uses half-brick row tiles when the table names option 64 (the two
    halves of a brick row in sibling blocks, ordered over the whole launch of
    both runs); each 4096^2 run equals the same run stepped alone through the
    single-run API, bitwise, mean velocity and energies included."""
    N = 4096
    kw = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)
    p = af.params(N, dt=5e-5, precision=32, **kw)
    ws = [af.ic_bandlimited(N, seed=s, amp=2.0, kmax=12.0) for s in (31, 32)]
    small = af.params(256, dt=2e-4, precision=32, **kw)
    eng = af.Engine(1 << 34)
    slots = [eng.add(p, w, (0.05 * N * N, -0.02 * N * N)) for w in ws]
    eng.add(small, af.ic_bandlimited(256, seed=33, amp=1.0))
    eng.step(3)
    for w0, s in zip(ws, slots):
        we, me = eng.get_state(s)
        ee = eng.energy(s, 3)
        run = af.Run(p)
        run.set_state(w0, (0.05 * N * N, -0.02 * N * N))
        er = run.step(3, energy=True)
        wr, mr = run.get_state()
        run.close()
        assert np.array_equal(we, wr) and np.array_equal(me, mr), s
        assert np.array_equal(ee, er), s
    eng.close()
    print(f"\n4096^2 single-run schedule {af.active_schedule(32, N)}")
