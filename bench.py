#!/usr/bin/env python3
"""Benchmark of the ActFlow time step on one MI355X.

A "step" is one full IF-RK step (4 stages = 8 fused kernel launches) of the
James-Wilczek vorticity equation plus the k=0 mean-velocity ODE and the
per-step energy, on synthetic input (seeded band-limited initial condition)
already resident in HBM.

Headline workload (`value`): single run, 1024 x 1024 grid, FP32 fields and
FFTs with FP64 integrating factors and reductions, dt = 1e-4, lambda0 = 3.5,
alpha = -0.8, beta = 0.01, Gamma0 = -0.045, Gamma2 = 2.5e-5.  The same
workload in FP64 and at 2048^2 / 4096^2 is reported under "extra".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt-steps/s @1024²/2048², FP32 vs FP64; % HBM roofline; served p50/p99"
HBM_PEAK_GBS = 8000.0
PHYS = dict(lx=2 * np.pi, lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def field_bytes(n, prec):
    """One stored complex field: Ny x Nx/2 elements (Nyquist column not stored)."""
    return n * (n // 2) * (8 if prec == 32 else 16)


def step_model_bytes(n, prec=32):
    """Bytes per step of the 3-pass model of one IF-RK step: 76 transfers of an
    Ny x (Nx/2+1) complex field (what a schedule of separate FFT, pointwise and
    update passes moves; this design moves 50 transfers of Ny x Nx/2)."""
    return 76 * n * (n // 2 + 1) * (8 if prec == 32 else 16)


# Compulsory HBM traffic per launch in complex-field transfers (read + write):
#   row pass: reads v_x, v_y, omega (mixed), writes P, Q             -> 5
#   column pass: reads P, Q, omega_old (stages 1-3), acc (stages 2-4);
#                writes acc (or the new state) and v_x, v_y, omega   -> 7, 8, 8, 7
#                (+ the FP32 integrating-factor table, see col_fields)
ROW_FIELDS = 5.0
COL_FIELDS = (7 + 8 + 8 + 7) / 4.0
# The ensemble mix instantiated for single-launch mode: per grid, the schedule
# (by, bx, row radix, row group, col radix, col fwd group, col inv group,
# col width, options) with 256-lane blocks in both passes.
MIX_SCHEDULES = {1024: (4, 2, 8, 1, 8, 1, 1, 2, 0), 512: (4, 4, 4, 1, 8, 1, 1, 4, 0), 256: (8, 4, 4, 1, 4, 1, 1, 4, 0)}


def col_fields(af, n, prec):
    """Column-pass transfers per launch, plus the integrating-factor table an
    FP32 run reads when its column radix is <= 8 or its schedule sets option
    16 (4 B per mode per plane: 2 planes at stages 1 and 3, 1 at stage 2, in
    units of one FP32 field)."""
    extra = 0.0
    sc = af.active_schedule(prec, n)
    if prec == 32 and (sc[4] <= 8 or sc[8] & 16):
        extra = (2 + 1 + 2 + 0) / 4.0 * 0.5
    return COL_FIELDS + extra


def read_traffic(workload_tag):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE), from the committed summary profiles/traffic_<tag>.txt."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_tag}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    return vals or None


def rel_l2(a, b):
    """Real-space relative L2 of two half spectra (Parseval weights)."""
    wt = np.full(a.shape, 2.0)
    wt[:, 0] = 1.0
    wt[:, -1] = 1.0
    return float(np.sqrt(np.sum(wt * np.abs(a - b) ** 2) / np.sum(wt * np.abs(b) ** 2)))


def fp32_error(af, p32, ic, M0, steps, w32):
    """Relative L2 on omega of an FP32 result against the same run in FP64
    (the FP64 GPU path, equal to the FP64 oracle to <= 1e-12 per step in the
    parity tests), after the same number of steps from the same state."""
    q = af.AfParams.from_buffer_copy(p32)
    q.precision = 64
    run = af.Run(q)
    run.set_state(ic, M0)
    run.step(steps)
    w64, _ = run.get_state()
    run.close()
    return {"fp32_rel_l2_vs_fp64": rel_l2(w32, w64), "after_steps": int(steps)}


def run_workload(af, n, prec, steps, warmup, seed=1, host_io=False, device_warmup=None):
    """One single run: `warmup` untimed steps, then `steps` timed steps.
    device_warmup (optional): an untimed call that keeps the device busy right
    before the warm-up steps (the run's own creation leaves it idle)."""
    p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, precision=prec, **PHYS)
    ic = af.ic_bandlimited(n, seed=seed, amp=1e-3, kmax=10.0)
    run = af.Run(p)
    run.set_state(ic)
    if device_warmup:
        device_warmup()
    if warmup:
        run.step(warmup)
    t0 = time.perf_counter()
    e = run.step(steps, energy=True)          # returns after the device finished
    wall = time.perf_counter() - t0
    assert np.all(np.isfinite(e))
    # per-kernel durations: HIP events at each kernel's own start and end on
    # the run's stream, over steps launched right after the timed ones (the
    # device is in the same state; the timed steps run as graphs, which take
    # no per-launch events, so each event-timed step follows a 4-step graph
    # that keeps the device busy while the host enqueues it)
    kprof = max(10, min(steps, 200))
    row_ms, col_ms, tot_ms = run.profile(kprof)
    err = None
    if prec == 32:
        err = fp32_error(af, p, ic, (0.0, 0.0), run.steps_done, run.get_state()[0])
    io = None
    if host_io:
        # the C ABI hands the state over in host memory (FP64 half spectrum):
        # time a download and an upload of the current state (median of 5)
        gets, sets = [], []
        for _ in range(5):
            t0 = time.perf_counter()
            w, m = run.get_state()
            gets.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            run.set_state(w, m)
            sets.append(time.perf_counter() - t0)
        io = dict(get_state_ms=float(np.median(gets)) * 1e3, set_state_ms=float(np.median(sets)) * 1e3,
                  state_bytes=int(w.nbytes + m.nbytes))
    run.close()
    return dict(n=n, prec=prec, steps=steps, wall=wall, ms_per_step=wall / steps * 1e3,
                value=n * n * steps / wall, row_ms=row_ms, col_ms=col_ms, prof_steps=kprof,
                prof_ms_per_step=tot_ms / (5 * kprof), host_io=io, fp32_error=err)


def pass_bytes_ms(af, r):
    """Algorithmic bytes per launch and mean event-timed launch duration of the
    two passes of run r."""
    fb = field_bytes(r["n"], r["prec"])
    return {"col_pass": (col_fields(af, r["n"], r["prec"]) * fb, r["col_ms"]),
            "row_pass": (ROW_FIELDS * fb, r["row_ms"])}


def roofline(af, r):
    kernels = pass_bytes_ms(af, r)
    name = max(kernels, key=lambda k: kernels[k][1])
    alg, ms = kernels[name]
    achieved = alg / (ms * 1e-3) / 1e9
    return name, alg, ms, achieved


def ensemble_leg(af, steps=200, seed=7, measured_sequential=False):
    """The mixed ensemble: 64 instances (32 x 256^2, 24 x 512^2, 8 x 1024^2), seeded random
    alpha, beta, lambda0, FP32, stepped together in shared launches; compared
    with the same instances run one after another through the single-run path
    (timed per size class and summed, or with measured_sequential every one
    of the 64 runs stepped alone and timed, and its final state compared)."""
    rng = np.random.default_rng(seed)
    mix = [(256, 32), (512, 24), (1024, 8)]
    eng = af.Engine(200 << 30)
    pts = 0
    runs = []
    for n, cnt in mix:
        for i in range(cnt):
            p = af.params(n, dt=2e-4 if n <= 256 else 1e-4, lam=float(rng.uniform(1, 5)),
                          alpha=float(rng.uniform(-1.2, -0.4)), beta=float(rng.uniform(0.005, 0.02)),
                          gamma0=-0.045, gamma2=2.5e-5, precision=32)
            ic = af.ic_bandlimited(n, seed=int(rng.integers(1, 1 << 30)), amp=1e-3, kmax=10.0)
            runs.append((eng.add(p, ic), p, ic))
            pts += n * n
    eng.step(10)
    t0 = time.perf_counter()
    eng.step(steps)
    wall = time.perf_counter() - t0
    value = pts * steps / wall
    out = {}
    if measured_sequential:
        seq, worst = 0.0, 0.0
        for slot, p, ic in runs:
            run = af.Run(p)
            run.set_state(ic)
            run.step(10)
            t0 = time.perf_counter()
            run.step(steps)
            seq += time.perf_counter() - t0
            w, _ = run.get_state()
            run.close()
            we, _ = eng.get_state(slot)
            worst = max(worst, float(np.linalg.norm(we - w) / np.linalg.norm(w)))
        # the engine's batched schedules differ from the single-run ones, so the
        # two agree to FP32 rounding grown over the steps, not bitwise
        out["max_rel_diff_engine_vs_alone"] = worst
    # The same 64 runs with the 256-lane schedules of the instantiated mix:
    # per-class launches (concurrent streams) against single-launch mode (one
    # launch per pass over all three classes)
    modes = {}
    try:
        for n, sc in MIX_SCHEDULES.items():
            af.select_schedule(32, n, n, af.schedules(32, n).index(sc))
        for single in (False, True):
            e2 = af.Engine(200 << 30)
            e2.set_single_launch(single)
            for _, p, ic in runs:
                e2.add(p, ic)
            e2.step(10)
            t0 = time.perf_counter()
            e2.step(steps)
            w2 = time.perf_counter() - t0
            modes["single_launch" if single else "per_class_launches"] = {
                "value": pts * steps / w2, "ms_per_step": w2 / steps * 1e3, "single_launch_active": e2.single_launch_active}
            e2.close()
    finally:
        for n in MIX_SCHEDULES:
            af.select_schedule(32, n, n, -1)
    out["mix_schedules"] = {"schedules": {f"{n}^2": list(sc) for n, sc in MIX_SCHEDULES.items()}, **modes}
    # FP32 error of the first instance of each size class against FP64
    first = {}
    for slot, p, ic in runs:
        first.setdefault(p.nx, (slot, p, ic))
    out["fp32_error"] = {f"{n}^2": fp32_error(af, p, ic, (0.0, 0.0), steps + 10, eng.get_state(slot)[0])
                         for n, (slot, p, ic) in first.items()}
    eng.close()
    if not measured_sequential:
        seq = 0.0
        for n, cnt in mix:
            r = run_workload(af, n, 32, max(50, steps // 2), 10)
            seq += cnt * (r["ms_per_step"] * 1e-3) * steps
    fb = sum(cnt * field_bytes(n, 32) for n, cnt in mix)
    return {"instances": 64, "mix": "32x256^2 + 24x512^2 + 8x1024^2 FP32", "steps": steps,
            "sequential": "measured, each run alone" if measured_sequential else "per-size timings x counts",
            **out,
            "value": value, "ms_per_step": wall / steps * 1e3, "sequential_value": pts * steps / seq,
            "speedup_vs_sequential": seq / wall,
            "step_bytes_50_transfers": 50 * fb,
            "hbm_frac_50_transfers": 50 * fb / (wall / steps) / 1e9 / HBM_PEAK_GBS,
            "step_model_frac": sum(cnt * step_model_bytes(n) for n, cnt in mix) / (wall / steps) / 1e9
            / HBM_PEAK_GBS}


def large_grid_leg(af, steps=5000):
    """The large single runs at full length: FP32 runs of 2048^2 and 4096^2 for
    lambda0 in {2, 3.5, 5} (alpha -0.8, beta 0.01, Gamma0 -0.045, Gamma2 2.5e-5),
    dt 5e-5, `steps` steps with per-step energy from the seeded band-limited
    IC; one line per run."""
    out = []
    for n in (2048, 4096):
        for lam in (2.0, 3.5, 5.0):
            p = af.params(n, dt=5e-5, lam=lam, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5, precision=32)
            ic = af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0)
            run = af.Run(p)
            run.set_state(ic)
            t0 = time.perf_counter()
            e = run.step(steps, energy=True)
            wall = time.perf_counter() - t0
            err = fp32_error(af, p, ic, (0.0, 0.0), steps, run.get_state()[0])
            run.close()
            out.append({"grid": [n, n], "lambda0": lam, "steps": steps, "value": n * n * steps / wall,
                        "ms_per_step": wall / steps * 1e3, "energy_first": float(e[0]), "energy_last": float(e[-1]),
                        "finite": bool(np.all(np.isfinite(e))), "fp32_error": err})
            print(json.dumps({"large_grid": out[-1]}), flush=True)
    return out


def batched_leg(af, n, count, steps, seed=3):
    """Large grids stepped as a batch: `count` FP32 instances of n x n with
    lambda0 cycling over {2, 3.5, 5} (other parameters as the headline), in
    shared launches of the ensemble engine (its tuned batched schedule)."""
    eng = af.Engine(200 << 30)
    slots = []
    for i in range(count):
        p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, lam=(2.0, 3.5, 5.0)[i % 3], alpha=-0.8, beta=0.01,
                      gamma0=-0.045, gamma2=2.5e-5, precision=32)
        ic = af.ic_bandlimited(n, seed=seed + i, amp=1e-3, kmax=10.0)
        slots.append((eng.add(p, ic), p, ic))
    eng.step(5)
    t0 = time.perf_counter()
    eng.step(steps)
    wall = time.perf_counter() - t0
    slot, p, ic = slots[0]
    err = fp32_error(af, p, ic, (0.0, 0.0), steps + 5, eng.get_state(slot)[0])
    eng.close()
    fb = field_bytes(n, 32) * count
    return {"grid": [n, n], "instances": count, "steps": steps, "value": count * n * n * steps / wall,
            "ms_per_step": wall / steps * 1e3, "fp32_error": {"instance 0": err},
            "hbm_frac_50_transfers": 50 * fb / (wall / steps) / 1e9 / HBM_PEAK_GBS,
            "step_model_frac": count * step_model_bytes(n) / (wall / steps) / 1e9 / HBM_PEAK_GBS}


def serving_leg(af, rate, requests=16, seed=11, horizon_scale=1.0):
    """The served workload: Poisson arrivals at `rate`/s over a window of requests/rate s
    (about `requests` arrivals); N from {128..2048}^2 with weights
    0.3/0.3/0.2/0.15/0.05; horizon U{1k..50k} steps (times horizon_scale);
    snapshot every 500 steps; alpha, beta, lambda0 as the ensemble; FP32;
    200 GB budget; each request's dt from served_dt (1e-4 up to 1024^2, 5e-5
    at 2048^2, lower where an advective CFL estimate asks).  The FP32 error
    is stated for the first request of the smallest grid: FP32 vs FP64 single
    runs of its parameters after 200 steps (the full horizons outlast the
    flow's decorrelation, where only statistics compare)."""
    rng = np.random.default_rng(seed)
    sizes, wts = [128, 256, 512, 1024, 2048], [0.3, 0.3, 0.2, 0.15, 0.05]
    window = requests / rate
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window:
            break
        n = int(rng.choice(sizes, p=wts))
        lam, alpha, beta = float(rng.uniform(1, 5)), float(rng.uniform(-1.2, -0.4)), float(rng.uniform(0.005, 0.02))
        p = af.params(n, dt=af.served_dt(n, lam, alpha, beta), lam=lam, alpha=alpha, beta=beta,
                      gamma0=-0.045, gamma2=2.5e-5, precision=32)
        h = max(1, int(rng.integers(1000, 50001) * horizon_scale))
        reqs.append(af.Request(arrival_s=t, params=p, horizon=h, snapshot_every=500,
                               seed=int(rng.integers(1, 1 << 30))))
    nsnap = [0]
    res, st, _ = af.serve(reqs, hbm_budget=200 << 30, max_tick_steps=50,
                          on_snapshot=lambda *a: nsnap.__setitem__(0, nsnap[0] + 1))
    failed = sum(1 for r in res if r["status"] != 0)   # e.g. a run whose state went NaN
    out = {"rate_per_s": rate, "window_s": window, "requests": len(reqs), "completed": st["completed"],
           "failed": failed, "horizon": f"U{{1k..50k}} x {horizon_scale}", "throughput": st["throughput"],
           "step_p50_ms": st["step_p50_s"] * 1e3, "step_p99_ms": st["step_p99_s"] * 1e3,
           "ttc_p50_s": st["ttc_p50_s"], "ttc_p99_s": st["ttc_p99_s"], "queue_p50_s": st["queue_p50_s"],
           "queue_p99_s": st["queue_p99_s"], "max_resident": st["max_resident"], "snapshots": nsnap[0],
           "chunks": st["ticks"], "wall_s": st["wall_s"]}
    if reqs:
        i = min(range(len(reqs)), key=lambda j: (reqs[j].params.nx, j))
        r = reqs[i]
        ic = af.ic_bandlimited(r.params.nx, seed=r.seed, amp=r.ic_amp, kmax=r.ic_kmax)
        run = af.Run(r.params)
        run.set_state(ic)
        k = min(200, r.horizon)
        run.step(k)
        err = fp32_error(af, r.params, ic, (0.0, 0.0), k, run.get_state()[0])
        run.close()
        out["fp32_error"] = {"request": i, "grid": [r.params.nx, r.params.ny], **err}
    return out


def cpu_workloads():
    """The oracle on the 256^2 FP64, 1024^2 FP64 / FP32 and 2048^2 FP32
    workloads, on all host cores (up to 16) and on one thread: median time of single steps (energy on)
    after warm-up steps (10 on all cores, 2 on one thread), one JSON line per
    leg.  Steps on one thread are fewer to bound the run."""
    from oracle.oracle_py import Oracle
    af = importlib.import_module("2d_active_matter_simulation_amd")
    threads = min(os.cpu_count() or 1, 16)
    legs = [("256^2 FP64, dt 2e-4, all 2000 steps", 256, 64, 2e-4, 2000, 200),
            ("1024^2 FP64, dt 1e-4, first 200 steps", 1024, 64, 1e-4, 200, 5),
            ("1024^2 FP32, dt 1e-4, first 200 steps", 1024, 32, 1e-4, 200, 5),
            ("2048^2 FP32, dt 5e-5, 10 steps", 2048, 32, 5e-5, 10, 2)]
    out = []
    for name, n, prec, dt, steps, steps1 in legs:
        for nt, k, warm in ((threads, steps, 10), (1, steps1, 2)):
            o = Oracle(n, dt=dt, lam=PHYS["lam"], beta=PHYS["beta"], c0=-PHYS["alpha"], c2=-PHYS["gamma0"],
                       c4=-PHYS["gamma2"], precision=prec, nthreads=nt)
            o.set_state(af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0))
            o.step(warm)
            times = []
            for _ in range(k):
                t0 = time.perf_counter()
                o.step(1, energy=True)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            out.append({"workload": name, "threads": nt, "steps_timed": k, "warmup": warm,
                        "median_ms_per_step": med * 1e3, "value": n * n / med, "unit": "grid-pt-steps/s"})
            print(json.dumps({"cpu_oracle": out[-1]}), flush=True)
    return out


def cpu_baseline(n=1024, prec=32, target_s=10.0):
    """The oracle (C++ restatement of the reference step) on the host cores,
    timed over enough steps for ~target_s of CPU work (at least 3)."""
    from oracle.oracle_py import Oracle
    af = importlib.import_module("2d_active_matter_simulation_amd")
    threads = min(os.cpu_count() or 1, 16)
    o = Oracle(n, dt=1e-4, lam=PHYS["lam"], beta=PHYS["beta"], c0=-PHYS["alpha"], c2=-PHYS["gamma0"],
               c4=-PHYS["gamma2"], precision=prec, nthreads=threads)
    o.set_state(af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0))
    t0 = time.perf_counter()
    o.step(1)
    first = time.perf_counter() - t0
    steps = max(3, int(target_s / max(first, 1e-3)))
    t0 = time.perf_counter()
    o.step(steps, energy=True)
    wall = time.perf_counter() - t0
    out = {"value": n * n * steps / wall, "unit": "grid-pt-steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} steps of the {n}^2 FP{prec} oracle after 1 warm-up step "
                     f"({wall:.1f} s, {threads} OpenMP threads)"}
    # the same on one thread, a few steps (~target_s / 2)
    o1 = Oracle(n, dt=1e-4, lam=PHYS["lam"], beta=PHYS["beta"], c0=-PHYS["alpha"], c2=-PHYS["gamma0"],
                c4=-PHYS["gamma2"], precision=prec, nthreads=1)
    o1.set_state(af.ic_bandlimited(n, seed=1, amp=1e-3, kmax=10.0))
    t0 = time.perf_counter()
    o1.step(1)
    first1 = time.perf_counter() - t0
    steps1 = max(2, int(target_s / 2 / max(first1, 1e-3)))
    t0 = time.perf_counter()
    o1.step(steps1, energy=True)
    wall1 = time.perf_counter() - t0
    out["single_thread"] = {"value": n * n * steps1 / wall1, "cores": 1, "sample": f"{steps1} steps ({wall1:.1f} s)"}
    out["host"] = {"cpu_count": os.cpu_count(), "model": cpu_model()}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip FP64 / 2048^2 / 4096^2 lines")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble and batched legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-workload leg")
    ap.add_argument("--grid", type=int, default=1024,
                    help="grid of the single-run leg (default 1024, the headline; other sizes for profiling)")
    ap.add_argument("--batch", type=int, default=0,
                    help="profiling aid: only step --batch FP32 instances of --grid in the ensemble engine "
                         "(--steps steps) and print their throughput")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 256 MiB (the reference kernel for scaling "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes), then exit")
    ap.add_argument("--serving-requests", type=int, default=0,
                    help="only run the served workload at 10, 50 and 200 requests/s with the full "
                         "U{1k..50k} horizons, each arrival window sized for about this many requests, "
                         "printing one JSON line per rate")
    ap.add_argument("--ensemble-full", type=int, default=0,
                    help="only run the 64-instance ensemble for this many steps and compare it with each "
                         "of the 64 runs stepped alone (measured), printing one JSON line")
    ap.add_argument("--large-grids", type=int, default=0,
                    help="only run single FP32 runs of 2048^2 and 4096^2 for lambda0 in {2, 3.5, 5} for this "
                         "many steps (from the first step, energy on), one JSON line per run")
    ap.add_argument("--cpu-workloads", action="store_true",
                    help="only time the CPU oracle on the 256^2, 1024^2 and 2048^2 workloads (all host cores and one "
                         "thread), one JSON line per leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(instances would shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)
    if args.cpu_workloads:
        print(json.dumps({"host": {"cpu_count": os.cpu_count(), "model": cpu_model()}}), flush=True)
        cpu_workloads()
        return
    if args.copy_calibration:
        import torch
        n = 64 << 20   # floats: 256 MiB per buffer
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied 5 x {n * 4} bytes", flush=True)
        return
    af = importlib.import_module("2d_active_matter_simulation_amd")
    af.lib()

    N = args.grid
    if args.large_grids > 0:
        large_grid_leg(af, args.large_grids)
        return
    if args.ensemble_full > 0:
        print(json.dumps({"ensemble": ensemble_leg(af, steps=args.ensemble_full, measured_sequential=True)}))
        return
    if args.serving_requests > 0:
        for r in (10, 50, 200):
            print(json.dumps({"serving": serving_leg(af, r, requests=args.serving_requests)}),
                  flush=True)
        return
    if args.batch > 0:
        print(json.dumps({"batched": batched_leg(af, N, args.batch, args.steps)}))
        return
    # Device warm-up, untimed: after idle the device needs ~200 steps of work
    # (about 20 ms at 1024^2) before it steps at full speed (measured: a 20-step
    # call right after other GPU work takes 1.88 ms, after 100 ms of idle 1.95
    # ms).  Every leg's own W warm-up steps then follow other GPU work, as in a
    # process that has been stepping, and the legs compare like for like.
    warm = af.Run(af.params(N, dt=1e-4 if N <= 1024 else 5e-5, precision=32, **PHYS))
    warm.set_state(af.ic_bandlimited(N, seed=2, amp=1e-3, kmax=10.0))
    warm.step(200)
    # The other single-run workloads go next, the headline after them.
    extra = {}
    if not args.no_extra:
        for n, prec, steps in ((1024, 64, max(200, args.steps // 4)), (2048, 32, max(100, args.steps // 8)),
                               (2048, 64, max(50, args.steps // 16)), (4096, 32, max(20, args.steps // 32))):
            r = run_workload(af, n, prec, steps, min(args.warmup, 20))
            nm, a, ms, ach = roofline(af, r)
            extra[f"{n}_fp{prec}"] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "dominant": nm,
                                      "dominant_ms": ms, "achieved_GBs": ach,
                                      "frac": ach / HBM_PEAK_GBS,
                                      "step_model_frac": step_model_bytes(n, prec)
                                      / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
            if r["fp32_error"]:
                extra[f"{n}_fp{prec}"]["fp32_error"] = r["fp32_error"]
    # The headline run's creation (graph capture and upload) and state upload
    # leave the device idle for some milliseconds: the separate warm-up run
    # steps again (untimed) right before the headline's own W warm-up steps.
    head = run_workload(af, N, 32, args.steps, args.warmup, host_io=True,
                        device_warmup=lambda: warm.step(200))
    warm.close()
    name, alg, kms, achieved = roofline(af, head)
    traffic = read_traffic(f"{N}_fp32")
    tr_bytes = None
    if traffic and f"{name}_bytes_per_launch" in traffic:
        tr_bytes = traffic[f"{name}_bytes_per_launch"]
    # both passes of the headline run, each against the HBM peak (the dominant
    # one is the roofline entry itself)
    per_pass = {}
    for k, (b, ms) in pass_bytes_ms(af, head).items():
        gbs = b / (ms * 1e-3) / 1e9
        per_pass[k] = {"achieved": gbs, "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": b,
                       "avg_launch_ms": ms,
                       "traffic": traffic.get(f"{k}_bytes_per_launch") if traffic else None}
    if not args.no_extra and N == 1024:
        extra["fp32_over_fp64_1024"] = head["value"] / extra["1024_fp64"]["value"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"single run {N}x{N}, FP32 fields/FFTs, FP64 IF factors + reductions, "
                               f"dt={'1e-4' if N <= 1024 else '5e-5'}, "
                               "lambda0=3.5 alpha=-0.8 beta=0.01 Gamma0=-0.045 Gamma2=2.5e-5, band-limited IC "
                               "(seed 1, amp 1e-3, |k|<=10), per-step energy on",
                   "grid": [N, N], "instances": 1},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": tr_bytes,
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": kms,
                     "events_over_steps": head["prof_steps"],
                     "events_ms_per_step": head["prof_ms_per_step"],
                     "step_model_frac": step_model_bytes(N) / (head["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "passes": per_pass},
        "fp32_error": head["fp32_error"],
        "extra": extra,
    }
    io = head["host_io"]
    if io:
        # PCIe-inclusive rates (never `value`): one upload + one download around
        # the timed steps, and a download every 500 steps (the served snapshot cadence)
        step_s = head["ms_per_step"] * 1e-3
        io["with_one_upload_and_download"] = N * N * args.steps / (
            args.steps * step_s + (io["get_state_ms"] + io["set_state_ms"]) * 1e-3)
        io["with_download_every_500_steps"] = N * N * 500 / (500 * step_s + io["get_state_ms"] * 1e-3)
        out["host_io"] = io
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(af)
        out["batched"] = {}
        for n, cnt, steps in ((1024, 8, 400), (2048, 4, 100), (2048, 8, 60)):
            b = batched_leg(af, n, cnt, steps)
            fp64 = extra.get(f"{n}_fp64")
            if fp64:
                b["over_fp64_single_run"] = b["value"] / fp64["value"]
            out["batched"][f"{n}x{cnt}"] = b
    if not args.no_serving:
        out["serving"] = [serving_leg(af, r) for r in (10, 50, 200)]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
