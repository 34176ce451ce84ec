#!/usr/bin/env python3
"""Autotuner for the FFT radix / brick schedules of libactflow.so.

For every requested (precision, N) it times each instantiated schedule on the
GPU -- a single run, or an ensemble of `--instances` copies stepped in shared
launches -- checks that every schedule reproduces the default schedule's state
to rounding, and writes the fastest into a table beside the library that new
processes pick up automatically: schedules.txt (single runs) by default, or
with --instances > 1 schedules_batched.txt (runs of the ensemble engine and
the serving loop).

    python3 2d_active_matter_simulation_amd/autotune.py --sizes 256 512 1024 2048 4096
    python3 2d_active_matter_simulation_amd/autotune.py --instances 8 --precisions 32 --sizes 256 512 1024 2048

With --mix N:COUNT ... (FP32) the engine table is tuned for one mix of size
classes stepped together, where the classes run concurrently and a schedule
that is fastest alone need not be fastest in the mix: coordinate descent over
each class's schedules with the other classes fixed, timing the whole mix in
one engine, starting from the current table; the result is then timed against
the starting table in alternating repetitions.

    python3 2d_active_matter_simulation_amd/autotune.py --mix 256:32 512:24 1024:8
"""
import argparse
import importlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
af = importlib.import_module(os.path.basename(HERE))

PHYS = dict(lam=3.5, alpha=-0.8, beta=0.01, gamma0=-0.045, gamma2=2.5e-5)


def _rel(a, b):
    wt = np.full(a.shape, 2.0)
    wt[:, 0] = 1
    wt[:, -1] = 1
    return float(np.sqrt(np.sum(wt * np.abs(a - b) ** 2) / np.sum(wt * np.abs(b) ** 2)))


def time_schedule(prec, n, idx, steps, instances, check_steps=4):
    af.select_schedule(prec, n, n, idx)
    p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, precision=prec, **PHYS)
    w0 = af.ic_bandlimited(n, seed=1, amp=1.0, kmax=10.0)
    if instances <= 1:
        run = af.Run(p)
        run.set_state(w0)
        run.step(check_steps)
        state = run.get_state()[0]
        run.step(5)
        t0 = time.perf_counter()
        run.step(steps)
        dt = time.perf_counter() - t0
        run.close()
    else:
        eng = af.Engine(1 << 40)
        slots = [eng.add(p, w0) for _ in range(instances)]
        eng.step(check_steps)
        state = eng.get_state(slots[0])[0]
        eng.step(5)
        t0 = time.perf_counter()
        eng.step(steps)
        dt = time.perf_counter() - t0
        eng.close()
    return dt / steps * 1e3, state


def _table_index(path, prec, n, scheds):
    """Index in scheds of the schedule a table names for n x n, or 0."""
    if os.path.exists(path):
        for line in open(path):
            v = line.split("#")[0].split()
            if len(v) >= 10 and (int(v[0]), int(v[1]), int(v[2])) == (prec, n, n):
                sc = tuple(int(x) for x in v[3:])
                if len(sc) == 7:
                    sc += (sc[1],)
                if len(sc) == 8:
                    sc += (0,)
                if sc in scheds:
                    return scheds.index(sc)
    return 0


def time_mix(mix, choice, steps, seed=7):
    """ms per step of the whole mix [(n, count)] in one engine, every class n
    with schedule index choice[n] (FP32, seeded parameters as the ensemble
    benchmark: alpha in [-1.2, -0.4], beta in [0.005, 0.02], lambda0 in [1, 5])."""
    for n, i in choice.items():
        af.select_schedule(32, n, n, i)
    rng = np.random.default_rng(seed)
    eng = af.Engine(1 << 40)
    try:
        for n, count in mix:
            for _ in range(count):
                p = af.params(n, dt=1e-4 if n <= 1024 else 5e-5, precision=32, lam=float(rng.uniform(1, 5)),
                              alpha=float(rng.uniform(-1.2, -0.4)), beta=float(rng.uniform(0.005, 0.02)),
                              gamma0=-0.045, gamma2=2.5e-5)
                eng.add(p, af.ic_bandlimited(n, seed=int(rng.integers(1, 1 << 30)), amp=1e-3, kmax=10.0))
        eng.step(20)
        t0 = time.perf_counter()
        eng.step(steps)
        return (time.perf_counter() - t0) / steps * 1e3
    finally:
        eng.close()
        for n in choice:
            af.select_schedule(32, n, n, -1)


def tune_mix(mix, table, steps, passes=2, confirm=3):
    sizes = [n for n, _ in mix]
    scheds = {n: af.schedules(32, n) for n in sizes}
    start = {n: _table_index(table, 32, n, scheds[n]) for n in sizes}
    choice = dict(start)
    best = time_mix(mix, choice, steps)
    print(f"mix {mix}: start {[scheds[n][choice[n]] for n in sizes]} {best:.4f} ms/step", flush=True)
    for _ in range(passes):
        improved = False
        for n in sizes:
            for i in range(len(scheds[n])):
                if i == choice[n]:
                    continue
                ms = time_mix(mix, {**choice, n: i}, steps)
                print(f"  {n:5d} sched {i} {scheds[n][i]} {ms:8.4f} ms/step", flush=True)
                if ms < best * 0.995:
                    best, choice[n], improved = ms, i, True
                    print(f"  -> {n}: {scheds[n][i]} {ms:.4f}", flush=True)
        if not improved:
            break
    # the result against the starting table, alternating, on longer runs
    a, b = [], []
    for _ in range(confirm):
        a.append(time_mix(mix, start, 4 * steps))
        b.append(time_mix(mix, choice, 4 * steps))
    print(f"confirm: start {min(a):.4f} ms/step, tuned {min(b):.4f} ms/step", flush=True)
    return (choice, min(b)) if min(b) < min(a) * 0.995 else (start, min(a)), scheds


def merge_table(lines, existing):
    """Table text from the header + tuned `lines`, keeping the lines of the
    existing table text for (precision, nx, ny) not tuned now."""
    tuned = {tuple(ln.split()[:3]) for ln in lines[1:]}
    out = list(lines)
    for ln in existing.splitlines():
        if ln.strip() and not ln.startswith("#") and tuple(ln.split()[:3]) not in tuned:
            out.append(ln)
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[256, 512, 1024, 2048, 4096])
    ap.add_argument("--precisions", type=int, nargs="+", default=[32, 64])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--instances", type=int, default=1)
    ap.add_argument("--out", default=None,
                    help="table to write (default: schedules.txt, or schedules_batched.txt with --instances > 1)")
    ap.add_argument("--dry", action="store_true", help="print, do not write the table")
    ap.add_argument("--mix", nargs="+", default=None, metavar="N:COUNT",
                    help="tune the engine table for this mix of FP32 size classes stepped together")
    args = ap.parse_args()
    if args.mix:
        mix = [tuple(int(x) for x in m.split(":")) for m in args.mix]
        out = args.out or os.path.join(HERE, "schedules_batched.txt")
        (choice, ms), scheds = tune_mix(mix, out, args.steps)
        note = "+".join(f"{c}x{n}^2" for n, c in mix)
        lines = ["# prec nx ny by bx row_radix row_group col_radix col_fwd_group col_inv_group col_width options  (ms/step)"]
        lines += [f"32 {n} {n} " + " ".join(map(str, scheds[n][choice[n]])) + f"   # {ms:.4f} (the mix {note})"
                  for n, _ in mix]
        text = merge_table(lines, open(out).read() if os.path.exists(out) else "")
        if args.dry:
            print(text)
        else:
            with open(out, "w") as f:
                f.write(text)
            print("wrote", out)
        return
    if args.out is None:
        args.out = os.path.join(HERE, "schedules_batched.txt" if args.instances > 1 else "schedules.txt")
    lines = ["# prec nx ny by bx row_radix row_group col_radix col_fwd_group col_inv_group col_width options  (ms/step)"]
    for prec in args.precisions:
        for n in args.sizes:
            scheds = af.schedules(prec, n)
            best = None
            ref = None
            for i, sc in enumerate(scheds):
                steps = max(10, args.steps * (1024 * 1024) // (n * n)) if n < 1024 else args.steps
                ms, state = time_schedule(prec, n, i, steps, args.instances)
                err = 0.0 if ref is None else _rel(state, ref)
                if ref is None:
                    ref = state
                tol = 1e-12 if prec == 64 else 1e-5
                ok = err <= tol
                print(f"fp{prec} {n:5d} sched {i} {sc} {ms:8.4f} ms/step  rel-to-default {err:.1e}"
                      f"{'' if ok else '  REJECTED'}", flush=True)
                if ok and (best is None or ms < best[0]):
                    best = (ms, sc)
            if best:
                note = f" ({args.instances} instances)" if args.instances > 1 else ""
                lines.append(f"{prec} {n} {n} " + " ".join(map(str, best[1])) + f"   # {best[0]:.4f}{note}")
                print(f"  -> best {best[1]} {best[0]:.4f} ms/step", flush=True)
    text = merge_table(lines, open(args.out).read() if os.path.exists(args.out) else "")
    if args.dry:
        print(text)
    else:
        with open(args.out, "w") as f:
            f.write(text)
        print("wrote", args.out)


if __name__ == "__main__":
    main()
