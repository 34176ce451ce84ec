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
    args = ap.parse_args()
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
