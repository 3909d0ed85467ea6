#!/usr/bin/env python3
"""Benchmark: linear SWE RK4 steps on one MI355X.

Workload (the single large solve): N=384 cells per panel edge (6*384^2 =
884,736 cells), Ch staggering, sbp42, a=6.371e6 m, g=9.81, H0=1000 m,
Omega=7.292e-5, f = 2 Omega sin(phi); h0 = seeded degree-32 spherical-harmonic
field (20 m RMS), balanced pole-tapered winds; dt at CFL 0.5.

value = FP32 Mpts*steps/s (Mpts = 6 N^2 / 1e6 cells) over exactly --steps RK4
steps with the state resident in HBM, timed with HIP events on the plan's
stream (and wall clock around it, both bracketed by device synchronisation).
The same steps are run in FP64; the line reports both rates, the FP32-vs-FP64
relative L2 error of (h, u, v) after warmup+steps, and the energy drift.
cpu_baseline: the CPU oracle (scalar C, FP64, one core) on a bounded sample of
the same workload.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_DEFAULT, ORDER = 384, 42
HBM_PEAK_GBS = 8000.0


def dof(N):
    return 6 * (N + 1) * (3 * N + 1)


class HipEvents:
    """hipEvent timing on an explicit stream (the plan's)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def run_gpu(cs, precision, N, state0, dt, steps, warmup, ev):
    op = cs.create_linear_swm_operator(N, ORDER, precision)
    st = op.create_stvec().upload(*state0)
    rk4 = cs.create_timescheme("rk4")
    m0, e0 = op.integrals(st)
    if warmup:
        rk4.step(st, op, dt, warmup)
    op.synchronize()
    stream = op.stream
    a, b = ev.create(), ev.create()
    t0 = time.perf_counter()
    ev.record(a, stream)
    rk4.step(st, op, dt, steps)
    ev.record(b, stream)
    op.synchronize()
    wall = time.perf_counter() - t0
    ms = ev.elapsed_ms(a, b)
    m1, e1 = op.integrals(st)
    return dict(ms=ms, wall=wall, state=st.download(), energy_drift=e1 / e0 - 1.0)


def rel_l2(a, b, w):
    return float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b * b)))


def cpu_baseline(N, state0, dt, nsteps):
    from oracle.swe_oracle import Oracle

    o = Oracle(N, ORDER)
    t = time.perf_counter()
    o.rk4(*state0, dt, nsteps)
    sec = time.perf_counter() - t
    return 6 * N * N * nsteps / sec / 1e6, sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=N_DEFAULT)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from ch_swe_amd import testcases as tc

    N = args.N
    t0 = time.perf_counter()
    state0 = tc.sh_state(N, ORDER, lmax=32, seed=0, rms=20.0)  # before any device work
    gen_s = time.perf_counter() - t0
    dt = tc.cfl_dt(N, 0.5)
    print(f"# initial state N={N} generated in {gen_s:.1f}s, dt={dt:.2f}s", file=sys.stderr, flush=True)

    import ch_swe_amd as cs

    ev = HipEvents()
    r32 = run_gpu(cs, cs.FP32, N, state0, dt, args.steps, args.warmup, ev)
    print(f"# fp32: {r32['ms']:.3f} ms for {args.steps} steps", file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        r64 = run_gpu(cs, cs.FP64, N, state0, dt, args.steps, args.warmup, ev)
        print(f"# fp64: {r64['ms']:.3f} ms for {args.steps} steps", file=sys.stderr, flush=True)

    cells = 6 * N * N
    ms_step = r32["ms"] / args.steps
    value = cells * args.steps / (r32["ms"] * 1e-3) / 1e6
    step_bytes = 16 * dof(N) * 4
    achieved = step_bytes / (ms_step * 1e-3) / 1e9

    errs = None
    if r64 is not None:
        w = {m: tc.sbp_weights(N, m, ORDER) for m in ("u", "v", "h")}
        errs = {m: rel_l2(a, b, w[m]) for m, a, b in zip(("u", "v", "h"), r32["state"], r64["state"])}

    cpu_val, cpu_sec = cpu_baseline(N, state0, dt, args.cpu_steps)

    line = {
        "metric": "Mpts*steps/s (FP32 stencils, FP64 metric and integrals), N=384 sbp42 single solve",
        "value": round(value, 2),
        "unit": "Mpts*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "wall_ms_per_step": round(r32["wall"] * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded degree-32 spherical-harmonic h0, 20 m RMS, balanced winds)",
        "config": {"workload": f"single large solve: N={N} (6x{N}^2 cells), Ch, sbp42, RK4 at CFL 0.5",
                   "N": N, "order": ORDER, "dt_s": round(dt, 4), "cells": cells, "dof": dof(N)},
        "fp64": None if r64 is None else {
            "value": round(cells * args.steps / (r64["ms"] * 1e-3) / 1e6, 2),
            "ms_per_step": round(r64["ms"] / args.steps, 5)},
        "accuracy": None if errs is None else {
            "rel_l2_fp32_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
            "after_steps": args.warmup + args.steps,
            "energy_drift_fp32": float(f"{r32['energy_drift']:.3e}"),
            "energy_drift_fp64": float(f"{r64['energy_drift']:.3e}")},
        "roofline": {
            "bound": "hbm",
            "kernel": "one RK4 step (all kernels of the step; unit = 16 state vectors)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": {
            "value": round(cpu_val, 3),
            "unit": "Mpts*steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"CPU oracle (scalar C, FP64), same N={N} sbp42 state, {args.cpu_steps} RK4 steps "
                      f"({cpu_sec:.1f} s)",
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
