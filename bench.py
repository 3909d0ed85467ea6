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
cpu_baseline: the CPU oracle (C, FP64, OpenMP row loops on the host cores) on
a bounded sample of the same workload.

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
    out = dict(ms=ms, wall=wall, state=st.download(), energy_drift=e1 / e0 - 1.0)
    # per-kernel durations from a separate pass in which every stage /
    # records launch carries dispatch-timestamped start and stop events
    scratch = op.create_stvec().upload(*out["state"])
    op.kernel_timing(True)
    rk4.step(scratch, op, dt, min(steps, 200))
    out["stage_ms"], out["n_stage"], out["records_ms"], out["n_records"] = op.kernel_times()
    op.kernel_timing(False)
    # host-boundary rate: upload of the FP64 host state, the steps, download
    # (what a caller handing over host buffers sees), outside the timed region
    t0 = time.perf_counter()
    st2 = op.create_stvec().upload(*state0)
    rk4.step(st2, op, dt, steps)
    st2.download()
    out["host_boundary_s"] = time.perf_counter() - t0
    return out


TRAFFIC_FILE = "stage_traffic_N384_sbp42_fp32.json"


def measured_traffic(N, dtype):
    """HBM bytes per stage launch from the committed counter runs of this
    workload (N=384 sbp42 FP32), or None for any other workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", TRAFFIC_FILE)
    if N != 384 or ORDER != 42 or dtype != "f32" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def rel_l2(a, b, w):
    return float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b * b)))


def cpu_baseline(N, state0, dt, nsteps, threads):
    """The CPU oracle (C, FP64, OpenMP over rows) on nsteps of the same run."""
    from oracle import swe_oracle

    swe_oracle.set_num_threads(threads)
    o = swe_oracle.Oracle(N, ORDER)
    t = time.perf_counter()
    o.rk4(*state0, dt, nsteps)
    sec = time.perf_counter() - t
    return 6 * N * N * nsteps / sec / 1e6, sec, swe_oracle.num_threads()


def host_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS if set, else the cores
    this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def bench_ensemble(args, ev):
    """Uniform batched ensemble: 256 members at N=96, sbp42, FP32 state, all
    members' stages in one launch each, the step loop replayed from captured
    HIP graphs; 500 timed steps."""
    from ch_swe_amd import testcases as tc
    import ch_swe_amd as cs

    N, members, steps = 96, args.members, args.steps
    states = tc.ensemble_states(N, ORDER, members, lmax=getattr(args, "lmax", 16), seed=1)
    dt = tc.cfl_dt(N, 0.5)
    op = cs.create_linear_swm_operator(N, ORDER, cs.FP32)
    mem = [op.create_stvec().upload(*s) for s in states]
    ens = cs.Ensemble(mem)
    ens.rk4(dt, args.warmup, graph=True)
    ens.synchronize()
    a, b = ev.create(), ev.create()
    t0 = time.perf_counter()
    ev.record(a, ens.stream)
    ens.rk4(dt, steps, graph=True)
    ev.record(b, ens.stream)
    ens.synchronize()
    wall = time.perf_counter() - t0
    ms = ev.elapsed_ms(a, b)
    cells = 6 * N * N
    value = members * cells * steps / (ms * 1e-3) / 1e6
    # one member alone, same steps, for the per-member agreement check
    solo = op.create_stvec().upload(*states[0])
    cs.create_timescheme("rk4").step(solo, op, dt, args.warmup + steps)
    d = max(float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-30)) for x, y in zip(mem[0].download(), solo.download()))
    # per-member FP32-vs-FP64 error: an FP64 ensemble of the first members, same steps
    n64 = min(4, members)
    op64 = cs.create_linear_swm_operator(N, ORDER, cs.FP64)
    mem64 = [op64.create_stvec().upload(*states[k]) for k in range(n64)]
    ens64 = cs.Ensemble(mem64)
    ens64.rk4(dt, args.warmup + steps, graph=True)
    ens64.synchronize()
    w = {m: tc.sbp_weights(N, m, ORDER) for m in ("u", "v", "h")}
    member_err = [{m: float(f"{rel_l2(a, b, w[m]):.3e}") for m, a, b in zip("uvh", mem[k].download(), mem64[k].download())}
                  for k in range(n64)]
    return {
        "metric": "aggregate Mpts*steps/s, uniform batched ensemble (FP32 stencils)",
        "value": round(value, 2), "unit": "Mpts*steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(ms / steps, 5), "wall_ms_per_step": round(wall * 1e3 / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded degree-16 spherical-harmonic fields + Gaussian bumps, one per member)",
        "config": {"workload": f"uniform ensemble: {members} members, N={N}, sbp42, RK4 at CFL 0.5, HIP-graph step loop",
                   "members": members, "N": N, "order": ORDER, "dt_s": round(dt, 4)},
        "member0_vs_alone_max_rel_diff": d,
        "member_rel_l2_fp32_vs_fp64": member_err,
        "after_steps": args.warmup + steps,
    }


def bench_uneven(args, ev):
    """Uneven batch: 64 instances, N in {24,48,96,192}, sbp21/sbp42 mixed,
    own CFL dt, steps log-uniform in [100, 2000], FP32; batched run against
    the same instances solved one at a time."""
    from ch_swe_amd import testcases as tc
    import ch_swe_amd as cs

    rng = np.random.default_rng(4)
    n_inst = args.members
    Ns = rng.choice([24, 48, 96, 192], n_inst)
    orders = rng.choice([21, 42], n_inst)
    steps = np.round(np.exp(rng.uniform(np.log(100), np.log(2000), n_inst))).astype(int)
    if args.steps_scale != 1.0:
        steps = np.maximum(1, np.round(steps * args.steps_scale)).astype(int)
    dts = np.array([tc.cfl_dt(int(N), 0.5) for N in Ns])
    plans, init = {}, []
    for i in range(n_inst):
        key = (int(Ns[i]), int(orders[i]))
        if key not in plans:
            plans[key] = cs.create_linear_swm_operator(key[0], key[1], cs.FP32)
        init.append(tc.ensemble_states(key[0], key[1], 1, lmax=getattr(args, "lmax", 16), seed=100 + i, basis=1)[0])
    cells = np.array([6 * int(N) * int(N) for N in Ns])
    work = float(np.sum(cells * steps))
    # batched
    mem = [plans[(int(Ns[i]), int(orders[i]))].create_stvec().upload(*init[i]) for i in range(n_inst)]
    ens = cs.Ensemble(mem)
    ens.rk4(dts, np.minimum(steps, 5), graph=False)  # warm-up (compiles nothing; touches memory)
    ens.synchronize()
    for i in range(n_inst):
        mem[i].upload(*init[i])
    a, b = ev.create(), ev.create()
    ev.record(a, ens.stream)
    ens.rk4(dts, steps, graph=True)
    ev.record(b, ens.stream)
    ens.synchronize()
    ms_batch = ev.elapsed_ms(a, b)
    # one at a time
    solo = [plans[(int(Ns[i]), int(orders[i]))].create_stvec().upload(*init[i]) for i in range(n_inst)]
    rk = cs.create_timescheme("rk4")
    for p in plans.values():
        p.synchronize()
    t0 = time.perf_counter()
    for i in range(n_inst):
        op = plans[(int(Ns[i]), int(orders[i]))]
        rk.step(solo[i], op, float(dts[i]), int(steps[i]))
        op.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3
    d = max(float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-30))
            for i in range(n_inst) for x, y in zip(mem[i].download(), solo[i].download()))
    # per-instance FP32-vs-FP64 error: FP64 copies of a sample of the instances, batched
    sample = list(range(min(8, n_inst)))
    plans64 = {}
    mem64 = []
    for i in sample:
        key = (int(Ns[i]), int(orders[i]))
        if key not in plans64:
            plans64[key] = cs.create_linear_swm_operator(key[0], key[1], cs.FP64)
        mem64.append(plans64[key].create_stvec().upload(*init[i]))
    ens64 = cs.Ensemble(mem64)
    ens64.rk4(dts[sample], steps[sample], graph=True)
    ens64.synchronize()
    inst_err = []
    for j, i in enumerate(sample):
        N_i, o_i = int(Ns[i]), int(orders[i])
        w = {m: tc.sbp_weights(N_i, m, o_i) for m in ("u", "v", "h")}
        e = {m: float(f"{rel_l2(a, b, w[m]):.3e}") for m, a, b in zip("uvh", mem[i].download(), mem64[j].download())}
        inst_err.append({"N": N_i, "order": o_i, "steps": int(steps[i]), **e})
    return {
        "metric": "aggregate Mpts*steps/s, uneven batch (FP32 stencils)",
        "value": round(work / (ms_batch * 1e-3) / 1e6, 2), "unit": "Mpts*steps/s", "n_gpus": 1,
        "steps": int(steps.max()), "warmup": 5, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded degree-16 spherical-harmonic field + Gaussian bump per instance)",
        "config": {"workload": f"uneven batch: {n_inst} instances, N in {{24,48,96,192}}, sbp21/42, own dt, "
                               "steps log-uniform [100, 2000], HIP-graph step loop",
                   "instances": n_inst, "cells_x_steps": work},
        "batched_ms": round(ms_batch, 3), "one_at_a_time_ms": round(ms_seq, 3),
        "one_at_a_time_value": round(work / (ms_seq * 1e-3) / 1e6, 2),
        "speedup_vs_one_at_a_time": round(ms_seq / ms_batch, 2),
        "batched_vs_alone_max_rel_diff": d,
        "instance_rel_l2_fp32_vs_fp64": inst_err,
    }


def bench_serving(args, ev):
    """Serving: Poisson request streams (rates swept) of solves with N in
    {32,64,128,256,512} (weights 0.3/0.3/0.2/0.15/0.05), steps log-uniform in
    [100, 2000], FP32 or FP64 per request, seeded h0 (scaled degree-16
    spherical-harmonic field with balanced winds + Gaussian bump); admitted
    by a device-memory budget, continuously batched.  Reports p50/p99
    latency (arrival to result) and aggregate Mpts*steps/s per rate."""
    from ch_swe_amd import testcases as tc
    import ch_swe_amd as cs

    sizes, weights = [32, 64, 128, 256, 512], [0.3, 0.3, 0.2, 0.15, 0.05]
    base = {N: tc.sh_state(N, ORDER, lmax=getattr(args, "lmax", 16), seed=7, rms=20.0, workers=1) for N in sizes}
    geo = {N: tc.mesh_geometry(N, "h")[0] for N in sizes}
    rates = [float(r) for r in args.rates.split(",")]
    n_req = args.members or 48
    rows = []
    for ri, lam in enumerate(rates):
        # the request stream, built before the clock starts (client side)
        rng = np.random.default_rng(100 + ri)
        arr = np.cumsum(rng.exponential(1.0 / lam, n_req))
        Ns = rng.choice(sizes, n_req, p=weights)
        steps = np.round(np.exp(rng.uniform(np.log(100), np.log(2000), n_req))).astype(int)
        precs = rng.choice([cs.FP32, cs.FP64], n_req)
        scale = rng.standard_normal(n_req)
        amp, width = rng.uniform(1, 20, n_req), rng.uniform(2e5, 1e6, n_req) / 6.371e6
        cen = rng.standard_normal((n_req, 3))
        cen /= np.linalg.norm(cen, axis=1, keepdims=True)
        states = []
        for k in range(n_req):
            u, v, h = base[int(Ns[k])]
            d = np.arccos(np.clip(geo[int(Ns[k])] @ cen[k], -1.0, 1.0))
            states.append((scale[k] * u, scale[k] * v, scale[k] * h + amp[k] * np.exp(-((d / width[k]) ** 2))))
        srv = cs.Server(int(args.budget_gb * (1 << 30)))
        ids, t_arr, t_done = {}, {}, {}
        nxt, t0 = 0, time.perf_counter()
        while len(t_done) < n_req:
            now = time.perf_counter() - t0
            while nxt < n_req and arr[nxt] <= now:
                N = int(Ns[nxt])
                rid = srv.submit(*states[nxt], tc.cfl_dt(N, 0.5), int(steps[nxt]), N, ORDER, int(precs[nxt]))
                ids[rid] = nxt
                t_arr[rid] = arr[nxt]
                nxt += 1
            q, a, _ = srv.stats()
            if q == 0 and a == 0:
                if nxt < n_req:
                    time.sleep(max(0.0, arr[nxt] - (time.perf_counter() - t0)))
                continue
            srv.tick(args.chunk)
            tnow = time.perf_counter() - t0
            for rid in [r for r in ids if r not in t_done]:
                if srv.take(rid) is not None:
                    t_done[rid] = tnow
        makespan = max(t_done.values()) - arr[0]
        lat = np.array([t_done[r] - t_arr[r] for r in t_done]) * 1e3
        work = float(np.sum([6 * int(N) * int(N) * int(s) for N, s in zip(Ns, steps)]))
        rows.append({"rate_per_s": lam, "requests": n_req, "p50_ms": round(float(np.percentile(lat, 50)), 2),
                     "p99_ms": round(float(np.percentile(lat, 99)), 2),
                     "mpts_steps_per_s": round(work / makespan / 1e6, 2),
                     "fp64_share": round(float(np.mean(precs == cs.FP64)), 2)})
        print(f"# serving rate {lam}/s: {rows[-1]}", file=sys.stderr, flush=True)
        srv.close()
    top = rows[-1]
    return {
        "metric": "p50/p99 latency (ms) and aggregate Mpts*steps/s, served mixed-size solves",
        "value": top["mpts_steps_per_s"], "unit": "Mpts*steps/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic Poisson request stream (seeded sizes, steps, precisions, initial states)",
        "config": {"workload": f"serving: Poisson arrivals at {args.rates} req/s, {n_req} requests per rate, "
                               f"budget {args.budget_gb} GB, {args.chunk}-step ticks",
                   "sizes": sizes, "weights": weights},
        "per_rate": rows,
    }


def bench_implicit(args):
    """The implicit solve (swe_solve: CG on the Helmholtz problem for h,
    operator_swm_lin_mod.f90 L110-146) at CFL 4 on a seeded random
    right-hand side, rel_tol 1e-8: microseconds per CG iteration for N = 96
    and N = args.N, FP64 and FP32 (wall time of one solve after a warm-up
    solve, the solve ends synchronised)."""
    import ch_swe_amd as cs
    rows = []
    for N, prec in ((96, cs.FP64), (args.N, cs.FP64), (args.N, cs.FP32)):
        op = cs.create_linear_swm_operator(N, ORDER, prec, radius=6371229.0)
        rng = np.random.default_rng(0)
        r = op.create_stvec().upload(rng.standard_normal((6, N + 1, N)), rng.standard_normal((6, N, N + 1)),
                                     rng.standard_normal((6, N + 1, N + 1)))
        x = op.create_stvec()
        dt = 4 * (0.5 * np.pi / N) * 6371229.0 / np.sqrt(9.81 * 1000.0)
        op.solve(x, r, dt, rel_tol=1e-8, max_iter=300)
        op.synchronize()
        t0 = time.perf_counter()
        it, rel = op.solve(x, r, dt, rel_tol=1e-8, max_iter=300)
        op.synchronize()
        t = time.perf_counter() - t0
        rows.append({"N": N, "dtype": "f64" if prec == cs.FP64 else "f32", "cg_iterations": int(it),
                     "rel_residual": float(rel), "us_per_iteration": round(t * 1e6 / max(it, 1), 1)})
        print(f"# implicit solve {rows[-1]}", file=sys.stderr, flush=True)
    top = rows[-1]
    return {
        "metric": "CG iterations per second of the implicit solve (Helmholtz problem, CFL 4, rel_tol 1e-8)",
        "value": round(1e6 / top["us_per_iteration"], 1), "unit": "iterations/s", "n_gpus": 1, "steps": None,
        "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic seeded random right-hand side",
        "config": {"workload": f"implicit solve: N={args.N} sbp42 FP32 (rows: N=96 FP64, N={args.N} FP64 / FP32)"},
        "per_case": rows,
    }


def accuracy_leg(cs, tc, N, state0, dt, steps):
    """FP32 (plain and with compensated summation) against FP64 after
    `steps` RK4 steps from the benchmark state, untimed: relative SBP-L2
    error of u, v, h and the relative drift of the linear energy."""
    w = {m: tc.sbp_weights(N, m, ORDER) for m in ("u", "v", "h")}
    res = {}
    for name, prec, comp in (("fp64", cs.FP64, False), ("fp32", cs.FP32, False), ("fp32_compensated", cs.FP32, True)):
        op = cs.create_linear_swm_operator(N, ORDER, prec, compensated=comp)
        st = op.create_stvec().upload(*state0)
        _, e0 = op.integrals(st)
        cs.create_timescheme("rk4").step(st, op, dt, steps)
        _, e1 = op.integrals(st)
        res[name] = (st.download(), e1 / e0 - 1.0)
    out = {"after_steps": steps, "energy_drift_fp64": float(f"{res['fp64'][1]:.3e}")}
    for name in ("fp32", "fp32_compensated"):
        out["rel_l2_" + name + "_vs_fp64"] = {m: float(f"{rel_l2(a, b, w[m]):.3e}")
                                             for m, a, b in zip("uvh", res[name][0], res["fp64"][0])}
        out["energy_drift_" + name] = float(f"{res[name][1]:.3e}")
    return out


def oracle_check(cs, N, state0, dt, steps, threads):
    """The GPU path against the CPU oracle on the benchmark state, `steps`
    RK4 steps: FP64 max relative error (bar 1e-12) and FP32 relative L2."""
    from ch_swe_amd import testcases as tc
    from oracle import swe_oracle

    swe_oracle.set_num_threads(threads)
    ref = swe_oracle.Oracle(N, ORDER).rk4(*state0, dt, steps)
    w = {m: tc.sbp_weights(N, m, ORDER) for m in ("u", "v", "h")}
    out = {"steps": steps}
    for name, prec in (("fp64", cs.FP64), ("fp32", cs.FP32)):
        op = cs.create_linear_swm_operator(N, ORDER, prec)
        st = op.create_stvec().upload(*state0)
        cs.create_timescheme("rk4").step(st, op, dt, steps)
        got = st.download()
        if prec == cs.FP64:
            out["fp64_max_rel_err"] = float(f"{max(np.abs(g - r).max() / np.abs(r).max() for g, r in zip(got, ref)):.3e}")
        else:
            out["fp32_rel_l2"] = {m: float(f"{rel_l2(g, r, w[m]):.3e}") for m, g, r in zip("uvh", got, ref)}
    return out


def sample_workloads(args, ev):
    """Small versions of the batch and serving workloads, run after the
    headline measurement: the uniform ensemble (256 x N=96, 100 timed
    steps), the uneven batch (64 instances, step counts scaled by 1/4) and
    the serving loop at 10, 100 and 500 requests/s (24 requests each).  The
    host-side initial states use spherical harmonics of degree 6 instead of
    16 (generation time; the solves' cost does not depend on it)."""
    import copy

    out = {}
    a = copy.copy(args)
    a.members, a.steps, a.warmup, a.lmax = 256, 100, 10, 6
    e = bench_ensemble(a, ev)
    out["ensemble"] = {k: e[k] for k in ("value", "unit", "ms_per_step", "member_rel_l2_fp32_vs_fp64", "after_steps",
                                         "member0_vs_alone_max_rel_diff")}
    out["ensemble"]["config"] = e["config"]["workload"]
    print(f"# ensemble sample: {e['value']} Mpts*steps/s", file=sys.stderr, flush=True)
    a = copy.copy(args)
    a.members, a.steps_scale, a.lmax = 64, 0.25, 6
    u = bench_uneven(a, ev)
    out["uneven"] = {k: u[k] for k in ("value", "unit", "one_at_a_time_value", "speedup_vs_one_at_a_time",
                                       "batched_vs_alone_max_rel_diff", "instance_rel_l2_fp32_vs_fp64")}
    out["uneven"]["config"] = u["config"]["workload"] + " (step counts x 0.25)"
    print(f"# uneven sample: {u['value']} Mpts*steps/s", file=sys.stderr, flush=True)
    a = copy.copy(args)
    a.members, a.rates, a.lmax = 24, "10,100,500", 6
    v = bench_serving(a, ev)
    out["serving"] = {"per_rate": v["per_rate"], "config": v["config"]["workload"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=N_DEFAULT)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-threads", type=int, default=host_threads())
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", choices=["single", "ensemble", "uneven", "serving", "implicit"], default="single",
                    help="single large solve (default, the headline), uniform ensemble, uneven batch, serving, "
                         "implicit solve")
    ap.add_argument("--rates", default="5,10,20,40", help="serving: request rates (per second)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serving: device-memory budget")
    ap.add_argument("--chunk", type=int, default=25, help="serving: steps per tick")
    ap.add_argument("--members", type=int, default=None, help="ensemble members / batch instances")
    ap.add_argument("--steps-scale", type=float, default=1.0, help="uneven batch: scale every step count")
    ap.add_argument("--no-samples", action="store_true",
                    help="single workload: skip the accuracy, oracle and batch/serving sample legs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (3 launches) so that a "
                         "counter profile can scale FETCH_SIZE / WRITE_SIZE by a known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload != "single":
        ev = HipEvents()
        if args.workload == "ensemble":
            args.members = args.members or 256
            if args.steps == 1000:
                args.steps = 500
            line = bench_ensemble(args, ev)
        elif args.workload == "serving":
            line = bench_serving(args, ev)
        elif args.workload == "implicit":
            line = bench_implicit(args)
        else:
            args.members = args.members or 64
            line = bench_uneven(args, ev)
        print(json.dumps(line), flush=True)
        return

    from ch_swe_amd import testcases as tc

    N = args.N
    t0 = time.perf_counter()
    state0 = tc.sh_state(N, ORDER, lmax=32, seed=0, rms=20.0)  # before any device work
    gen_s = time.perf_counter() - t0
    dt = tc.cfl_dt(N, 0.5)
    print(f"# initial state N={N} generated in {gen_s:.1f}s, dt={dt:.2f}s", file=sys.stderr, flush=True)

    import ch_swe_amd as cs

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy calibration: 3 x {args.copy_calibration} MiB read + write", file=sys.stderr,
              flush=True)
        del src, dst

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
    # Dominant kernel: the fused stage kernel, 4 launches per step.  The
    # algorithmic traffic of classic RK4 done stage by stage is 3, 5, 5, 3
    # state vectors (stage 1 reads y writes acc, y2; stages 2-3 read y, v0,
    # acc and write acc, y'; stage 4 reads y, acc and writes v0): 4 vectors
    # per launch on average, one vector = DOF * 4 bytes.  (The kernels move
    # 17 vectors a step: the last stage also reads v0, because acc holds only
    # the increment.)  The timed region also holds the edge-record launches
    # (one after each stage).  Durations come from the pass with
    # dispatch-timestamped events (the kernels' own run times).
    launch_bytes = 4 * dof(N) * 4
    launch_us = r32["stage_ms"] * 1e3 / r32["n_stage"]
    records_us = r32["records_ms"] * 1e3 / max(r32["n_records"], 1)
    achieved = launch_bytes / (launch_us * 1e-6) / 1e9
    traffic = measured_traffic(N, "f32")

    errs = None
    if r64 is not None:
        w = {m: tc.sbp_weights(N, m, ORDER) for m in ("u", "v", "h")}
        errs = {m: rel_l2(a, b, w[m]) for m, a, b in zip(("u", "v", "h"), r32["state"], r64["state"])}

    cpu_val, cpu_sec, cpu_threads = (cpu_baseline(N, state0, dt, args.cpu_steps, args.cpu_threads)
                                     if args.cpu_steps > 0 else (None, 0.0, 0))
    # the same oracle on one core, a tenth of the steps
    one_steps = max(1, args.cpu_steps // 10)
    cpu1_val, cpu1_sec, _ = (cpu_baseline(N, state0, dt, one_steps, 1) if args.cpu_steps > 0 else (None, 0.0, 0))

    extra = {}
    if not args.no_samples:
        extra["accuracy_1000_steps"] = accuracy_leg(cs, tc, N, state0, dt, 1000)
        print(f"# accuracy leg: {extra['accuracy_1000_steps']}", file=sys.stderr, flush=True)
        extra["oracle_check"] = oracle_check(cs, N, state0, dt, 20, args.cpu_threads)
        print(f"# oracle check: {extra['oracle_check']}", file=sys.stderr, flush=True)
        extra.update(sample_workloads(args, ev))

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
        "host_boundary_value": round(cells * args.steps / r32["host_boundary_s"] / 1e6, 2),
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
            "kernel": "k_stage (fused RK4 stage), mean of the 4 stage launches per step",
            "bytes_per_launch": launch_bytes,
            "us_per_launch": round(launch_us, 3),
            "records_us_per_launch": round(records_us, 3),
            "step_fraction_in_stage_kernel": round(4 * launch_us * 1e-3 / ms_step, 3),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else traffic["hbm_bytes_per_launch_mean"],
            "traffic_source": None if traffic is None else
            "rocprofv3 TCC_EA0 request counts by size (32/64/128 B reads, 32/64 B writes; a 512 MiB copy "
            "counts 512 MiB each way), mean of the stage launches of a step, profiles/" + TRAFFIC_FILE,
        },
        "cpu_baseline": {
            "value": None if cpu_val is None else round(cpu_val, 3),
            "unit": "Mpts*steps/s",
            "cores": cpu_threads,
            "kind": "port",
            "sample": f"CPU oracle (C, FP64, OpenMP row loops, {cpu_threads} threads), same N={N} "
                      f"sbp42 state, {args.cpu_steps} RK4 steps ({cpu_sec:.1f} s)",
            "single_core_value": None if cpu1_val is None else round(cpu1_val, 3),
            "single_core_sample": f"the same oracle on 1 thread, {one_steps} RK4 steps ({cpu1_sec:.1f} s)",
        },
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
