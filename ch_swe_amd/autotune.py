"""Autotuner of the fused stage kernel's tiling.

For one (N, order, precision) it times the RK4 step -- the stage and records
launches, each carrying dispatch-timestamped events -- for every first-tile
width a plan accepts along x and y (swe_config.tile_x0 / tile_y0: how the
points that the full 32 x TY tiles leave are split between the two edge
tiles of an axis) and for both block orders (XCD-aware and plain), keeps the
fastest, and stores it in a JSON table that create_linear_swm_operator reads
(ch_swe_amd/tuning.json unless CH_SWE_TUNING names another file).  The
choices change results only at the level of rounding (a point may move
between the inner and the edge path of the kernel, which round
differently).

    python -m ch_swe_amd.autotune 96 384 --order 42 --precision fp32 [--out FILE]

prints one line per candidate and writes the table.
"""
import argparse
import json
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(_HERE, "tuning.json")
F_TX = 32


def table_path():
    return os.environ.get("CH_SWE_TUNING", DEFAULT_TABLE)


def key(N, order, precision):
    from . import FP32
    return f"{int(N)}/{int(order)}/{'fp32' if precision == FP32 else 'fp64'}"


def load_table(path=None):
    path = path or table_path()
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(N, order, precision, path=None):
    """The tuned entry for (N, order, precision), or None."""
    return load_table(path).get(key(N, order, precision))


def axis_choices(N, tw):
    """First-tile widths an axis of N + 1 points can take (0: the library's
    split); the library refuses those whose tiles exceed its limits."""
    n = N + 1
    k = (n - 17) // tw
    if n < 17 + tw or k < 1:
        return [0]
    m = n - k * tw
    return [0] + list(range(8, m - 8))


def time_step(cs, N, order, precision, state, steps, reps, **kw):
    """Mean microseconds of one RK4 step's launches (stage + records), the
    best of reps runs of steps steps."""
    op = cs.create_linear_swm_operator(N, order, precision, tuned=False, **kw)
    st = op.create_stvec().upload(*state)
    rk4 = cs.create_timescheme("rk4")
    rk4.step(st, op, 1.0, 3)
    best = None
    for _ in range(reps):
        op.kernel_timing(True)
        rk4.step(st, op, 1.0, steps)
        sm, ns, rm, nr = op.kernel_times()
        op.kernel_timing(False)
        us = (sm + rm) * 1e3 / steps
        best = us if best is None else min(best, us)
    return best


def tune(N, order=42, precision=None, steps=40, reps=3, log=print):
    """Sweep the candidates for (N, order, precision); returns the table entry."""
    import ch_swe_amd as cs
    if precision is None:
        precision = cs.FP32
    ty = 16 if precision == cs.FP32 else 8
    rng = np.random.default_rng(0)
    state = (1e-3 * rng.standard_normal((6, N + 1, N)), 1e-3 * rng.standard_normal((6, N, N + 1)),
             rng.standard_normal((6, N + 1, N + 1)))
    results = []
    for x0 in axis_choices(N, F_TX):
        for y0 in axis_choices(N, ty):
            for linear in (False, True):
                try:
                    us = time_step(cs, N, order, precision, state, steps, reps, tile_x0=x0, tile_y0=y0,
                                   linear_blocks=linear)
                except cs.SweError as e:
                    if "exceeds" in str(e) or "out of range" in str(e):
                        continue  # a tiling outside the kernel's limits
                    raise
                results.append((us, x0, y0, linear))
                log(f"N={N} order={order} {key(N, order, precision).split('/')[2]} tile_x0={x0} tile_y0={y0} "
                    f"linear_blocks={linear}: {us:.2f} us/step")
    default = next(r for r in results if r[1] == 0 and r[2] == 0 and not r[3])
    best = min(results)
    return {"tile_x0": best[1], "tile_y0": best[2], "linear_blocks": best[3],
            "us_per_step": round(best[0], 3), "default_us_per_step": round(default[0], 3),
            "candidates": len(results), "steps": steps, "reps": reps}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("N", type=int, nargs="+")
    ap.add_argument("--order", type=int, default=42)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None, help="table to update (default: the one the package reads)")
    args = ap.parse_args(argv)
    import ch_swe_amd as cs
    prec = cs.FP32 if args.precision == "fp32" else cs.FP64
    path = args.out or table_path()
    table = load_table(path)
    for N in args.N:
        entry = tune(N, args.order, prec, args.steps, args.reps, log=lambda s: print(s, flush=True))
        table[key(N, args.order, prec)] = entry
        print(f"# {key(N, args.order, prec)}: {json.dumps(entry)}", flush=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"# wrote {path}", file=sys.stderr)


if __name__ == "__main__":
    main()
