"""GPU: the serving loop (swe_server_*).  Requests of mixed size, order and
precision, admitted at different ticks under a memory budget, each give
exactly what the same solve gives alone."""
import sys

import numpy as np
import pytest

import ch_swe_amd as cs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _state(N, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((6, N + 1, N)), rng.standard_normal((6, N, N + 1)),
            rng.standard_normal((6, N + 1, N + 1)))


def _alone(N, order, precision, s0, dt, nsteps):
    op = cs.create_linear_swm_operator(N, order, precision)
    st = op.create_stvec().upload(*s0)
    cs.create_timescheme("rk4").step(st, op, dt, nsteps)
    return st.download()


# (N, order, precision, dt, nsteps, tick at which it is submitted)
REQS = [(16, 42, cs.FP64, 900.0, 23, 0), (12, 21, cs.FP32, 1200.0, 9, 0), (24, 42, cs.FP32, 700.0, 40, 1),
        (16, 42, cs.FP64, 600.0, 0, 1), (20, 21, cs.FP64, 800.0, 17, 3), (16, 42, cs.FP32, 500.0, 31, 4)]


@pytest.mark.parametrize("budget", [1 << 34, 4 << 20])
def test_mixed_requests_match_alone(oracle_mod, budget):
    """Ample budget, and one so small (a few MB) that requests wait in the
    queue and idle plans are released."""
    _say(f"server budget={budget}")
    srv = cs.Server(budget)
    ids, s0, results = {}, {}, {}
    tick, saw_queue = 0, False
    while len(results) < len(REQS):
        for k, (N, order, prec, dt, ns, at) in enumerate(REQS):
            if at == tick:
                s0[k] = _state(N, 50 + k)
                ids[k] = srv.submit(*s0[k], dt, ns, N, order, prec)
        srv.tick(7)
        saw_queue |= srv.stats()[0] > 0
        for k, rid in ids.items():
            if k not in results:
                res = srv.take(rid)
                if res is not None:
                    results[k] = res
        tick += 1
        assert tick < 500
    for k, (N, order, prec, dt, ns, _) in enumerate(REQS):
        alone = _alone(N, order, prec, s0[k], dt, ns)
        for g, r in zip(results[k], alone):
            assert _rel(g, r) == 0.0, (k, N, order, prec)
        if prec == cs.FP64 and ns:
            for g, r in zip(results[k], oracle_mod.Oracle(N, order).rk4(*s0[k], dt, ns)):
                assert _rel(g, r) < 1e-12
    if budget < (1 << 30):
        assert saw_queue
    assert srv.stats()[:2] == (0, 0)


def test_server_rejects_bad_requests():
    _say("server errors")
    srv = cs.Server(1 << 20)
    s = _state(16, 1)
    with pytest.raises(cs.SweError, match="budget"):
        srv.submit(*s, 600.0, 3, 16, 42, cs.FP64)
    big = cs.Server(1 << 32)
    with pytest.raises(cs.SweError, match="order"):
        big.submit(*s, 600.0, 3, 16, 63, cs.FP64)
    with pytest.raises(cs.SweError, match="unknown request"):
        cs.lib()
        import ctypes

        r = ctypes.c_int()
        cs._check(cs.lib().swe_server_take(big.h, 12345, None, None, None, ctypes.byref(r)))
