"""GPU: swe_integrals_noalloc (partial sums in the library's static device
array) against swe_integrals (the plan's buffers), called through the C ABI.
Two calls in a row per configuration, the second on a smaller plan, both
entry points on the same state; the results agree bitwise (same kernels,
same grid, same summation order) and match the oracle's FP64 integrals."""
import ctypes
import sys

import numpy as np
import pytest

import ch_swe_amd as cs
from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(op, st):
    """(swe_integrals, swe_integrals_noalloc) on one state, through the C ABI."""
    L = cs.lib()
    out = []
    for fn in (L.swe_integrals, L.swe_integrals_noalloc):
        m, e = ctypes.c_double(), ctypes.c_double()
        cs._check(fn(op.plan.h, st.h, ctypes.byref(m), ctypes.byref(e)))
        out.append((m.value, e.value))
    return out


# (order, precision, unfused, N of the first call, N of the second, smaller)
CONFIGS = [(42, cs.FP64, False, 97, 20), (21, cs.FP32, False, 64, 9), (42, cs.FP64, True, 24, 12)]


@pytest.mark.parametrize("order,precision,unfused,N1,N2", CONFIGS)
def test_noalloc_integrals_match_allocating_entry(oracle_mod, order, precision, unfused, N1, N2):
    for k, N in enumerate((N1, N2)):
        _say(f"integrals_noalloc order={order} precision={precision} unfused={unfused} N={N} (call {k + 1})")
        o = oracle_mod.Oracle(N, order)
        rng = np.random.default_rng(100 * N + order)
        u, v, h = (rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V)),
                   rng.standard_normal(o.shape(MESH_XY)))
        op = cs.create_linear_swm_operator(N, order, precision, unfused=unfused)
        st = op.create_stvec().upload(u, v, h)
        (m, e), (mn, en) = _both(op, st)
        assert (mn, en) == (m, e)
        bar = 1e-12 if precision == cs.FP64 else 1e-6
        assert abs(mn - o.mass(h)) < bar * np.abs(h).max() * 4 * np.pi
        assert abs(en / o.energy(u, v, h) - 1) < bar


def test_noalloc_rejects_bad_arguments():
    _say("integrals_noalloc errors")
    op = cs.create_linear_swm_operator(8, 42, cs.FP64)
    other = cs.create_linear_swm_operator(8, 42, cs.FP64)
    st = other.create_stvec()
    m, e = ctypes.c_double(), ctypes.c_double()
    L = cs.lib()
    with pytest.raises(cs.SweError, match="another plan"):
        cs._check(L.swe_integrals_noalloc(op.plan.h, st.h, ctypes.byref(m), ctypes.byref(e)))
    with pytest.raises(cs.SweError, match="NULL plan"):
        cs._check(L.swe_integrals_noalloc(None, st.h, ctypes.byref(m), ctypes.byref(e)))
