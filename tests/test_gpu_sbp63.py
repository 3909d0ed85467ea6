"""GPU: the sbp63 Ch operator (div_ch_sbp_sat_63_proj, grad_ch_sbp_sat_63,
co2contra_ch_sbp63, coriolis_Ch_sbp63 and SBP_C63_quadrature:
div_factory_mod.f90 L185-214, grad_factory_mod.f90 L178, co2contra_factory_mod.f90
L148-152, coriolis_factory_mod.f90 L57-61, quadrature_factory_mod.f90 L118, with
the stencils of sbp_diff_i2c_63_mod.f90, sbp_diff_c2i_63_mod.f90,
sbp_interp_i2c_63_mod.f90 and sbp_interp_c2i_63_mod.f90), which runs on the
multi-kernel path, against the oracle on the same seeded inputs.  FP64 bar:
1e-12 relative (max error / max magnitude of the field); FP32: 1e-5."""
import sys

import numpy as np
import pytest

import ch_swe_amd as cs
from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

pytestmark = pytest.mark.gpu

A = 6371229.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _random_state(o, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V)),
            rng.standard_normal(o.shape(MESH_XY)))


# N = 12: the closures of both ends meet (rows 1..6 and 7..12 of the centres);
# N = 13: one interior row; larger N: interior stencils between them
@pytest.mark.parametrize("N", [12, 13, 20, 33, 48])
@pytest.mark.parametrize("cor_const", [False, True])
def test_apply_fp64_matches_oracle(oracle_mod, N, cor_const):
    _say(f"sbp63 apply fp64 N={N} cor_const={cor_const}")
    kw = dict(coriolis_const=cor_const, omega=1e-4 if cor_const else 7.292e-5)
    o = oracle_mod.Oracle(N, 63, **kw)
    u, v, h = _random_state(o, 630 + N)
    op = cs.create_linear_swm_operator(N, 63, cs.FP64, **kw)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    for g, r in zip(vout.download(), o.apply(u, v, h)):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("N", [14, 24])
def test_rk4_fp64_matches_oracle(oracle_mod, N):
    _say(f"sbp63 rk4 fp64 N={N}")
    o = oracle_mod.Oracle(N, 63)
    u, v, h = _random_state(o, 7)
    u, v = 10 * u, 10 * v
    ref = o.rk4(u, v, h, 300.0, 20)
    op = cs.create_linear_swm_operator(N, 63, cs.FP64)
    st = op.create_stvec().upload(u, v, h)
    cs.create_timescheme("rk4").step(st, op, 300.0, 20)
    for g, r in zip(st.download(), ref):
        assert _rel(g, r) < 1e-12


def test_apply_fp32_close_to_oracle(oracle_mod):
    N = 32
    _say(f"sbp63 apply fp32 N={N}")
    o = oracle_mod.Oracle(N, 63)
    u, v, h = _random_state(o, 11)
    op = cs.create_linear_swm_operator(N, 63, cs.FP32)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    for g, r in zip(vout.download(), o.apply(u, v, h)):
        assert _rel(g, r) < 1e-5


@pytest.mark.parametrize("N", [12, 25])
def test_integrals_match_oracle(oracle_mod, N):
    """Mass and energy with the sbp63 quadrature weights."""
    _say(f"sbp63 integrals N={N}")
    o = oracle_mod.Oracle(N, 63)
    u, v, h = _random_state(o, 21 + N)
    op = cs.create_linear_swm_operator(N, 63, cs.FP64)
    m, e = op.integrals(op.create_stvec().upload(u, v, h))
    assert abs(m - o.mass(h)) < 1e-12 * np.abs(h).max() * 4 * np.pi
    assert abs(e / o.energy(u, v, h) - 1) < 1e-12


def test_solve_matches_oracle(oracle_mod):
    N, dt = 20, 3000.0
    _say(f"sbp63 solve N={N}")
    o = oracle_mod.Oracle(N, 63, radius=A, grav=9.80616, h_mean=1000.0)
    op = cs.create_linear_swm_operator(N, 63, cs.FP64, radius=A, grav=9.80616, h_mean=1000.0)
    rng = np.random.default_rng(63)
    rhs = (rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V)),
           rng.standard_normal(o.shape(MESH_XY)))
    ref = o.solve(*rhs, dt, 1e-12, 500)
    x = op.create_stvec()
    it, rel = op.solve(x, op.create_stvec().upload(*rhs), dt, rel_tol=1e-12, max_iter=500)
    assert rel < 1e-12 and abs(it - ref[3]) <= 2, (it, ref[3], rel)
    for g, e in zip(x.download(), ref[:3]):
        assert _rel(g, e) < 1e-9


def test_energy_is_conserved_without_rotation(oracle_mod):
    """The sbp63 operator is energy neutral for a consistent state (the SBP
    property with the projection SAT): RK4 with omega = 0 changes the energy
    of a smooth hill by no more than the scheme's own dissipation."""
    N = 24
    _say(f"sbp63 energy N={N}")
    o = oracle_mod.Oracle(N, 63, omega=0.0)
    h0 = o.gaussian_hill(0.3, 0.4, 0.2, 10.0)
    op = cs.create_linear_swm_operator(N, 63, cs.FP64, omega=0.0)
    st = op.create_stvec().upload(o.zeros(MESH_U), o.zeros(MESH_V), h0)
    e0 = op.integrals(st)[1]
    cs.create_timescheme("rk4").step(st, op, 300.0, 40)
    assert abs(op.integrals(st)[1] / e0 - 1) < 1e-7


@pytest.mark.parametrize("order", [63, 42])
def test_multi_kernel_rk4_graph_replays_match_oracle(oracle_mod, order):
    """The multi-kernel RK4 loop replays chunks of 16 steps from a captured
    graph: 40 steps (two replays and 8 direct steps), then 33 steps with
    another dt (a new graph) on the same state, against the oracle."""
    N = 14
    _say(f"multi-kernel rk4 graph replays order={order}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 5)
    u, v = 10 * u, 10 * v
    ref = o.rk4(u, v, h, 300.0, 40)
    ref = o.rk4(*ref, 200.0, 33)
    op = cs.create_linear_swm_operator(N, order, cs.FP64, unfused=True)
    st = op.create_stvec().upload(u, v, h)
    ts = cs.create_timescheme("rk4")
    ts.step(st, op, 300.0, 40)
    ts.step(st, op, 200.0, 33)
    for g, r in zip(st.download(), ref):
        assert _rel(g, r) < 1e-12
