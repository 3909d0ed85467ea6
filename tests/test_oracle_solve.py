"""Pins of the oracle's implicit solve (swo_solve: operator_swm_lin_t%solve,
operator_swm_lin_mod.f90 L110-146, with cg_solver_t, cg_solver_mod.f90
L19-90): the result satisfies the implicit Euler equation of the linear
operator without Coriolis, v - dt L0 v = rhs (checked with the oracle's
apply, which the golden vectors pin), to the CG tolerance in h and to
rounding in u, v; the energy identities that follow from L0 being
skew-adjoint in the SBP energy (implicit Euler: E(v) = <v, rhs>_E <= E(rhs);
Crank-Nicolson: E conserved) hold; a planted mistake in the Helmholtz
operator breaks them."""
import numpy as np
import pytest

from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

A = 6371229.0


def _state(o, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V)),
            rng.standard_normal(o.shape(MESH_XY)))


def _smooth_state(o):
    """A state the operator keeps consistent (continuous h, synchronised
    edge winds): a Gaussian hill after a few RK4 steps; L0 is skew-adjoint
    in the SBP energy on such states."""
    h0 = o.gaussian_hill(0.3, 0.4, 0.2, 10.0)
    return o.rk4(o.zeros(MESH_U), o.zeros(MESH_V), h0, 600.0, 5)


def _energy_dot(o, x, y):
    """<x, y>_E by polarisation of the oracle's quadratic energy."""
    p = o.energy(*[a + b for a, b in zip(x, y)])
    m = o.energy(*[a - b for a, b in zip(x, y)])
    return 0.25 * (p - m)


@pytest.mark.parametrize("order", [21, 42])
@pytest.mark.parametrize("tol", [1e-8, 1e-12])
def test_solve_is_implicit_euler(oracle_mod, order, tol):
    o = oracle_mod.Oracle(20, order, radius=A, omega=0.0, grav=9.80616, h_mean=1000.0)
    rhs = _state(o, 3)
    dt = 3000.0
    u, v, h, it, rel = o.solve(*rhs, dt, tol, 400)
    assert 0 < it <= 400 and rel < tol
    du, dv, dh = o.apply(u, v, h)
    for x, d, r, bar in ((u, du, rhs[0], 1e-14), (v, dv, rhs[1], 1e-14), (h, dh, rhs[2], 200 * tol)):
        assert np.abs(x - dt * d - r).max() <= bar * np.abs(r).max()


def test_solve_ignores_coriolis(oracle_mod):
    """The solve inverts I - dt L0 whatever the operator's rotation."""
    a = oracle_mod.Oracle(14, 42, radius=A, omega=7.292e-5)
    b = oracle_mod.Oracle(14, 42, radius=A, omega=0.0)
    rhs = _state(a, 5)
    ra, rb = a.solve(*rhs, 1800.0, 1e-12, 300), b.solve(*rhs, 1800.0, 1e-12, 300)
    for x, y in zip(ra[:3], rb[:3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("order", [21, 42])
def test_implicit_euler_energy(oracle_mod, order):
    """v - dt L0 v = b with <v, L0 v>_E = 0 gives E(v) = <v, b>_E <= E(b),
    on a consistent state (on random ones L0 is not skew: duplicated edge
    values that disagree)."""
    o = oracle_mod.Oracle(16, order, radius=A, omega=0.0)
    b = _smooth_state(o)
    u, v, h, it, rel = o.solve(*b, 4000.0, 1e-13, 400)
    x = (u, v, h)
    e, eb = o.energy(*x), o.energy(*b)
    assert abs(e - _energy_dot(o, x, b)) <= 1e-10 * eb
    assert e < eb


@pytest.mark.parametrize("order", [21, 42])
def test_crank_nicolson_conserves_energy(oracle_mod, order):
    """crank_nikolson_mod.f90 L20-32: v1 = solve(v0 + dt/2 L0 v0, dt/2) keeps
    the energy of a skew-adjoint L0."""
    o = oracle_mod.Oracle(16, order, radius=A, omega=0.0)
    v0 = _smooth_state(o)
    dt = 4000.0
    d = o.apply(*v0)
    rhs = [a + 0.5 * dt * b for a, b in zip(v0, d)]
    u, v, h, it, rel = o.solve(*rhs, 0.5 * dt, 1e-13, 400)
    assert abs(o.energy(u, v, h) / o.energy(*v0) - 1) < 1e-12


def test_zero_right_hand_side(oracle_mod):
    o = oracle_mod.Oracle(8, 21, radius=A)
    z = (o.zeros(MESH_U), o.zeros(MESH_V), o.zeros(MESH_XY))
    u, v, h, it, rel = o.solve(*z, 100.0, 1e-10, 10)
    assert it == 0 and not np.any(u) and not np.any(v) and not np.any(h)


def test_planted_helmholtz_sign_breaks_the_identity(oracle_mod):
    """Mutation 13 flips gamma: the implicit Euler residual in h is then of
    the order of the right-hand side (or CG fails outright)."""
    o = oracle_mod.Oracle(12, 42, radius=A, omega=0.0)
    rhs = _state(o, 11)
    oracle_mod.set_mutation(13)
    try:
        u, v, h, it, rel = o.solve(*rhs, 3000.0, 1e-10, 200)
    finally:
        oracle_mod.set_mutation(0)
    du, dv, dh = o.apply(u, v, h)
    res = np.abs(h - 3000.0 * dh - rhs[2]).max() / np.abs(rhs[2]).max()
    assert not np.isfinite(res) or res > 1e-2
