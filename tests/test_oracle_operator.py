"""Pins the oracle's full operator (operator_swm_lin_mod.f90 apply L66-108)
and RK4 step (rk4_mod.f90 L20-42) with invariants and closed forms:

* semi-discrete energy neutrality <grad E, A(y)> = 0 for continuous h
  (SBP + SAT + projection + covariant edge sync + symmetric co2contra);
* exact Coriolis neutrality and exact SBP mass conservation;
* continuity of the h tendency across duplicated edge points (projection);
* convergence of grad / div / co2contra to analytic fields;
* Williamson test 2 in linear mode is a steady state (its own KAT:
  models/shallow_water/testcases/Williamson_test2_mod.f90 L113-143 scores
  against the initial state);
* RK4 on a linear operator equals the degree-4 Taylor polynomial of exp(dt A);
* planted mistakes in the oracle (swo_set_mutation) break these checks.
"""
import numpy as np
import pytest

from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

CORNER = (np.pi / 4, np.arcsin(1 / np.sqrt(3)))


def _smooth_h(o):
    return (o.gaussian_hill(*CORNER, 0.3, 10.0) + o.gaussian_hill(0.3, 0.2, 0.5, 3.0)
            + o.gaussian_hill(2.0, -0.9, 0.4, -4.0))


def _random_uv(o, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V))


def _energy_rate(o, u, v, h):
    du, dv, dh = o.apply(u, v, h)
    ut, vt = o.co2contra(u, v)
    wu, wv, wh = o.weights(MESH_U), o.weights(MESH_V), o.weights(MESH_XY)
    dE = o.grav * np.sum(h * dh * wh) + o.h_mean * (np.sum(ut * du * wu) + np.sum(vt * dv * wv))
    scale = (o.grav * np.sqrt(np.sum(h * h * wh) * np.sum(dh * dh * wh))
             + o.h_mean * np.sqrt((np.sum(ut * ut * wu) + np.sum(vt * vt * wv))
                                  * (np.sum(du * du * wu) + np.sum(dv * dv * wv))))
    return dE / scale


def _dup_jump(o, f):
    """max |f(own edge point) - f(neighbour's duplicate)| over the xy map."""
    hm = o.halo_map(0)
    N = o.N
    ei = np.clip(hm["ti"], 1, N + 1) - 1
    ej = np.clip(hm["tj"], 1, N + 1) - 1
    return np.abs(f[hm["tp"], ej, ei] - f[hm["sp"], hm["sj"] - 1, hm["si"] - 1]).max()


@pytest.mark.parametrize("order,N", [(21, 12), (21, 17), (42, 12), (42, 20), (63, 14), (63, 24)])
@pytest.mark.parametrize("cor_const", [False, True])
def test_energy_neutrality(oracle_mod, order, N, cor_const):
    o = oracle_mod.Oracle(N, order, coriolis_const=cor_const, omega=7.292e-5 if not cor_const else 1e-4)
    u, v = _random_uv(o)
    assert abs(_energy_rate(o, u, v, _smooth_h(o))) < 1e-13


@pytest.mark.parametrize("order", [21, 42, 63])
def test_coriolis_neutral_and_mass_conserved(oracle_mod, order):
    o = oracle_mod.Oracle(16, order)
    u, v = _random_uv(o, 3)
    rng = np.random.default_rng(4)
    h = rng.standard_normal(o.shape(MESH_XY))
    ut, vt = o.co2contra(u, v)
    cu, cv = o.coriolis(ut, vt)
    wu, wv, wh = o.weights(MESH_U), o.weights(MESH_V), o.weights(MESH_XY)
    work = np.sum(ut * cu * wu) + np.sum(vt * cv * wv)
    norm = np.sqrt((np.sum(ut * ut * wu) + np.sum(vt * vt * wv)) * (np.sum(cu * cu * wu) + np.sum(cv * cv * wv)))
    assert abs(work / norm) < 1e-14
    du, dv, dh = o.apply(u, v, h)
    assert abs(np.sum(dh * wh)) < 1e-14 * np.sqrt(np.sum(dh * dh * wh))
    # the projection makes the h tendency continuous across panel edges
    assert _dup_jump(o, dh) < 1e-14 * np.abs(dh).max()


@pytest.mark.parametrize("order", [21, 42, 63])
def test_co2contra_symmetric_positive(oracle_mod, order):
    o = oracle_mod.Oracle(14, order)
    wu, wv = o.weights(MESH_U), o.weights(MESH_V)
    a1, a2 = _random_uv(o, 5)
    b1, b2 = _random_uv(o, 6)
    ca1, ca2 = o.co2contra(a1, a2)
    cb1, cb2 = o.co2contra(b1, b2)
    lhs = np.sum(a1 * cb1 * wu) + np.sum(a2 * cb2 * wv)
    rhs = np.sum(ca1 * b1 * wu) + np.sum(ca2 * b2 * wv)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)
    assert np.sum(a1 * ca1 * wu) + np.sum(a2 * ca2 * wv) > 0


def _conv(errs):
    return np.log2(np.array(errs[:-1]) / np.array(errs[1:]))


@pytest.mark.parametrize("order,rate", [(21, 1.4), (42, 1.8), (63, 2.6)])
def test_grad_convergence(oracle_mod, order, rate):
    """h = c . r on the unit sphere: covariant gradient = c . a / radius."""
    c = np.array([0.3, -0.5, 0.8])
    errs = []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, order)
        g = o.metric(MESH_XY)
        h = c[0] * g["rx"] + c[1] * g["ry"] + c[2] * g["rz"]
        h = np.sin(2 * h) + h
        gx, gy = o.grad(h)
        gu, gv = o.metric(MESH_U), o.metric(MESH_V)
        a1u = o.basis(MESH_U)[0].reshape(gx.shape + (3,))
        a2v = o.basis(MESH_V)[1].reshape(gy.shape + (3,))
        ru = np.stack([gu["rx"], gu["ry"], gu["rz"]], -1)
        rv = np.stack([gv["rx"], gv["ry"], gv["rz"]], -1)
        dfu = (2 * np.cos(2 * (ru @ c)) + 1)
        dfv = (2 * np.cos(2 * (rv @ c)) + 1)
        ex = dfu * (a1u @ c) / o.radius
        ey = dfv * (a2v @ c) / o.radius
        errs.append(max(np.abs(gx - ex).max(), np.abs(gy - ey).max()) * o.radius)
    assert np.all(_conv(errs) > rate), errs


C_VEC = np.array([0.2, 0.7, -0.4])


def _analytic(o, m):
    """V = c - (c.r) r (tangential part of a constant vector) at mesh m, with
    the covariant basis a1, a2 and the contravariant basis b1, b2 = Qi a."""
    g = o.metric(m)
    r = np.stack([g["rx"], g["ry"], g["rz"]], -1)
    V = C_VEC - (r @ C_VEC)[..., None] * r
    a1, a2 = (a.reshape(r.shape) for a in o.basis(m))
    b1 = g["Qi11"][..., None] * a1 + g["Qi12"][..., None] * a2
    b2 = g["Qi12"][..., None] * a1 + g["Qi22"][..., None] * a2
    return dict(r=r, V=V, a1=a1, a2=a2, b1=b1, b2=b2)


def _subop_errors(o):
    """max errors of co2contra, div and Coriolis against closed forms:
    contravariant components V.b; div V = -2 (c.r)/a on a sphere of radius
    a; covariant Coriolis term (-f r x V).a with f = 2 Omega (axis . r)."""
    Fu, Fv, Fp = _analytic(o, MESH_U), _analytic(o, MESH_V), _analytic(o, MESH_XY)
    u = (Fu["V"] * Fu["a1"]).sum(-1)
    v = (Fv["V"] * Fv["a2"]).sum(-1)
    ut_x = (Fu["V"] * Fu["b1"]).sum(-1)
    vt_x = (Fv["V"] * Fv["b2"]).sum(-1)
    ut, vt = o.co2contra(u, v)
    e_c2c = max(np.abs(ut - ut_x).max(), np.abs(vt - vt_x).max())
    d = o.div(ut_x, vt_x)
    e_div = np.abs(d * o.radius + 2 * (Fp["r"] @ C_VEC)).max()
    cu, cv = o.coriolis(ut_x, vt_x)
    ex = []
    for F, a in ((Fu, "a1"), (Fv, "a2")):
        f = 2 * o.omega * F["r"][..., 2]
        ex.append((-f[..., None] * np.cross(F["r"], F["V"]) * F[a]).sum(-1))
    e_cor = max(np.abs(cu - ex[0]).max(), np.abs(cv - ex[1]).max()) / (2 * o.omega)
    return e_c2c, e_div, e_cor


@pytest.mark.parametrize("order,rate", [(21, 0.75), (42, 1.8), (63, 1.95)])
def test_co2contra_div_coriolis_convergence(oracle_mod, order, rate):
    """Max-norm errors are set by the boundary closures: first order for
    sbp21, second order for sbp42, between second and third for sbp63 (and
    3 to 10 times smaller than sbp42's at the same N)."""
    errs = np.array([_subop_errors(oracle_mod.Oracle(N, order)) for N in (16, 32, 64)])
    for k in range(3):
        assert np.all(_conv(errs[:, k]) > rate), errs


@pytest.mark.parametrize("order", [21, 42, 63])
def test_ts2_linear_steady_state(oracle_mod, order):
    """Geostrophically balanced solid rotation: the tendency is pure
    truncation error and converges to zero (max norm: first order, set by
    the points next to the cube corners; quadrature L2: ~1.5)."""
    emax, el2 = [], []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, order, grav=9.80616, h_mean=29400 / 9.80616,
                              radius=6371229.0, omega=2 * np.pi / 86164.0)
        u0 = np.pi * o.radius / 6 / 86164.0
        u, v, h = o.ts2_linear(u0, o.h_mean)
        du, dv, dh = o.apply(u, v, h)
        sc = o.h_mean * u0 / o.radius
        wh = o.weights(MESH_XY)
        emax.append(np.abs(dh).max() / sc)
        el2.append(np.sqrt(np.sum(dh**2 * wh) / np.sum(wh)) / sc)
    assert emax[-1] < 0.03 and el2[-1] < 3e-3
    assert np.all(_conv(emax) > 0.8), emax
    assert np.all(_conv(el2) > 1.3), el2


@pytest.mark.parametrize("order", [21, 42, 63])
def test_rk4_is_taylor_polynomial(oracle_mod, order):
    o = oracle_mod.Oracle(10 if order != 63 else 14, order, coriolis_const=True, omega=1e-4)
    u, v = _random_uv(o, 7)
    h = _smooth_h(o)
    dt = 900.0
    terms = [(u, v, h)]
    for k in range(4):
        terms.append(tuple(dt * x / (k + 1) for x in o.apply(*terms[-1])))
    expect = [sum(t[c] for t in terms) for c in range(3)]
    got = o.rk4(u, v, h, dt, 1)
    for c in range(3):
        assert np.abs(got[c] - expect[c]).max() < 1e-13 * np.abs(expect[c]).max()


def test_gaussian_energy_drift_is_time_scheme_only(oracle_mod):
    """Config-1 run (N=32, sbp21, corner Gaussian, 60000 s): the space
    discretisation conserves energy, so the drift is RK4's own damping,
    which falls as dt^4 over a fixed time (measured: ~x31 from dt=600 s to
    300 s).  Mass is conserved to roundoff."""
    o = oracle_mod.Oracle(32, 21)
    h0 = o.gaussian_hill(*CORNER, 5e5 / o.radius, 10.0)
    z = (o.zeros(MESH_U), o.zeros(MESH_V))
    e0, m0 = o.energy(*z, h0), o.mass(h0)
    drift = []
    for dt, n in ((600.0, 100), (300.0, 200)):
        u, v, h = o.rk4(*z, h0, dt, n)
        drift.append(abs(o.energy(u, v, h) / e0 - 1))
        assert abs(o.mass(h) / m0 - 1) < 1e-12
    assert drift[0] < 1e-3
    assert drift[1] < drift[0] / 16


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(oracle_mod, mutation):
    """Each planted mistake (rotated-edge flux sign, dropped sync metric term,
    missing h halo, one wrong closure coefficient, missing projection) must
    fail at least one of the checks above."""
    try:
        oracle_mod.set_mutation(mutation)
        o = oracle_mod.Oracle(12, 42)
        u, v = _random_uv(o)
        h = _smooth_h(o)
        caught = abs(_energy_rate(o, u, v, h)) > 1e-10
        du, dv, dh = o.apply(u, v, h)
        caught |= _dup_jump(o, dh) > 1e-10 * np.abs(dh).max()
        ut, vt = o.co2contra(u, v)
        cu, cv = o.coriolis(ut, vt)
        wu, wv = o.weights(MESH_U), o.weights(MESH_V)
        caught |= abs(np.sum(ut * cu * wu) + np.sum(vt * cv * wv)) > 1e-10 * np.sqrt(
            (np.sum(ut * ut * wu) + np.sum(vt * vt * wv)) * (np.sum(cu * cu * wu) + np.sum(cv * cv * wv)))
        # consistency: errors against closed forms no longer fall with N
        e16 = np.array(_subop_errors(oracle_mod.Oracle(16, 42)))
        e32 = np.array(_subop_errors(oracle_mod.Oracle(32, 42)))
        caught |= bool(np.any(e32 > e16 / 2.5))
        assert caught
    finally:
        oracle_mod.set_mutation(0)
