"""Pins of the oracle's integral diagnostics (print_integrals_diag /
calc_energies, swm_output_diag_mod.f90 L156-234) and its Ch curl
(curl_ch_sbp_sat, curl_sbp_sat_mod.f90 L26-74): closed forms on the sphere,
the discrete Stokes identity the SAT terms make exact, and the identities
that tie the diagnostics to the oracle's own mass and quadrature."""
import numpy as np
import pytest

from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

A = 6371229.0


def _conv(errs):
    return np.log2(np.array(errs[:-1]) / np.array(errs[1:]))


def _solid_rotation(o, u0, axis):
    """Covariant components of V = u0 axis x r (axis normalised) on the u
    and v points."""
    ax = np.asarray(axis, float) / np.linalg.norm(axis)
    comps = []
    for m in (MESH_U, MESH_V):
        a1, a2 = o.basis(m)
        r = np.cross(a1, a2)  # normal direction = r (unit sphere, J > 0)
        r /= np.linalg.norm(r, axis=-1, keepdims=True)
        V = u0 * np.cross(ax, r)
        comps.append(np.einsum("...k,...k->...", V, a1 if m == MESH_U else a2).reshape(o.shape(m)))
    return comps


def _centres(o):
    """Unit vectors at the cell centres, (6, N, N, 3)."""
    N = o.N
    hx = 0.5 * np.pi / N
    t = np.tan(-0.25 * np.pi + (np.arange(1, N + 1) - 0.5) * hx)
    X, Y = np.meshgrid(t, t)  # [j, i]
    ex = np.array([[-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 0, 1], [1, 0, 0], [1, 0, 0]])
    ey = np.array([[0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]])
    nn = np.array([[0, 0, 1], [-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 1, 0], [0, -1, 0]])
    s = np.sqrt(1 + X**2 + Y**2)
    r = (ex[:, None, None, :] * X[..., None] + ey[:, None, None, :] * Y[..., None] - nn[:, None, None, :]) / s[..., None]
    return r


@pytest.mark.parametrize("order,rate", [(21, 0.9), (42, 1.8)])
def test_curl_of_solid_rotation(oracle_mod, order, rate):
    """curl of u0 axis x r on a sphere of radius a = 2 u0 (axis . r) / a;
    max-norm convergence (closures: first order sbp21, second sbp42)."""
    axis = (0.3, -0.5, 0.8)
    errs = []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, order, radius=A)
        u0 = 20.0
        u, v = _solid_rotation(o, u0, axis)
        r = _centres(o)
        exact = 2 * u0 * (r @ (np.array(axis) / np.linalg.norm(axis))) / A
        errs.append(np.abs(o.curl(u, v) - exact).max() / (2 * u0 / A))
    assert errs[-1] < 1e-2
    assert np.all(_conv(errs) > rate), errs


@pytest.mark.parametrize("order", [21, 42])
def test_curl_discrete_stokes(oracle_mod, order):
    """For any field, sum over all centres of w_o * curl vanishes: the SAT
    terms replace each panel's edge values by the average with the
    neighbour's duplicate, so every edge's circulation cancels between its
    two panels.  A wrong sign, component or index in the tangential halo
    breaks this."""
    o = oracle_mod.Oracle(24, order, radius=A)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V))
    c = o.curl(u, v)
    # the centre weights (J hx^2 Ac Ac), checked against the oracle's own: at
    # rest with f = 1 and h = 1 the enstrophy is half their sum
    oc = oracle_mod.Oracle(24, order, radius=A, omega=1.0, coriolis_const=True)
    wsum = 2 * oc.integrals_diag(0 * u, 0 * v, np.ones(o.shape(MESH_XY)))["Enstrophy"]
    w_o = _centre_weights(o)
    assert abs(w_o.sum() - wsum) < 1e-12 * wsum
    total = np.sum(w_o * c)
    scale = np.sum(w_o * np.abs(c))
    assert abs(total) < 1e-13 * scale, (total, scale)


def _centre_weights(o):
    N = o.N
    hx = 0.5 * np.pi / N
    t = np.tan(-0.25 * np.pi + (np.arange(1, N + 1) - 0.5) * hx)
    X, Y = np.meshgrid(t, t)
    J = (1 + X**2) * (1 + Y**2) / np.sqrt(1 + X**2 + Y**2) ** 3
    a = np.ones(N)
    if o.order == 42:
        a[:4] = a[-4:][::-1] = [13 / 12, 7 / 8, 25 / 24, 1]
    W = J * hx * hx * a[None, :] * a[:, None]
    return np.broadcast_to(W, (6, N, N))


@pytest.mark.parametrize("order", [21, 42])
def test_centre_weights_integrate_the_sphere(oracle_mod, order):
    """The centre quadrature integrates the unit sphere's area, at least
    second order."""
    errs = [abs(_centre_weights(oracle_mod.Oracle(N, order)).sum() - 4 * np.pi) for N in (16, 32, 64)]
    assert errs[-1] < 1e-3
    assert np.all(_conv(errs) > 1.9), errs


@pytest.mark.parametrize("order", [21, 42])
def test_diag_at_rest(oracle_mod, order):
    """u = v = 0, h = c: KE = 0, PE = g c^2 / 2 * mass(1) (the oracle's own
    quadrature), hmin = hmax = c, mass = c * mass(1), enstrophy = f^2 / (2 c)
    * sum of the centre weights (constant f), TE = PE."""
    o = oracle_mod.Oracle(20, order, radius=A, omega=1e-4, coriolis_const=True)
    c, g = 250.0, 9.80616
    h = np.full(o.shape(MESH_XY), c)
    z = (o.zeros(MESH_U), o.zeros(MESH_V))
    d = o.integrals_diag(*z, h, g)
    m1 = o.mass(np.ones_like(h))
    assert d["hmin"] == c and d["hmax"] == c
    assert d["mass"] == pytest.approx(c * m1, rel=1e-14)
    assert d["KE"] == 0.0
    assert d["PE"] == pytest.approx(0.5 * g * c * c * m1, rel=1e-13)
    assert d["TE"] == d["PE"]
    assert d["Enstrophy"] == pytest.approx(0.5 * 1e-4**2 / c * _centre_weights(o).sum(), rel=1e-13)


@pytest.mark.parametrize("order,rate", [(21, 0.9), (42, 1.8)])
def test_diag_solid_rotation_closed_forms(oracle_mod, order, rate):
    """Solid rotation about z with h = c and f = 2 omega sin(lat):
    KE -> c u0^2 / 2 * 8 pi / 3, enstrophy -> 2 (u0/a + omega)^2 / c * 4 pi / 3
    (unit-sphere quadratures), converging with N."""
    c, u0, om = 300.0, 30.0, 7.292e-5
    ke_exact = 0.5 * c * u0**2 * 8 * np.pi / 3
    en_exact = 2 * (u0 / A + om) ** 2 / c * 4 * np.pi / 3
    eke, een = [], []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, order, radius=A, omega=om)
        u, v = _solid_rotation(o, u0, (0, 0, 1))
        d = o.integrals_diag(u, v, np.full(o.shape(MESH_XY), c))
        eke.append(abs(d["KE"] / ke_exact - 1))
        een.append(abs(d["Enstrophy"] / en_exact - 1))
    assert eke[-1] < 1e-3 and een[-1] < 1e-3, (eke, een)
    assert np.all(_conv(eke) > rate), eke
    assert np.all(_conv(een) > rate), een


def test_diag_extrema_and_mass_follow_h(oracle_mod):
    o = oracle_mod.Oracle(12, 42, radius=A)
    rng = np.random.default_rng(5)
    h = 100 + rng.standard_normal(o.shape(MESH_XY))
    u, v = rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V))
    d = o.integrals_diag(u, v, h, 9.80616)
    assert d["hmin"] == h.min() and d["hmax"] == h.max()
    assert d["mass"] == o.mass(h)
    assert d["TE"] == d["KE"] + d["PE"]
