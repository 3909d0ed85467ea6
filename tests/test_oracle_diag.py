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


def _fd5(f, eps):
    """Five-point derivative at 0 (exact for quartics)."""
    return (8 * (f(eps) - f(-eps)) - (f(2 * eps) - f(-2 * eps))) / (12 * eps)


def _tends_state(o, seed, h_var):
    rng = np.random.default_rng(seed)
    u, v = 10 * rng.standard_normal(o.shape(MESH_U)), 10 * rng.standard_normal(o.shape(MESH_V))
    h = 500.0 + h_var * rng.standard_normal(o.shape(MESH_XY))
    return u, v, h


@pytest.mark.parametrize("order", [21, 42])
def test_tends_are_derivatives_of_the_diagnostics(oracle_mod, order):
    """print_integral_tends against the time derivative of
    print_integrals_diag along the operator's own tendency F = apply(x),
    taken by finite differences of the oracle's diagnostics at x + eps F.

    Enstrophy: d/dt sum w (z + f)^2 / (2 hq) is exactly the two terms
    print_integral_tends sums, and hq from the u path alone equals the
    diagnostics' average of the u and v paths (the i2c interpolations along x
    and y commute), for any h.  TE: with a constant layer h = c (I(h) = c) the
    KE part is c <co2contra(u), u_t> + I(h_t) terms, the derivative of
    c <co2contra(u), u> / 2 because the Ch co2contra is symmetric in the u/v
    quadrature (checked here too); for a varying layer only the PE part
    stays exact, checked against the dropped-PE variant's error.  Planted
    mistakes (a dropped KE term, PE term, or hq_t term, or f left out of the
    absolute vorticity) each fail one of these."""
    o = oracle_mod.Oracle(16, order, radius=A, omega=7.292e-5, grav=9.80616, h_mean=1000.0)
    g = 9.80616
    rng = np.random.default_rng(11)
    a = rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V))
    b = rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V))
    ca, cb = o.co2contra(*a), o.co2contra(*b)
    wu, wv = o.weights(MESH_U).reshape(o.shape(MESH_U)), o.weights(MESH_V).reshape(o.shape(MESH_V))
    s1 = np.sum(wu * ca[0] * b[0]) + np.sum(wv * ca[1] * b[1])
    s2 = np.sum(wu * cb[0] * a[0]) + np.sum(wv * cb[1] * a[1])
    assert abs(s1 - s2) < 1e-13 * abs(s1)

    def check(u, v, h, mutation=0):
        ut, vt, ht = o.apply(u, v, h)
        eps = 1e-3 * 500 / np.abs(ht).max()
        d = {k: _fd5(lambda e: o.integrals_diag(u + e * ut, v + e * vt, h + e * ht, g)[k], eps)
             for k in ("TE", "Enstrophy")}
        oracle_mod.set_mutation(mutation)
        try:
            te, en = o.integral_tends(u, v, h, g)
        finally:
            oracle_mod.set_mutation(0)
        wh = o.weights(MESH_XY).reshape(h.shape)
        pe_scale = g * np.sum(wh * np.abs(h * ht))
        return abs(d["TE"] - te) / pe_scale, abs(d["Enstrophy"] - en) / abs(en)

    uc, vc, hc = _tends_state(o, 12, 0.0)
    uvar, vvar, hvar = _tends_state(o, 13, 100.0)
    te_c, en_c = check(uc, vc, hc)
    te_v, en_v = check(uvar, vvar, hvar)
    assert te_c < 1e-9 and en_c < 1e-7 and en_v < 1e-7, (te_c, en_c, en_v)
    assert te_v < 2e-4, te_v  # the KE terms' asymmetry under a varying I(h)
    assert check(uc, vc, hc, 6)[0] > 1e3 * max(te_c, 1e-12)   # KE term of h_t dropped
    assert check(uvar, vvar, hvar, 7)[0] > 20 * te_v           # PE term dropped
    assert check(uc, vc, hc, 8)[1] > 0.1                        # hq_t term dropped
    assert check(uvar, vvar, hvar, 9)[1] > 0.1                  # f left out


def test_tends_at_rest_and_scales(oracle_mod):
    """At rest with a constant layer the tendency vanishes (up to the
    rounding of a gradient of a constant) and so do both results; the
    returned scales bound the results."""
    o = oracle_mod.Oracle(12, 42, radius=A, omega=7.292e-5)
    z = (o.zeros(MESH_U), o.zeros(MESH_V))
    te, en, ste, sen = o.integral_tends(*z, np.full(o.shape(MESH_XY), 300.0), scales=True)
    assert abs(te) < 1e-30 and abs(en) < 1e-50, (te, en)
    u, v, h = _tends_state(o, 14, 30.0)
    te, en, ste, sen = o.integral_tends(u, v, h, scales=True)
    assert abs(te) <= ste and abs(en) <= sen and ste > 0 and sen > 0


def test_cfl_at_rest(oracle_mod):
    """print_CFL_diag at rest on a layer of depth c: the gravity-wave CFL is
    sqrt(g c) dt / dx with dx = a pi / (2N), the advective ones vanish."""
    N, c, dt, g = 20, 800.0, 600.0, 9.80616
    o = oracle_mod.Oracle(N, 42, radius=A)
    cg, cx, cy = o.cfl_diag(o.zeros(MESH_U), o.zeros(MESH_V), np.full(o.shape(MESH_XY), c), dt, g)
    assert cg == pytest.approx(np.sqrt(g * c) * dt / (A * np.pi / (2 * N)), rel=1e-14)
    assert cx == 0.0 and cy == 0.0


@pytest.mark.parametrize("order,rate", [(21, 0.9), (42, 1.8)])
def test_cfl_solid_rotation_converges(oracle_mod, order, rate):
    """Solid rotation about a tilted axis: the advective CFLs converge to
    sqrt(max |V.a^1 V.a_1|) dt / dx with the exact contravariant component
    V.a^1 = g^11 V.a_1 + g^12 V.a_2 from the basis vectors (the oracle's
    co2contra is the SBP approximation of it)."""
    axis, u0, dt = (0.2, 0.4, 0.9), 40.0, 300.0
    errs = []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, order, radius=A)
        u, v = _solid_rotation(o, u0, axis)
        exact = []
        for m, cov in ((MESH_U, u), (MESH_V, v)):
            a1, a2 = o.basis(m)
            g11, g12, g22 = (np.einsum("...k,...k->...", a, b) for a, b in ((a1, a1), (a1, a2), (a2, a2)))
            det = g11 * g22 - g12 * g12
            ax = np.array(axis) / np.linalg.norm(axis)
            r = np.cross(a1, a2)
            r /= np.linalg.norm(r, axis=-1, keepdims=True)
            V = u0 * np.cross(ax, r)
            c1, c2 = np.einsum("...k,...k->...", V, a1), np.einsum("...k,...k->...", V, a2)
            contra = (g22 * c1 - g12 * c2) / det if m == MESH_U else (-g12 * c1 + g11 * c2) / det
            own = c1 if m == MESH_U else c2
            exact.append(np.sqrt(np.abs(contra * own).max()) * dt / (A * np.pi / (2 * N)))
        got = o.cfl_diag(u, v, np.full(o.shape(MESH_XY), 100.0), dt)[1:]
        errs.append(max(abs(got[k] / exact[k] - 1) for k in range(2)))
    assert errs[-1] < 1e-3 and np.all(_conv(errs) > rate), errs
