"""Pins of the oracle's lat-lon output of h (swo_latlon_h): what the
shallow-water output writes to h.dat (latlon_outputer_mod.f90 L33-63,
latlon_regrid_mod.f90, ecs_halo_xy_mod.f90, ecs_halo_factory_mod.f90
L154-269): the exchanged xy field with its ECS halo (edge and corner copies
averaged, virtual rows interpolated along the neighbour's rows), cubic
interpolation at the lat-lon points, float32 output.

The checks: a smooth function on the sphere converges at fourth order to
its exact lat-lon values (any wrong panel, orientation, row or weight
breaks this); lat-lon points that fall on grid nodes return the node value;
the output depends on the copies of an edge or corner point only through
their sum (the averaging); constants are reproduced.  The winds (u.dat,
v.dat: swo_latlon_uv) of solid rotation converge to the exact zonal and
meridional components, and a planted mistake in the vector halo's frame
change breaks that."""
import numpy as np
import pytest

from oracle.swe_oracle import MESH_XY

EX = np.array([[-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 0, 1], [1, 0, 0], [1, 0, 0]])
EY = np.array([[0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]])
NN = np.array([[0, 0, 1], [-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 1, 0], [0, -1, 0]])


def xy_points(N):
    """Unit vectors of the xy points, (6, N+1, N+1, 3) indexed [p, j, i]."""
    t = np.tan(-0.25 * np.pi + np.arange(N + 1) * 0.5 * np.pi / N)
    X, Y = np.meshgrid(t, t)
    s = np.sqrt(1 + X**2 + Y**2)
    return (EX[:, None, None, :] * X[..., None] + EY[:, None, None, :] * Y[..., None]
            - NN[:, None, None, :]) / s[..., None]


def latlon_points(nlat, nlon):
    phi = -0.5 * np.pi + np.arange(nlat) * np.pi / (nlat - 1)
    lam = np.arange(nlon) * 2 * np.pi / nlon
    P, L = np.meshgrid(phi, lam, indexing="ij")
    return np.stack([np.cos(P) * np.cos(L), np.cos(P) * np.sin(L), np.sin(P)], -1)


def smooth(r):
    x, y, z = r[..., 0], r[..., 1], r[..., 2]
    return np.exp(x) + 2 * y * z + np.sin(3 * z)


def test_smooth_field_converges(oracle_mod):
    """Max error of a smooth field against its exact lat-lon values falls by
    >= 2^3.3 per doubling of N (cubic interpolation in the panels and in the
    ECS halo) down to float32 rounding."""
    errs = []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, 42)
        out = o.latlon_h(smooth(xy_points(N)))
        assert out.shape == (2 * N + 1, 4 * N) and out.dtype == np.float32
        errs.append(np.abs(out - smooth(latlon_points(2 * N + 1, 4 * N))).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[-1] < 2e-6 and np.all(rates > 3.3), (errs, rates)


def test_constant_is_reproduced(oracle_mod):
    o = oracle_mod.Oracle(10, 21)
    out = o.latlon_h(np.full(o.shape(MESH_XY), 7.25))
    assert np.all(out == np.float32(7.25))


def test_nodes_return_node_values(oracle_mod):
    """With N even the equator and the lat-lon meridians at multiples of
    90/N degrees pass through grid nodes; at every lat-lon point that lands
    on a node inside a panel the output is that node's value (within one
    float32 rounding), for a random field."""
    N = 12
    o = oracle_mod.Oracle(N, 42)
    rng = np.random.default_rng(4)
    h = rng.standard_normal(o.shape(MESH_XY))
    out = o.latlon_h(h)
    nodes = xy_points(N).reshape(-1, 3)
    inner = np.zeros((6, N + 1, N + 1), bool)
    inner[:, 1:N, 1:N] = True
    pts = latlon_points(2 * N + 1, 4 * N).reshape(-1, 3)
    hits = 0
    for k, r in enumerate(pts):
        d = np.linalg.norm(nodes - r, axis=1)
        m = np.argmin(d)
        if d[m] < 1e-12 and inner.ravel()[m]:
            hits += 1
            assert abs(out.ravel()[k] - np.float32(h.ravel()[m])) <= np.spacing(np.float32(abs(h.ravel()[m]))), k
    assert hits > 6 * N  # the equator, the meridians and the poles


def _copies(N):
    """Groups of packed indices of the same point on different panels."""
    r = xy_points(N).reshape(-1, 3)
    key = np.round(r * 1e9).astype(np.int64)
    groups = {}
    for k, kk in enumerate(map(tuple, key)):
        groups.setdefault(kk, []).append(k)
    return [g for g in groups.values() if len(g) > 1]


@pytest.mark.parametrize("mutation,invariant", [(0, True), (10, False)])
def test_output_depends_on_copies_through_their_sum(oracle_mod, mutation, invariant):
    """Moving value between the copies of an edge point (two panels) or a
    cube corner (three) leaves the output unchanged: the ECS halo averages
    the copies before anything reads them.  Edge points within two of a
    corner are left alone (the corner diagonals read one panel's copy of
    them as it is).  Without the averaging (mutation 10) it changes."""
    N = 16
    o = oracle_mod.Oracle(N, 21)
    rng = np.random.default_rng(8)
    h = rng.standard_normal(o.shape(MESH_XY))
    hp = h.copy().ravel()
    ij = np.stack(np.meshgrid(np.arange(N + 1), np.arange(N + 1)), -1)  # [j, i] -> (i, j)
    ij = np.broadcast_to(ij, (6, N + 1, N + 1, 2)).reshape(-1, 2)
    moved = 0
    for g in _copies(N):
        if len(g) == 2:
            a = np.minimum(ij[g[0]], N - ij[g[0]]).max()  # distance along the edge to a corner
            if a < 3:
                continue
            d = rng.standard_normal()
            hp[g[0]] += d
            hp[g[1]] -= d
        else:
            d1, d2 = rng.standard_normal(2)
            hp[g[0]] += d1
            hp[g[1]] += d2
            hp[g[2]] -= d1 + d2
        moved += 1
    assert moved == 8 + 12 * (N + 1 - 6)
    oracle_mod.set_mutation(mutation)
    try:
        a = o.latlon_h(h)
        b = o.latlon_h(hp.reshape(h.shape))
    finally:
        oracle_mod.set_mutation(0)
    diff = np.abs(a.astype(np.float64) - b).max()
    if invariant:
        assert diff <= 4 * np.spacing(np.float32(np.abs(a).max())), diff
    else:
        assert diff > 1e-2


@pytest.mark.parametrize("mutation", [11, 12])
def test_planted_mistakes_break_convergence(oracle_mod, mutation):
    """The neighbour's row one short (11) or alpha and beta exchanged (12)
    leave errors of order 1e-2 or more that do not fall with N."""
    oracle_mod.set_mutation(mutation)
    try:
        o = oracle_mod.Oracle(32, 42)
        err = np.abs(o.latlon_h(smooth(xy_points(32))) - smooth(latlon_points(65, 128))).max()
    finally:
        oracle_mod.set_mutation(0)
    assert err > 5e-3


def _rotation_winds(o, u0, axis):
    """Covariant components of V = u0 axis x r on the u and v points."""
    from oracle.swe_oracle import MESH_U, MESH_V
    ax = np.asarray(axis, float) / np.linalg.norm(axis)
    comps = []
    for m in (MESH_U, MESH_V):
        a1, a2 = o.basis(m)
        r = np.cross(a1, a2)
        r /= np.linalg.norm(r, axis=-1, keepdims=True)
        V = u0 * np.cross(ax, r)
        comps.append(np.einsum("...k,...k->...", V, a1 if m == MESH_U else a2).reshape(o.shape(m)))
    return comps


def _exact_latlon_winds(N, u0, axis):
    nlat, nlon = 2 * N + 1, 4 * N
    lam = np.broadcast_to(np.arange(nlon) * 2 * np.pi / nlon, (nlat, nlon))
    phi = np.broadcast_to((-0.5 * np.pi + np.arange(nlat) * np.pi / (nlat - 1))[:, None], (nlat, nlon))
    V = u0 * np.cross(np.asarray(axis, float) / np.linalg.norm(axis), latlon_points(nlat, nlon))
    iv = np.stack([-np.sin(lam), np.cos(lam), 0 * lam], -1)
    jv = np.stack([-np.sin(phi) * np.cos(lam), -np.sin(phi) * np.sin(lam), np.cos(phi)], -1)
    return np.einsum("...k,...k->...", V, iv), np.einsum("...k,...k->...", V, jv)


def test_winds_converge(oracle_mod):
    """u.dat / v.dat of solid rotation about a tilted axis: the zonal and
    meridional winds converge to their exact lat-lon values at second order
    (the colocating c2i interpolation's closure), everywhere including the
    poles, the cube edges and corners (the vector halo's transform)."""
    u0, axis = 20.0, (0.3, -0.5, 0.8)
    errs = []
    for N in (16, 32, 64):
        o = oracle_mod.Oracle(N, 42)
        uo, vo = o.latlon_uv(*_rotation_winds(o, u0, axis))
        ez, em = _exact_latlon_winds(N, u0, axis)
        assert uo.dtype == np.float32 and uo.shape == (2 * N + 1, 4 * N)
        errs.append(max(np.abs(uo - ez).max(), np.abs(vo - em).max()) / u0)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[-1] < 3e-4 and np.all(rates > 1.7), (errs, rates)


def test_winds_planted_halo_frame_mistake(oracle_mod):
    """Leaving the exchanged components in the neighbour's frame (mutation
    14) puts errors of the size of the wind near the cube edges."""
    u0, axis, N = 20.0, (0.3, -0.5, 0.8), 32
    oracle_mod.set_mutation(14)
    try:
        o = oracle_mod.Oracle(N, 42)
        uo, vo = o.latlon_uv(*_rotation_winds(o, u0, axis))
    finally:
        oracle_mod.set_mutation(0)
    ez, em = _exact_latlon_winds(N, u0, axis)
    assert max(np.abs(uo - ez).max(), np.abs(vo - em).max()) > 0.1 * u0
