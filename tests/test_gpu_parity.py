"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  FP64 bar: 1e-12 relative (max error / max magnitude of
the field).  FP32 bar: stated per test.  Full-size cases (N=384) are checked
through size-independent invariants instead of the oracle."""
import sys

import numpy as np
import pytest

import ch_swe_amd as cs
from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY

pytestmark = pytest.mark.gpu

CENTRES = [(np.pi, 0.0),                                           # panel interior
           (np.pi / 4, np.arcsin(1 / np.sqrt(3))),                 # cube corner
           (0.27 * np.pi, np.arcsin(1 / np.sqrt(3)) - 0.02 * np.pi)]  # 1 cell off


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _random_state(o, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(o.shape(MESH_U)), rng.standard_normal(o.shape(MESH_V)),
            rng.standard_normal(o.shape(MESH_XY)))


def _op(N, order, precision=cs.FP64, **kw):
    return cs.create_linear_swm_operator(N, order, precision, **kw)


# sizes chosen to exercise the fused tiling: single-tile panels (N=8), last
# tile one point wide in y (FP64 tiles are 32x8, FP32 32x16: N=16, 32) and
# in x (N=32), exact multiples (N=15, 31, 47), and panels large enough for
# edge-aligned tiling with tiles on the interior fast path (N=48, 57, 80, 97)
APPLY_CASES = [(21, 8), (21, 16), (21, 32), (21, 33), (21, 48), (21, 80), (42, 8), (42, 15),
               (42, 16), (42, 31), (42, 40), (42, 47), (42, 57), (42, 97)]


@pytest.mark.parametrize("order,N", APPLY_CASES)
@pytest.mark.parametrize("cor_const", [False, True])
def test_apply_fp64_matches_oracle(oracle_mod, order, N, cor_const):
    _say(f"apply fp64 order={order} N={N} cor_const={cor_const}")
    kw = dict(coriolis_const=cor_const, omega=1e-4 if cor_const else 7.292e-5)
    o = oracle_mod.Oracle(N, order, **kw)
    u, v, h = _random_state(o, 1000 + N)
    ref = o.apply(u, v, h)
    op = _op(N, order, **kw)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    got = vout.download()
    for g, r in zip(got, ref):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("general,linear", [(True, False), (False, True)])
def test_launch_variants_match_oracle(oracle_mod, general, linear):
    """Every tile on the general path, and the plain block order, give the
    same operator."""
    N, order = 80, 42
    _say(f"apply fp64 N={N} general_tiles={general} linear={linear}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 5)
    op = _op(N, order, general_tiles=general, linear_blocks=linear)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    for g, r in zip(vout.download(), o.apply(u, v, h)):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("order,N", [(21, 16), (42, 16), (42, 33)])
def test_unfused_path_matches_oracle(oracle_mod, order, N):
    """The multi-kernel diagnostic path (halo gathers + point kernels)."""
    _say(f"apply unfused order={order} N={N}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 77)
    op = _op(N, order, unfused=True)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    for g, r in zip(vout.download(), o.apply(u, v, h)):
        assert _rel(g, r) < 1e-12
    st = op.create_stvec().upload(u, v, h)
    cs.create_timescheme("rk4").step(st, op, 300.0, 3)
    for g, r in zip(st.download(), o.rk4(u, v, h, 300.0, 3)):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rk4_config1_fp64_matches_oracle(oracle_mod, seed):
    """N=32, sbp21, FP64, a=6.371e6, H0=1000, g=9.81, f=2 Omega sin(phi),
    h = 10 m exp(-(d/500 km)^2) at the seed's centre, u=v=0, dt=600 s, 100
    RK4 steps."""
    lam0, phi0 = CENTRES[seed % 3]
    _say(f"rk4 config-1 seed={seed} centre=({lam0:.4f},{phi0:.4f})")
    o = oracle_mod.Oracle(32, 21)
    h0 = o.gaussian_hill(lam0, phi0, 5e5 / o.radius, 10.0)
    u0, v0 = o.zeros(MESH_U), o.zeros(MESH_V)
    ref = o.rk4(u0, v0, h0, 600.0, 100)
    op = _op(32, 21)
    st = op.create_stvec().upload(u0, v0, h0)
    cs.create_timescheme("rk4").step(st, op, 600.0, 100)
    got = st.download()
    for g, r in zip(got, ref):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("order,N", [(21, 24), (42, 24), (42, 40), (42, 80)])
def test_rk4_fp64_random_start_matches_oracle(oracle_mod, order, N):
    _say(f"rk4 random start order={order} N={N}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 7)
    u, v = 10 * u, 10 * v
    ref = o.rk4(u, v, h, 300.0, 20)
    op = _op(N, order)
    st = op.create_stvec().upload(u, v, h)
    cs.create_timescheme("rk4").step(st, op, 300.0, 20)
    for g, r in zip(st.download(), ref):
        assert _rel(g, r) < 1e-12


@pytest.mark.parametrize("order,N", [(21, 32), (42, 32), (42, 80)])
def test_apply_fp32_close_to_oracle(oracle_mod, order, N):
    """FP32 stencils with FP64-built coefficients: one apply on smooth input
    agrees with FP64 to 1e-5 relative (max error over max magnitude; dh
    involves cancellation in the divergence)."""
    _say(f"apply fp32 order={order}")
    o = oracle_mod.Oracle(N, order)
    h = sum(o.gaussian_hill(lam, phi, 0.4, 10.0) for lam, phi in CENTRES)
    ut = o.covariant_from_cart(*(0 * o.zeros(MESH_U) + c for c in (1.0, 2.0, 0.5)), MESH_U)
    vt = o.covariant_from_cart(*(0 * o.zeros(MESH_V) + c for c in (1.0, 2.0, 0.5)), MESH_V)
    ref = o.apply(ut, vt, h)
    op = _op(N, order, cs.FP32)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(ut, vt, h)
    op.apply(vout, vin)
    for g, r in zip(vout.download(), ref):
        assert _rel(g, r) < 1e-5


def test_stvec_assign_update(oracle_mod):
    """stvec_t%assign / %update semantics (grid_field_mod.f90 L460-534)."""
    N = 12
    o = oracle_mod.Oracle(N, 21)
    a = _random_state(o, 11)
    b = _random_state(o, 12)
    op = _op(N, 21)
    A, B, C = op.create_stvec(), op.create_stvec(), op.create_stvec()
    A.upload(*a)
    B.upload(*b)
    C.assign(2.0, A, -3.0, B)
    for g, x, y in zip(C.download(), a, b):
        assert np.abs(g - (2 * x - 3 * y)).max() < 1e-14
    C.update(0.5, A)
    for g, x, y in zip(C.download(), a, b):
        assert np.abs(g - (2 * x - 3 * y + 0.5 * x)).max() < 1e-14
    C.assign(4.0, B)
    for g, y in zip(C.download(), b):
        assert np.abs(g - 4 * y).max() < 1e-15


# (order, N, precision, flags): the fused tile-wise integrals over single-tile
# panels, uneven last tiles and more tiles than the kernel has blocks (N=97
# FP64: 4 x 13 tiles per panel); the unfused path; FP32 states (ut, vt in
# FP32, the sums in FP64: relative bar 1e-6)
INTEGRAL_CASES = [(21, 5, cs.FP64, False), (21, 20, cs.FP64, False), (21, 48, cs.FP64, False),
                  (42, 8, cs.FP64, False), (42, 20, cs.FP64, False), (42, 57, cs.FP64, False),
                  (42, 97, cs.FP64, False), (42, 20, cs.FP64, True), (21, 33, cs.FP64, True),
                  (42, 40, cs.FP32, False), (21, 31, cs.FP32, False)]


@pytest.mark.parametrize("order,N,precision,unfused", INTEGRAL_CASES)
def test_integrals_match_oracle(oracle_mod, order, N, precision, unfused):
    _say(f"integrals order={order} N={N} precision={precision} unfused={unfused}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 21 + N)
    op = _op(N, order, precision, unfused=unfused)
    st = op.create_stvec().upload(u, v, h)
    m, e = op.integrals(st)
    bar = 1e-12 if precision == cs.FP64 else 1e-6
    assert abs(m - o.mass(h)) < bar * np.abs(h).max() * 4 * np.pi
    assert abs(e / o.energy(u, v, h) - 1) < bar
    # a second call on the same state gives the same numbers
    assert op.integrals(st) == (m, e)


def test_errors_are_reported():
    op = _op(8, 21)
    other = _op(8, 21)
    a = op.create_stvec()
    b = other.create_stvec()
    with pytest.raises(cs.SweError, match="another plan"):
        op.apply(a, b)
    with pytest.raises(cs.SweError, match="must differ"):
        op.apply(a, a)


@pytest.mark.parametrize("precision", [cs.FP64, cs.FP32])
def test_full_size_invariants(oracle_mod, precision):
    """N=384, sbp42: the discrete operator conserves mass exactly and the
    SBP energy to roundoff (for continuous h); checked on the device output
    with the oracle's FP64 quadrature weights only (no oracle apply)."""
    N = 384
    _say(f"full-size invariants N={N} precision={precision}")
    o = oracle_mod.Oracle(N, 42)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(o.shape(MESH_U))
    v = rng.standard_normal(o.shape(MESH_V))
    h = sum(o.gaussian_hill(lam, phi, 0.3, 10.0) for lam, phi in CENTRES)
    op = _op(N, 42, precision)
    vin, vout = op.create_stvec(), op.create_stvec()
    vin.upload(u, v, h)
    op.apply(vout, vin)
    du, dv, dh = vout.download()
    wu, wv, wh = o.weights(MESH_U), o.weights(MESH_V), o.weights(MESH_XY)
    tol = 1e-12 if precision == cs.FP64 else 2e-5
    assert abs(np.sum(dh * wh)) < tol * np.sqrt(np.sum(dh * dh * wh) * np.sum(wh))
    # energy rate with the device's own co2contra: E = g/2<h,h> + H0/2<u,Cu>;
    # dE = g<h,dh> + H0<Cu, du>.  Cu from the oracle (FP64, same inputs).
    ut, vt = o.co2contra(u, v)
    dE = o.grav * np.sum(h * dh * wh) + o.h_mean * (np.sum(ut * du * wu) + np.sum(vt * dv * wv))
    scale = (o.grav * np.sqrt(np.sum(h * h * wh) * np.sum(dh * dh * wh))
             + o.h_mean * np.sqrt((np.sum(ut * ut * wu) + np.sum(vt * vt * wv))
                                  * (np.sum(du * du * wu) + np.sum(dv * dv * wv))))
    assert abs(dE / scale) < (1e-12 if precision == cs.FP64 else 1e-5)


def _continuous_h(N, order, h):
    """Make the duplicated edge values of h bitwise equal (copy from the
    lowest-numbered panel), as the reference's exchanged halos hold them."""
    m = cs.halo_map(N, order, 0)
    h = h.copy()
    for _ in range(3):
        for tp, ti, tj, sp, si, sj in zip(m["tp"], m["ti"], m["tj"], m["sp"], m["si"], m["sj"]):
            ti = 1 if ti == 0 else (N + 1 if ti == N + 2 else ti)
            tj = 1 if tj == 0 else (N + 1 if tj == N + 2 else tj)
            if sp < tp:
                h[tp, tj - 1, ti - 1] = h[sp, sj - 1, si - 1]
            else:
                h[sp, sj - 1, si - 1] = h[tp, tj - 1, ti - 1]
    return h


def _dup_mismatch(N, order, h):
    m = cs.halo_map(N, order, 0)
    worst = 0.0
    for tp, ti, tj, sp, si, sj in zip(m["tp"], m["ti"], m["tj"], m["sp"], m["si"], m["sj"]):
        ti = 1 if ti == 0 else (N + 1 if ti == N + 2 else ti)
        tj = 1 if tj == 0 else (N + 1 if tj == N + 2 else tj)
        worst = max(worst, abs(h[tp, tj - 1, ti - 1] - h[sp, sj - 1, si - 1]))
    return worst


@pytest.mark.parametrize("precision", [cs.FP64, cs.FP32])
@pytest.mark.parametrize("order,N", [(21, 33), (42, 48)])
def test_duplicated_h_stays_bitwise_continuous(oracle_mod, precision, order, N):
    """Every panel sharing an edge point computes the same projected h
    tendency from the same numbers, so duplicated h values stay identical
    (bitwise) through RK4 steps, in FP64 and FP32."""
    _say(f"continuity precision={precision} order={order} N={N}")
    o = oracle_mod.Oracle(N, order)
    u, v, _ = _random_state(o, 9)
    h = _continuous_h(N, order, o.gaussian_hill(0.8, 0.3, 0.4, 10.0))
    assert _dup_mismatch(N, order, h) == 0.0
    op = _op(N, order, precision)
    st = op.create_stvec().upload(u * 1e-3, v * 1e-3, h)
    cs.create_timescheme("rk4").step(st, op, 600.0, 20)
    assert _dup_mismatch(N, order, st.download()[2]) == 0.0


@pytest.mark.parametrize("precision", [cs.FP64, cs.FP32])
def test_rk4_graph_replays_match_direct_launches(oracle_mod, precision):
    """swe_rk4_steps replays its step loop from a captured graph for runs of
    64 steps or more: three 64-step calls (graph replays, with downloads and
    other work between them) give bitwise what 8-step calls (direct
    launches) give, after every call; FP64 also matches the oracle."""
    N, order, dt = 40, 42, 900.0
    _say(f"rk4 graph replay precision={precision}")
    o = oracle_mod.Oracle(N, order)
    s0 = _random_state(o, 77)
    op = _op(N, order, precision)
    rk4 = cs.create_timescheme("rk4")
    a = op.create_stvec().upload(*s0)
    b = op.create_stvec().upload(*s0)
    for rep in range(3):
        rk4.step(a, op, dt, 64)
        for _ in range(8):
            rk4.step(b, op, dt, 8)
        other = op.create_stvec().upload(*s0)  # eager work between replays
        op.apply(op.create_stvec(), other)
        for g, r in zip(a.download(), b.download()):
            assert np.array_equal(g, r), rep
    if precision == cs.FP64:
        ref = o.rk4(*s0, dt, 192)
        for g, r in zip(a.download(), ref):
            assert _rel(g, r) < 1e-12


def test_compensated_rk4_fp64_matches_oracle(oracle_mod):
    """Compensated summation of the step update keeps FP64 at the oracle's
    trajectory (1e-12), through the graph-replayed loop too."""
    N, order, dt = 24, 42, 1200.0
    _say("compensated rk4 fp64")
    o = oracle_mod.Oracle(N, order)
    s0 = _random_state(o, 31)
    op = _op(N, order, cs.FP64, compensated=True)
    st = op.create_stvec().upload(*s0)
    cs.create_timescheme("rk4").step(st, op, dt, 70)
    for g, r in zip(st.download(), o.rk4(*s0, dt, 70)):
        assert _rel(g, r) < 1e-12


def test_compensated_state_semantics():
    """assign gives a compensated state a fresh value (no stale
    compensation): stepping it equals stepping a freshly uploaded copy,
    bitwise; upload resets the compensation too."""
    N, order, dt = 20, 21, 900.0
    _say("compensated assign/upload semantics")
    rng = np.random.default_rng(3)
    s0 = (rng.standard_normal((6, N + 1, N)), rng.standard_normal((6, N, N + 1)),
          rng.standard_normal((6, N + 1, N + 1)))
    op = _op(N, order, cs.FP32, compensated=True)
    rk4 = cs.create_timescheme("rk4")
    a = op.create_stvec().upload(*s0)
    rk4.step(a, op, dt, 30)             # a now carries a compensation
    b = op.create_stvec().upload(*s0)
    a.assign(1.0, b)                     # fresh value again
    rk4.step(a, op, dt, 30)
    c = op.create_stvec().upload(*s0)
    rk4.step(c, op, dt, 30)
    for g, r in zip(a.download(), c.download()):
        assert np.array_equal(g, r)
    a.upload(*s0)
    rk4.step(a, op, dt, 30)
    for g, r in zip(a.download(), c.download()):
        assert np.array_equal(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [cs.FP64, cs.FP32])
@pytest.mark.parametrize("N,x0,y0,linear", [(96, 8, 0, False), (96, 24, 0, True), (100, 13, 9, False),
                                            (384, 20, 0, False)])
def test_tile_splits_agree(precision, N, x0, y0, linear):
    """Any first-tile widths a plan accepts, and either block order (the
    autotuner's choices), give the library's split's RK4 steps to rounding
    (a point moves between the inner and the edge path, which round
    differently): 1e-12 relative in FP64, 1e-5 in FP32."""
    _say(f"tile split N={N} x0={x0} y0={y0} linear={linear} precision={precision}")
    rng = np.random.default_rng(N)
    state = (rng.standard_normal((6, N + 1, N)), rng.standard_normal((6, N, N + 1)),
             rng.standard_normal((6, N + 1, N + 1)))
    out = []
    for kw in ({}, dict(tile_x0=x0, tile_y0=y0, linear_blocks=linear)):
        op = _op(N, 42, precision, tuned=False, **kw)
        st = op.create_stvec().upload(*state)
        cs.create_timescheme("rk4").step(st, op, 30.0, 3)
        out.append(st.download())
    tol = 1e-12 if precision == cs.FP64 else 1e-5
    for a, b in zip(*out):
        assert _rel(a, b) < tol


# N=32 FP64 and N=33 FP32: tilings that end one point short of a side (the
# centres before it read the curl's SAT halo)
DIAG_CASES = [(21, 20, cs.FP64, 0.0), (42, 48, cs.FP64, 7.292e-5), (42, 97, cs.FP64, 0.0), (21, 41, cs.FP32, 7.292e-5),
              (42, 96, cs.FP32, 0.0), (21, 32, cs.FP64, 0.0), (42, 33, cs.FP32, 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("order,N,precision,omega", DIAG_CASES)
def test_integrals_diag_matches_oracle(oracle_mod, order, N, precision, omega):
    """swe_integrals_diag (print_integrals_diag's hmin, hmax, mass, TE, KE, PE,
    enstrophy) against the oracle on a random state over a deep layer: FP64
    1e-12 relative, FP32 2e-6; omega = 0 leaves the enstrophy to the curl
    alone (its SAT halos included)."""
    _check_diag(oracle_mod, order, N, precision, omega, False)


# the multi-kernel diagnostics (unfused plans, and every sbp63 plan): N = 12,
# 13 put the sbp63 closures of both panel ends next to each other
MK_DIAG_CASES = [(21, 20, cs.FP64, 0.0), (42, 33, cs.FP64, 7.292e-5), (63, 12, cs.FP64, 0.0),
                 (63, 13, cs.FP64, 7.292e-5), (63, 40, cs.FP64, 7.292e-5), (63, 32, cs.FP32, 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("order,N,precision,omega", MK_DIAG_CASES)
def test_multi_kernel_diagnostics_match_oracle(oracle_mod, order, N, precision, omega):
    """The same sums on the multi-kernel path: integrals, CFL and the
    integral tendencies."""
    _check_diag(oracle_mod, order, N, precision, omega, True)
    _check_tends(oracle_mod, order, N, precision, omega, True)


def _check_diag(oracle_mod, order, N, precision, omega, unfused):
    _say(f"integrals_diag order={order} N={N} precision={precision} omega={omega} unfused={unfused}")
    o = oracle_mod.Oracle(N, order, radius=6371229.0, omega=omega)
    u, v, h = _random_state(o, 40 + N)
    h = 500.0 + 10.0 * h
    op = _op(N, order, precision, radius=6371229.0, omega=omega, unfused=unfused)
    st = op.create_stvec().upload(u, v, h)
    got = op.integrals_diag(st, 9.80616)
    ref = o.integrals_diag(u, v, h, 9.80616)
    tol = 1e-12 if precision == cs.FP64 else 2e-6
    for k, r in ref.items():
        assert abs(got[k] - r) <= tol * abs(r), (k, got[k], r)
    # print_CFL_diag from the same launch's maxima
    gc = op.cfl_diag(st, 600.0, 9.80616)
    rc = o.cfl_diag(u, v, h, 600.0, 9.80616)
    for k, r in zip(("gravity", "advective_x", "advective_y"), rc):
        assert abs(gc[k] - r) <= tol * abs(r), (k, gc[k], r)


@pytest.mark.gpu
@pytest.mark.parametrize("hw,precision", [(0, cs.FP64), (2, cs.FP64), (8, cs.FP32)])
def test_halo_tile_array_exchange(hw, precision):
    """swe_state_upload_halo / download_halo with the reference's halo'd
    tile arrays: the interior goes in and out as through the packed entry
    points, halo entries are ignored on upload and kept on download."""
    _say(f"halo tile arrays hw={hw} precision={precision}")
    N = 13
    op = _op(N, 42, precision)
    rng = np.random.default_rng(hw)
    shp = {"u": (6, N + 1, N), "v": (6, N, N + 1), "h": (6, N + 1, N + 1)}
    packed = {k: rng.standard_normal(s) for k, s in shp.items()}
    halo = {k: rng.standard_normal((6, s[1] + 2 * hw, s[2] + 2 * hw)) for k, s in shp.items()}
    for k, s in shp.items():
        halo[k][:, hw:hw + s[1], hw:hw + s[2]] = packed[k]
    st = op.create_stvec().upload_halo(hw, halo["u"], halo["v"], halo["h"])
    ref = op.create_stvec().upload(packed["u"], packed["v"], packed["h"])
    for a, b in zip(st.download(), ref.download()):
        assert np.array_equal(a, b)
    out = {k: np.full_like(a, 7.0) for k, a in halo.items()}
    st.download_halo(hw, out["u"], out["v"], out["h"])
    got = ref.download()
    for (k, s), g in zip(shp.items(), got):
        inner = out[k][:, hw:hw + s[1], hw:hw + s[2]]
        assert np.array_equal(inner, g)
        mask = np.ones(out[k].shape, bool)
        mask[:, hw:hw + s[1], hw:hw + s[2]] = False
        assert np.all(out[k][mask] == 7.0)


def _ulps(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


LATLON_CASES = [(42, 12, cs.FP64, None), (21, 33, cs.FP64, None), (42, 96, cs.FP64, None), (42, 48, cs.FP32, None),
                (21, 20, cs.FP64, (37, 50))]


@pytest.mark.gpu
@pytest.mark.parametrize("order,N,precision,grid", LATLON_CASES)
def test_latlon_h_matches_oracle(oracle_mod, order, N, precision, grid):
    """swe_latlon_h (h.dat's record: the ECS-halo'd cubic regrid) against the
    oracle on a random state: float32 outputs of FP64 sums, equal to within
    one unit in the last place (an FP32 plan's h is compared with the
    oracle's regrid of the same float values).  A second call on another
    grid rebuilds the plan's tables."""
    _say(f"latlon_h order={order} N={N} precision={precision} grid={grid}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 80 + N)
    op = _op(N, order, precision)
    st = op.create_stvec().upload(u, v, h)
    nlat, nlon = grid if grid else (2 * N + 1, 4 * N)
    hh = h if precision == cs.FP64 else h.astype(np.float32).astype(np.float64)
    for a, b in ((nlat, nlon), (nlat + 4, nlon + 3), (nlat, nlon)):
        got = op.latlon_h(st, a, b)
        ref = o.latlon_h(hh, a, b)
        assert got.shape == (a, b)
        assert _ulps(got, ref).max() <= 1, (a, b, _ulps(got, ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("order,N,precision,grid", LATLON_CASES)
def test_latlon_uv_matches_oracle(oracle_mod, order, N, precision, grid):
    """swe_latlon_uv (u.dat / v.dat: colocation, the vector ECS halo, cubic
    regrid, zonal / meridional turn) against the oracle on random winds:
    the halo's frame change is derived differently on each side, so the
    float outputs agree to 4e-7 of the largest value, not bitwise."""
    _say(f"latlon_uv order={order} N={N} precision={precision} grid={grid}")
    o = oracle_mod.Oracle(N, order)
    u, v, h = _random_state(o, 90 + N)
    op = _op(N, order, precision)
    st = op.create_stvec().upload(u, v, h)
    nlat, nlon = grid if grid else (2 * N + 1, 4 * N)
    if precision == cs.FP32:
        u, v = (a.astype(np.float32).astype(np.float64) for a in (u, v))
    for a, b in ((nlat, nlon), (nlat + 2, nlon + 5)):
        gu, gv = op.latlon_uv(st, a, b)
        ru, rv = o.latlon_uv(u, v, a, b)
        mx = max(np.abs(ru).max(), np.abs(rv).max())
        assert np.abs(gu - ru).max() <= 4e-7 * mx and np.abs(gv - rv).max() <= 4e-7 * mx


@pytest.mark.gpu
@pytest.mark.parametrize("order,N,precision,omega", DIAG_CASES)
def test_integral_tends_match_oracle(oracle_mod, order, N, precision, omega):
    """swe_integral_tends (print_integral_tends's TE_tend and Enstrophy_tend
    of the operator's tendency) against the oracle on a random state with
    winds of 10 m/s over a deep layer.  Both results are sums that cancel, so
    the error is measured against the sum of their terms' absolute values:
    FP64 1e-12, FP32 2e-6 of it."""
    _check_tends(oracle_mod, order, N, precision, omega, False)


def _check_tends(oracle_mod, order, N, precision, omega, unfused):
    _say(f"integral_tends order={order} N={N} precision={precision} omega={omega} unfused={unfused}")
    o = oracle_mod.Oracle(N, order, radius=6371229.0, omega=omega)
    u, v, h = _random_state(o, 60 + N)
    u, v, h = 10.0 * u, 10.0 * v, 500.0 + 10.0 * h
    op = _op(N, order, precision, radius=6371229.0, omega=omega, unfused=unfused)
    st = op.create_stvec().upload(u, v, h)
    got = op.integral_tends(st, 9.80616)
    te, en, ste, sen = o.integral_tends(u, v, h, 9.80616, scales=True)
    tol = 1e-12 if precision == cs.FP64 else 2e-6
    assert abs(got["TE_tend"] - te) <= tol * ste, (got, te, ste)
    assert abs(got["Enstrophy_tend"] - en) <= tol * sen, (got, en, sen)
    # the state is left as it was
    uu, vv, hh = st.download()
    assert np.array_equal(hh, h) if precision == cs.FP64 else np.allclose(hh, h, rtol=1e-6)
