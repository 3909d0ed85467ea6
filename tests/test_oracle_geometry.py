"""Pins the oracle's geometry, quadrature, halo maps and edge-sync weights.

Checks come from the geometry itself, not from the oracle's formulas:
* J and Qi against the cross product and inverse Gram matrix of the basis;
* the basis against finite differences of the mesh coordinates;
* SBP quadrature against the sphere's area 4*pi;
* every halo entry against the 3-D position of the point it copies
  (xy duplicates coincide; normal-flux entries sit on the neighbour's grid
  line through the shared edge point, at the right depth);
* vector halo signs against the basis vectors at the shared edge points;
* the covariant sync weights against the real neighbouring panel's
  contravariant basis (the oracle computes them on the prototype face).
"""
import numpy as np
import pytest

from oracle.swe_oracle import MESH_U, MESH_V, MESH_XY


def _o(m, N, order=21):
    return m.Oracle(N, order)


@pytest.mark.parametrize("mesh", [MESH_XY, MESH_U, MESH_V])
def test_metric_matches_basis(oracle_mod, mesh):
    o = _o(oracle_mod, 12)
    g = o.metric(mesh)
    a1, a2 = o.basis(mesh)
    J = np.linalg.norm(np.cross(a1, a2), axis=1)
    assert np.abs(J - g["J"].ravel()).max() < 1e-14
    G11 = (a1 * a1).sum(1)
    G12 = (a1 * a2).sum(1)
    G22 = (a2 * a2).sum(1)
    det = G11 * G22 - G12**2
    assert np.abs(G22 / det - g["Qi11"].ravel()).max() < 1e-12
    assert np.abs(-G12 / det - g["Qi12"].ravel()).max() < 1e-12
    assert np.abs(G11 / det - g["Qi22"].ravel()).max() < 1e-12
    r = np.stack([g["rx"].ravel(), g["ry"].ravel(), g["rz"].ravel()], 1)
    assert np.abs(np.linalg.norm(r, axis=1) - 1).max() < 1e-15
    # the basis is tangent to the sphere
    assert np.abs((a1 * r).sum(1)).max() < 1e-15
    assert np.abs((a2 * r).sum(1)).max() < 1e-15


def test_basis_matches_finite_differences(oracle_mod):
    N = 64
    o = _o(oracle_mod, N)
    g = o.metric(MESH_XY)
    a1, a2 = o.basis(MESH_XY)
    r = np.stack([g["rx"], g["ry"], g["rz"]], -1)  # [6, j, i, 3]
    a1 = a1.reshape(6, N + 1, N + 1, 3)
    a2 = a2.reshape(6, N + 1, N + 1, 3)
    hx = 0.5 * np.pi / N
    d1 = (r[:, :, 2:] - r[:, :, :-2]) / (2 * hx)
    d2 = (r[:, 2:, :] - r[:, :-2, :]) / (2 * hx)
    assert np.abs(d1 - a1[:, :, 1:-1]).max() < 5 * hx**2
    assert np.abs(d2 - a2[:, 1:-1, :]).max() < 5 * hx**2


@pytest.mark.parametrize("order", [21, 42])
def test_quadrature_area(oracle_mod, order):
    errs = []
    for N in (16, 32):
        o = oracle_mod.Oracle(N, order)
        errs.append([abs(o.weights(m).sum() / (4 * np.pi) - 1) for m in (MESH_XY, MESH_U, MESH_V)])
    errs = np.array(errs)
    assert errs[1].max() < (1e-3 if order == 21 else 1e-5)
    assert np.all(errs[1] < errs[0] / 3)


def _points(o, mesh):
    g = o.metric(mesh)
    return np.stack([g["rx"], g["ry"], g["rz"]], -1)


def test_xy_halo_duplicates_coincide(oracle_mod):
    N = 10
    o = _o(oracle_mod, N)
    r = _points(o, MESH_XY)
    hm = o.halo_map(0)
    assert len(hm["tp"]) == 6 * 4 * (N + 1)
    for k in range(len(hm["tp"])):
        tp, ti, tj = hm["tp"][k], hm["ti"][k], hm["tj"][k]
        # the halo point duplicates the target's own edge point
        ei = min(max(ti, 1), N + 1)
        ej = min(max(tj, 1), N + 1)
        src = r[hm["sp"][k], hm["sj"][k] - 1, hm["si"][k] - 1]
        own = r[tp, ej - 1, ei - 1]
        assert np.abs(src - own).max() < 1e-15
        assert hm["sp"][k] != tp


def test_ch_halo_geometry_and_signs(oracle_mod):
    """Normal-flux halo entries (div SAT): the source lies at depth d into the
    neighbour, on the neighbour's grid line that starts at the target's edge
    point, carries the neighbour's component that crosses the edge, and its
    sign makes it point like the target's unfolded axis."""
    N = 10
    o = _o(oracle_mod, N)
    rxy = _points(o, MESH_XY)
    a1xy, a2xy = (a.reshape(6, N + 1, N + 1, 3) for a in o.basis(MESH_XY))
    a1u = o.basis(MESH_U)[0].reshape(6, N + 1, N, 3)
    a2v = o.basis(MESH_V)[1].reshape(6, N, N + 1, 3)
    hm = o.halo_map(1)
    assert len(hm["tp"]) == 6 * 4 * 2 * (N + 1)  # sbp21: depth 2
    for k in range(len(hm["tp"])):
        tp, ti, tj, tc = hm["tp"][k], hm["ti"][k], hm["tj"][k], hm["tc"][k]
        sp, si, sj, sc, sg = hm["sp"][k], hm["si"][k], hm["sj"][k], hm["sc"][k], hm["sign"][k]
        if tc == 0:  # u halo across a left/right edge; tj is an xy row
            ei, ej = (1, tj) if ti < 1 else (N + 1, tj)
            depth = 1 - ti if ti < 1 else ti - N
            outward = (-1 if ti < 1 else 1) * a1xy[tp, ej - 1, ei - 1]
            axis_sign = -1 if ti < 1 else 1
        else:        # v halo across a bottom/top edge; ti is an xy column
            ei, ej = (ti, 1) if tj < 1 else (ti, N + 1)
            depth = 1 - tj if tj < 1 else tj - N
            outward = (-1 if tj < 1 else 1) * a2xy[tp, ej - 1, ei - 1]
            axis_sign = -1 if tj < 1 else 1
        E = rxy[tp, ej - 1, ei - 1]
        if sc == 0:   # source is a u point: x is the crossing (centre) axis
            assert si in (depth, N + 1 - depth)
            edge_pt = rxy[sp, sj - 1, 0 if si == depth else N]
            b = a1u[sp, sj - 1, si - 1]
        else:         # source is a v point: y is the crossing axis
            assert sj in (depth, N + 1 - depth)
            edge_pt = rxy[sp, 0 if sj == depth else N, si - 1]
            b = a2v[sp, sj - 1, si - 1]
        assert np.abs(edge_pt - E).max() < 1e-15
        # the source basis vector, signed, points like the unfolded target
        # axis: outward across right/top edges, inward across left/bottom
        assert sg * (b @ outward) * axis_sign > 0


def test_xy_vector_halo_signs(oracle_mod):
    """At a shared edge point the along-edge basis vectors of both panels are
    equal up to the recorded sign; the across-edge ones agree in direction."""
    N = 8
    o = _o(oracle_mod, N)
    a1, a2 = o.basis(MESH_XY)
    A = np.stack([a1.reshape(6, N + 1, N + 1, 3), a2.reshape(6, N + 1, N + 1, 3)], 0)
    hm = o.halo_map(2)
    for k in range(len(hm["tp"])):
        tp, ti, tj, tc = hm["tp"][k], hm["ti"][k], hm["tj"][k], hm["tc"][k]
        sp, si, sj, sc, sg = hm["sp"][k], hm["si"][k], hm["sj"][k], hm["sc"][k], hm["sign"][k]
        ei, ej = min(max(ti, 1), N + 1), min(max(tj, 1), N + 1)
        tgt = A[tc, tp, ej - 1, ei - 1]
        src = sg * A[sc, sp, sj - 1, si - 1]
        across = (tc == 0 and ti != ei) or (tc == 1 and tj != ej)
        if across:
            assert tgt @ src > 0.3 * np.linalg.norm(tgt) * np.linalg.norm(src)
        else:
            assert np.abs(tgt - src).max() < 1e-14


def test_sync_weights_from_real_neighbour(oracle_mod):
    """qb/qt/ql/qr (covariant sync) = own across-edge covariant vector dotted
    with the neighbour's along-edge contravariant vector, expressed in the
    unfolded frame; recomputed here from the actual neighbouring panels."""
    N = 8
    o = _o(oracle_mod, N)
    q = o.sync_weights()
    a1, a2 = o.basis(MESH_XY)
    g = o.metric(MESH_XY)
    Q11, Q12, Q22 = (g[k].reshape(-1) for k in ("Qi11", "Qi12", "Qi22"))
    b1 = Q11[:, None] * a1 + Q12[:, None] * a2
    b2 = Q12[:, None] * a1 + Q22[:, None] * a2
    A = np.stack([a1, a2]).reshape(2, 6, N + 1, N + 1, 3)
    B = np.stack([b1, b2]).reshape(2, 6, N + 1, N + 1, 3)
    hm = o.halo_map(2)
    for k in range(len(hm["tp"])):
        tp, ti, tj, tc = hm["tp"][k], hm["ti"][k], hm["tj"][k], hm["tc"][k]
        sp, si, sj, sc, sg = hm["sp"][k], hm["si"][k], hm["sj"][k], hm["sc"][k], hm["sign"][k]
        ei, ej = min(max(ti, 1), N + 1), min(max(tj, 1), N + 1)
        along_edge = (tc == 1 and ti != ei) or (tc == 0 and tj != ej)
        if not along_edge:
            continue
        # neighbour's contravariant vector dual to the unfolded along-edge axis
        bn = sg * B[sc, sp, sj - 1, si - 1]
        if ti < 1:
            ref, w = A[0, tp, ej - 1, ei - 1] @ bn, q["ql"][ej - 1]
        elif ti > N + 1:
            ref, w = A[0, tp, ej - 1, ei - 1] @ bn, q["qr"][ej - 1]
        elif tj < 1:
            ref, w = A[1, tp, ej - 1, ei - 1] @ bn, q["qb"][ei - 1]
        else:
            ref, w = A[1, tp, ej - 1, ei - 1] @ bn, q["qt"][ei - 1]
        assert abs(ref - w) < 1e-13, (tp, ti, tj, ref, w)
