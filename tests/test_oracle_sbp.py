"""Pins the oracle's 1-D SBP operators with identities the mathematics fixes.

A wrong closure coefficient, a wrong norm weight or a wrong SAT projection
breaks one of: the summation-by-parts identity, interpolation duality, or
polynomial exactness.  Coefficients follow
differential_operators/sbp_operators/sbp_{diff,interp}_{i2c,c2i}_{21,42}_mod.f90
and sbp_operators_collection_mod.f90 L169-200 (sbp63: sbp_*_63_mod.f90).
"""
import numpy as np
import pytest

R_PROJ = {21: [1.5, -0.5], 42: [15 / 8, -5 / 4, 3 / 8], 63: [35 / 16, -35 / 16, 21 / 16, -5 / 16]}
A_INT = {21: [0.5, 1.0], 42: [7 / 18, 9 / 8, 1.0, 71 / 72, 1.0],
         63: [95 / 288, 317 / 240, 23 / 30, 793 / 720, 157 / 160]}
A_CEN = {21: [1.0, 1.0], 42: [13 / 12, 7 / 8, 25 / 24, 1.0],
         63: [325363 / 276480, 144001 / 276480, 43195 / 27648, 86857 / 138240, 312623 / 276480, 271229 / 276480]}


def _mat(m, order, op, n):
    H = m.LINE_HALO
    nin = n + 1 if op in (m.OP_D_I2C, m.OP_I_I2C) else n
    cols = []
    for k in range(nin):
        f = np.zeros(nin + 2 * H)
        f[H + k] = 1.0
        cols.append(m.line_op(order, op, n, f, with_sat=False))
    return np.array(cols).T


def _weights(order, n):
    wi, wc = np.ones(n + 1), np.ones(n)
    ai, ac = A_INT[order], A_CEN[order]
    wi[: len(ai)] = ai
    wi[-len(ai):] = ai[::-1]
    wc[: len(ac)] = ac
    wc[-len(ac):] = ac[::-1]
    return wi, wc


@pytest.mark.parametrize("order", [21, 42, 63])
@pytest.mark.parametrize("n", [8, 13, 24])
def test_summation_by_parts(oracle_mod, order, n):
    if order == 63 and n < 12:
        pytest.skip("sbp63 closures need 12 centres")
    m = oracle_mod
    Di2c = _mat(m, order, m.OP_D_I2C, n)
    Dc2i = _mat(m, order, m.OP_D_C2I, n)
    wi, wc = _weights(order, n)
    B = np.diag(wc) @ Di2c + (np.diag(wi) @ Dc2i).T
    r = np.zeros(n)
    r[: len(R_PROJ[order])] = R_PROJ[order]
    expect = np.zeros((n, n + 1))
    expect[:, 0] = -r
    expect[:, -1] = r[::-1]
    assert np.abs(B - expect).max() < 1e-14


@pytest.mark.parametrize("order", [21, 42, 63])
@pytest.mark.parametrize("n", [8, 17])
def test_interpolation_duality(oracle_mod, order, n):
    if order == 63 and n < 12:
        pytest.skip("sbp63 closures need 12 centres")
    m = oracle_mod
    Ii2c = _mat(m, order, m.OP_I_I2C, n)
    Ic2i = _mat(m, order, m.OP_I_C2I, n)
    wi, wc = _weights(order, n)
    assert np.abs(np.diag(wc) @ Ii2c - (np.diag(wi) @ Ic2i).T).max() < 1e-14
    # interpolation of a constant is exact everywhere
    assert np.allclose(Ii2c.sum(1), 1.0, atol=1e-14)
    assert np.allclose(Ic2i.sum(1), 1.0, atol=1e-14)


@pytest.mark.parametrize("order,interior,boundary,interp_boundary",
                         [(21, 2, 1, 0), (42, 4, 2, 1), (63, 6, 3, 2)])
def test_polynomial_exactness(oracle_mod, order, interior, boundary, interp_boundary):
    """Interior rows of the staggered derivatives are exact for polynomials up
    to degree 2 (sbp21) / 4 (sbp42) / 6 (sbp63); closure rows up to degree 1
    / 2 / 3.  The 4th-order boundary interpolation is exact to degree 1 only:
    its free parameters c13, c14 minimise (not zero) the degree-2 error
    (sbp_interp_i2c_42_mod.f90 L12-19); the 6th-order one to degree 2 (its
    six free parameters minimise the degree-3 and 4 errors,
    sbp_interp_i2c_63_mod.f90 L9-20); the 2nd-order one copies the end
    value."""
    m = oracle_mod
    n = 24
    xi = np.arange(n + 1, dtype=float)  # interfaces
    xc = xi[:-1] + 0.5                    # centres
    Di2c = _mat(m, order, m.OP_D_I2C, n)
    Dc2i = _mat(m, order, m.OP_D_C2I, n)
    Ii2c = _mat(m, order, m.OP_I_I2C, n)
    Ic2i = _mat(m, order, m.OP_I_C2I, n)
    lo, hi = {21: (1, n - 1), 42: (4, n - 4), 63: (6, n - 6)}[order]
    for deg in range(interior + 1):
        # interior rows exact to degree `interior` for derivatives
        d_i2c = Di2c @ xi**deg
        exact_c = deg * xc ** (deg - 1) if deg else 0 * xc
        assert np.abs(d_i2c[lo:hi] - exact_c[lo:hi]).max() < 1e-9 * max(1, n**deg)
        d_c2i = Dc2i @ xc**deg
        exact_i = deg * xi ** (deg - 1) if deg else 0 * xi
        assert np.abs(d_c2i[lo + 1:hi] - exact_i[lo + 1:hi]).max() < 1e-9 * max(1, n**deg)
    for deg in range(boundary + 1):
        d_i2c = Di2c @ xi**deg
        exact_c = deg * xc ** (deg - 1) if deg else 0 * xc
        assert np.abs(d_i2c - exact_c).max() < 1e-9 * max(1, n**deg)
        d_c2i = Dc2i @ xc**deg
        exact_i = deg * xi ** (deg - 1) if deg else 0 * xi
        assert np.abs(d_c2i - exact_i).max() < 1e-9 * max(1, n**deg)
    for deg in range(interp_boundary + 1):
        assert np.abs(Ii2c @ xi**deg - xc**deg).max() < 1e-9 * max(1, n**deg)
        assert np.abs(Ic2i @ xc**deg - xi**deg).max() < 1e-9 * max(1, n**deg)
    # interior interpolation rows are exact to degree interior-1
    for deg in range(interior):
        assert np.abs((Ii2c @ xi**deg - xc**deg)[lo:hi]).max() < 1e-9 * max(1, n**deg)
        assert np.abs((Ic2i @ xc**deg - xi**deg)[lo + 1:hi]).max() < 1e-9 * max(1, n**deg)
