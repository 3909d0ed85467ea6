"""Initial conditions for the linear SWE workloads (host side, numpy, FP64).

Mesh coordinates and covariant basis vectors follow the equiangular
cubed-sphere metric (metric/metric_2d_ecs_mod.f90 L126-202, L338-411) with
the panel frames of topology/cubed_sphere_topology_mod.f90 L29-55; fields
are returned in the packed layout of include/ch_swe.h.

* gaussian_hill: test_fields_mod.f90 generate_gaussian_hill_scalar_field
  L491-514 (amplitude scaled), centres of gaussian_hill_linear_test_mod.f90
  L17-19 and linear_gauss_run.sh L41-52.
* sh_state: seeded smooth field for the large-solve and ensemble workloads: a
  real spherical-harmonic sum up to degree lmax with coefficients
  N(0,1) * l^-2, scaled to a given SBP-weighted RMS height; winds from the
  stream function psi = (g/f0) h cos^2(phi) (f-plane geostrophy at 45 deg,
  tapered by cos^2 phi so the field is smooth at the poles), V = k x grad psi.
"""
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

# panel frames: corner r of point (1,1), directions of increasing i / j,
# inward normal n (cubed_sphere_topology_mod.f90 L29-55)
_EX = np.array([[-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 0, 1], [1, 0, 0], [1, 0, 0]], float)
_EY = np.array([[0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]], float)
_NN = np.array([[0, 0, 1], [-1, 0, 0], [0, 0, -1], [1, 0, 0], [0, 1, 0], [0, -1, 0]], float)

MESHES = ("u", "v", "h")
CENTRES = {
    "interior": (np.pi, 0.0),
    "corner": (np.pi / 4, np.arcsin(1 / np.sqrt(3))),
    "near_corner": (0.27 * np.pi, np.arcsin(1 / np.sqrt(3)) - 0.02 * np.pi),
}


def shape(N, mesh):
    return {"u": (6, N + 1, N), "v": (6, N, N + 1), "h": (6, N + 1, N + 1)}[mesh]


def _angles(N, mesh):
    xt, yt = {"u": (1, 0), "v": (0, 1), "h": (0, 0)}[mesh]
    nx, ny = (N if xt else N + 1), (N if yt else N + 1)
    hx = 0.5 * np.pi / N
    al = -0.25 * np.pi + (np.arange(nx) + 0.5 * xt) * hx
    be = -0.25 * np.pi + (np.arange(ny) + 0.5 * yt) * hx
    return np.meshgrid(al, be, indexing="xy")  # [ny, nx]


def mesh_geometry(N, mesh):
    """Unit-sphere position r and covariant basis a1 = dr/dalpha,
    a2 = dr/dbeta at every point of a mesh: arrays [6, ny, nx, 3]."""
    A, B = _angles(N, mesh)
    X, Y = np.tan(A), np.tan(B)
    s2 = 1 + X**2 + Y**2
    s = np.sqrt(s2)
    s3 = s2 * s
    cx, cy = 1 + X**2, 1 + Y**2
    proto_r = np.stack([X / s, Y / s, 1 / s], -1)
    proto_a1 = np.stack([cx * cy / s3, -X * Y * cx / s3, -X * cx / s3], -1)
    proto_a2 = np.stack([-X * Y * cy / s3, cx * cy / s3, -Y * cy / s3], -1)
    out = []
    for v in (proto_r, proto_a1, proto_a2):
        # prototype face -> panel: ex*x + ey*y - n*z
        out.append(np.stack([v[..., 0:1] * _EX[p] + v[..., 1:2] * _EY[p] - v[..., 2:3] * _NN[p]
                             for p in range(6)], 0))
    return tuple(out)


def sbp_weights(N, mesh, order):
    """J hx hy Ax Ay (quadrature/sbp_quadrature_mod.f90 L22-41)."""
    A, B = _angles(N, mesh)
    X, Y = np.tan(A), np.tan(B)
    J = (1 + X**2) * (1 + Y**2) / (1 + X**2 + Y**2) ** 1.5
    hx = 0.5 * np.pi / N
    xt, yt = {"u": (1, 0), "v": (0, 1), "h": (0, 0)}[mesh]
    if order == 21:
        ai, ac = [0.5, 1.0], [1.0, 1.0]
    elif order == 42:
        ai, ac = [7 / 18, 9 / 8, 1.0, 71 / 72, 1.0], [13 / 12, 7 / 8, 25 / 24, 1.0]
    else:
        ai = [95 / 288, 317 / 240, 23 / 30, 793 / 720, 157 / 160]
        ac = [325363 / 276480, 144001 / 276480, 43195 / 27648, 86857 / 138240, 312623 / 276480, 271229 / 276480]

    def w1(n, centre):
        e = ac if centre else ai
        w = np.ones(n)
        w[: len(e)] = e
        w[n - len(e):] = e[::-1]
        return w

    wx = w1(A.shape[1], xt)
    wy = w1(A.shape[0], yt)
    return np.broadcast_to(J * hx * hx * wy[:, None] * wx[None, :], (6,) + A.shape).copy()


def _latlon(r):
    lam = np.arctan2(r[..., 1], r[..., 0])
    phi = np.arcsin(np.clip(r[..., 2], -1, 1))
    return lam, phi


def gaussian_hill(N, lam0, phi0, d0, amp=1.0):
    """h = amp * exp(-(dist/d0)^2), dist = great-circle angle to the centre."""
    r, _, _ = mesh_geometry(N, "h")
    c = np.array([np.cos(phi0) * np.cos(lam0), np.cos(phi0) * np.sin(lam0), np.sin(phi0)])
    d = np.arccos(np.clip(r @ c, -1.0, 1.0))
    return amp * np.exp(-((d / d0) ** 2))


def _legendre_sum(x, coef_c, coef_s, lmax):
    """A_m(x) = sum_l c_lm Pbar_l^m(x), B_m likewise, orthonormal real SH
    normalisation (4 pi); returns arrays [lmax+1, npts]."""
    npts = x.size
    Am = np.zeros((lmax + 1, npts))
    Bm = np.zeros((lmax + 1, npts))
    sx = np.sqrt(np.maximum(0.0, 1 - x * x))
    pmm = np.full(npts, np.sqrt(1 / (4 * np.pi)))
    for m in range(lmax + 1):
        if m > 0:
            pmm = pmm * np.sqrt((2 * m + 1) / (2 * m)) * sx
        fac = np.sqrt(2.0) if m > 0 else 1.0
        p_prev2, p_prev = None, pmm
        Am[m] += coef_c[m, m] * p_prev * fac
        Bm[m] += coef_s[m, m] * p_prev * fac
        if m + 1 <= lmax:
            p = np.sqrt(2 * m + 3) * x * pmm
            Am[m] += coef_c[m + 1, m] * p * fac
            Bm[m] += coef_s[m + 1, m] * p * fac
            p_prev2, p_prev = pmm, p
            for l in range(m + 2, lmax + 1):
                a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
                b = np.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                p = a * (x * p_prev - b * p_prev2)
                Am[m] += coef_c[l, m] * p * fac
                Bm[m] += coef_s[l, m] * p * fac
                p_prev2, p_prev = p_prev, p
    return Am, Bm


def _sh_eval(lam, phi, cc, cs_, lmax, with_grad=False, dphi=1e-5):
    lam, phi = lam.ravel(), phi.ravel()
    m = np.arange(lmax + 1)[:, None]
    cosm, sinm = np.cos(m * lam), np.sin(m * lam)
    Am, Bm = _legendre_sum(np.sin(phi), cc, cs_, lmax)
    h = np.sum(Am * cosm + Bm * sinm, 0)
    if not with_grad:
        return h
    h_lam = np.sum(m * (-Am * sinm + Bm * cosm), 0)
    Ap, Bp = _legendre_sum(np.sin(phi + dphi), cc, cs_, lmax)
    An, Bn = _legendre_sum(np.sin(phi - dphi), cc, cs_, lmax)
    h_phi = np.sum((Ap - An) * cosm + (Bp - Bn) * sinm, 0) / (2 * dphi)
    return h, h_lam, h_phi


def _sh_job(args):
    N, mesh, cc, cs_, lmax, grad = args
    r, _, _ = mesh_geometry(N, mesh)
    lam, phi = _latlon(r)
    res = _sh_eval(lam, phi, cc, cs_, lmax, grad)
    return tuple(x.reshape(lam.shape) for x in res) if grad else res.reshape(lam.shape)


def sh_state(N, order=42, lmax=32, seed=0, rms=20.0, grav=9.81, omega=7.292e-5, radius=6.371e6,
             workers=None):
    """(u, v, h): seeded degree-lmax spherical-harmonic height with `rms`
    metres SBP-weighted RMS and balanced, pole-tapered winds.  The three
    meshes are evaluated in worker processes (call before any device work)."""
    rng = np.random.default_rng(seed)
    cc = np.zeros((lmax + 1, lmax + 1))
    cs_ = np.zeros((lmax + 1, lmax + 1))
    for l in range(1, lmax + 1):
        cc[l, : l + 1] = rng.standard_normal(l + 1) / l**2
        cs_[l, 1: l + 1] = rng.standard_normal(l) / l**2
    jobs = [(N, "h", cc, cs_, lmax, False), (N, "u", cc, cs_, lmax, True), (N, "v", cc, cs_, lmax, True)]
    if workers is None:
        workers = 3 if N >= 96 else 1
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            res = list(ex.map(_sh_job, jobs))
    else:
        res = [_sh_job(j) for j in jobs]
    h = res[0]
    w = sbp_weights(N, "h", order)
    scale = rms / np.sqrt(np.sum(w * h * h) / np.sum(w))
    f0 = 2 * omega * np.sin(np.pi / 4)
    out = {"h": h * scale}
    for mesh, comp, (hv, hl, hp) in (("u", 1, res[1]), ("v", 2, res[2])):
        r, a1, a2 = mesh_geometry(N, mesh)
        lam, phi = _latlon(r)
        hv, hl, hp = hv * scale, hl * scale, hp * scale
        c = np.cos(phi)
        k = grav / f0
        psi_lam_over_cos = k * hl * c          # (1/cos) d/dlam (h cos^2)
        psi_phi = k * (hp * c * c - 2 * hv * c * np.sin(phi))
        e_lam = np.stack([-np.sin(lam), np.cos(lam), np.zeros_like(lam)], -1)
        e_phi = np.stack([-np.sin(phi) * np.cos(lam), -np.sin(phi) * np.sin(lam), c], -1)
        V = (psi_lam_over_cos[..., None] * e_phi - psi_phi[..., None] * e_lam) / radius
        out[mesh] = np.sum(V * (a1 if comp == 1 else a2), -1)
    return out["u"], out["v"], out["h"]


def ensemble_states(N, order, members, lmax=16, seed=0, basis=8, rms=20.0, radius=6.371e6, workers=None):
    """Seeded ensemble initial states: member m gets a degree-lmax
    spherical-harmonic field with coefficients ~ N(0,1) l^-2 (a unit-norm
    random combination of `basis` sh_state fields, which has the same
    coefficient distribution) with its balanced winds, plus a Gaussian bump of
    amplitude U(1, 20) m, width U(2e5, 1e6) m and a centre uniform on the
    sphere (no wind).  Returns a list of (u, v, h).  workers: as for
    sh_state (1 in a process that has already used the GPU)."""
    base = [sh_state(N, order, lmax=lmax, seed=seed * 1000 + k, rms=rms, workers=workers) for k in range(basis)]
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((members, basis))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    amp = rng.uniform(1.0, 20.0, members)
    width = rng.uniform(2e5, 1e6, members) / radius
    cen = rng.standard_normal((members, 3))
    cen /= np.linalg.norm(cen, axis=1, keepdims=True)
    r, _, _ = mesh_geometry(N, "h")
    out = []
    for m in range(members):
        u = sum(w[m, k] * base[k][0] for k in range(basis))
        v = sum(w[m, k] * base[k][1] for k in range(basis))
        h = sum(w[m, k] * base[k][2] for k in range(basis))
        d = np.arccos(np.clip(r @ cen[m], -1.0, 1.0))
        out.append((u, v, h + amp[m] * np.exp(-((d / width[m]) ** 2))))
    return out


def cfl_dt(N, courant=0.5, grav=9.81, h_mean=1000.0, radius=6.371e6):
    """dt with c dt / (a hx) = courant, c = sqrt(g H0) (the CFL definition of
    swm_output_diag_mod.f90 L138-149 with h -> H0)."""
    return courant * radius * (0.5 * np.pi / N) / np.sqrt(grav * h_mean)
