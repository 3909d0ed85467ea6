"""ctypes front end of the CPU oracle (oracle/libswe_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker (or the timed CPU baseline),
never as part of the product path.  Packed layouts are documented in
oracle/swe_oracle.h.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libswe_oracle.so")
_lib = None

MESH_XY, MESH_U, MESH_V = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.swo_create.restype = ctypes.c_void_p
    lib.swo_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]
    lib.swo_destroy.argtypes = [ctypes.c_void_p]
    lib.swo_npts.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.swo_npts.restype = ctypes.c_int
    for name, n in (("swo_apply", 6), ("swo_grad", 3), ("swo_co2contra", 4),
                    ("swo_coriolis", 4), ("swo_div", 3)):
        getattr(lib, name).argtypes = [ctypes.c_void_p] + [_dp] * n
    lib.swo_rk4.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]
    lib.swo_metric.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 7
    lib.swo_quad_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.swo_coriolis_fJ2.argtypes = [ctypes.c_void_p, _dp]
    lib.swo_sync_weights.argtypes = [ctypes.c_void_p] + [_dp] * 4
    lib.swo_energy.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.swo_energy.restype = ctypes.c_double
    lib.swo_mass.argtypes = [ctypes.c_void_p, _dp]
    lib.swo_mass.restype = ctypes.c_double
    lib.swo_curl.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.swo_integrals_diag.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp]
    lib.swo_integral_tends.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp]
    lib.swo_solve.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                              _dp, _dp, _dp, _ip, _dp]
    lib.swo_latlon_uv.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.swo_cfl_diag.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp]
    lib.swo_latlon_h.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib.swo_gaussian_hill.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp]
    lib.swo_ts2_linear.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
    lib.swo_covariant_from_cart.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, _dp]
    lib.swo_halo_map.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_ip] * 9
    lib.swo_halo_map.restype = ctypes.c_int
    lib.swo_line_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.swo_basis.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.swo_set_mutation.argtypes = [ctypes.c_int]
    lib.swo_set_num_threads.argtypes = [ctypes.c_int]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


LINE_HALO = 4
OP_D_I2C, OP_D_C2I, OP_I_I2C, OP_I_C2I = 0, 1, 2, 3


def line_op(order, op, n, f, with_sat=True):
    """Apply one 1-D SBP operator.  f is indexed from 1-LINE_HALO (a line of
    n+1+2*LINE_HALO values with index 1 at f[LINE_HALO]); n = centre count."""
    lib = _load()
    f = np.ascontiguousarray(f, dtype=np.float64)
    nout = n if op in (OP_D_I2C, OP_I_I2C) else n + 1
    out = np.zeros(nout)
    lib.swo_line_op(order, op, n, _p(f), _p(out), 1 if with_sat else 0)
    return out


def num_threads():
    """OpenMP threads the oracle's row loops use."""
    return _load().swo_num_threads()


def set_num_threads(n):
    _load().swo_set_num_threads(int(n))


def set_mutation(m):
    """Plant a known mistake in the oracle (0 restores it); see swe_oracle.c."""
    _load().swo_set_mutation(int(m))


class Oracle:
    """FP64 CPU restatement of operator_swm_lin_t%apply and rk4_t%step."""

    def __init__(self, N, order, radius=6.371e6, grav=9.81, h_mean=1000.0,
                 omega=7.292e-5, axis=(0.0, 0.0, 1.0), coriolis_const=False):
        lib = _load()
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        self._h = lib.swo_create(N, order, radius, grav, h_mean, omega, _p(ax),
                                 1 if coriolis_const else 0)
        if not self._h:
            raise ValueError(f"unsupported oracle configuration N={N} order={order}")
        self.N, self.order = N, order
        self.radius, self.grav, self.h_mean, self.omega = radius, grav, h_mean, omega
        self.lib = lib
        self.n_u = lib.swo_npts(self._h, MESH_U)
        self.n_v = lib.swo_npts(self._h, MESH_V)
        self.n_h = lib.swo_npts(self._h, MESH_XY)

    def __del__(self):
        if getattr(self, "_h", None):
            self.lib.swo_destroy(self._h)
            self._h = None

    def npts(self, mesh):
        return {MESH_U: self.n_u, MESH_V: self.n_v, MESH_XY: self.n_h}[mesh]

    def shape(self, mesh):
        N = self.N
        return {MESH_U: (6, N + 1, N), MESH_V: (6, N, N + 1), MESH_XY: (6, N + 1, N + 1)}[mesh]

    def zeros(self, mesh):
        return np.zeros(self.shape(mesh))

    @staticmethod
    def _c(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def apply(self, u, v, h):
        u, v, h = self._c(u), self._c(v), self._c(h)
        du, dv, dh = np.empty_like(u), np.empty_like(v), np.empty_like(h)
        self.lib.swo_apply(self._h, _p(u), _p(v), _p(h), _p(du), _p(dv), _p(dh))
        return du, dv, dh

    def rk4(self, u, v, h, dt, nsteps):
        u, v, h = self._c(u).copy(), self._c(v).copy(), self._c(h).copy()
        self.lib.swo_rk4(self._h, _p(u), _p(v), _p(h), dt, nsteps)
        return u, v, h

    def grad(self, h):
        h = self._c(h)
        gx, gy = self.zeros(MESH_U), self.zeros(MESH_V)
        self.lib.swo_grad(self._h, _p(h), _p(gx), _p(gy))
        return gx, gy

    def co2contra(self, u, v):
        u, v = self._c(u), self._c(v)
        ut, vt = np.empty_like(u), np.empty_like(v)
        self.lib.swo_co2contra(self._h, _p(u), _p(v), _p(ut), _p(vt))
        return ut, vt

    def coriolis(self, ut, vt):
        ut, vt = self._c(ut), self._c(vt)
        cu, cv = np.empty_like(ut), np.empty_like(vt)
        self.lib.swo_coriolis(self._h, _p(ut), _p(vt), _p(cu), _p(cv))
        return cu, cv

    def div(self, ut, vt):
        ut, vt = self._c(ut), self._c(vt)
        d = self.zeros(MESH_XY)
        self.lib.swo_div(self._h, _p(ut), _p(vt), _p(d))
        return d

    def metric(self, mesh):
        out = [self.zeros(mesh) for _ in range(7)]
        self.lib.swo_metric(self._h, mesh, *[_p(a) for a in out])
        return dict(zip(("J", "Qi11", "Qi12", "Qi22", "rx", "ry", "rz"), out))

    def weights(self, mesh):
        w = self.zeros(mesh)
        self.lib.swo_quad_weights(self._h, mesh, _p(w))
        return w

    def fJ2(self):
        f = self.zeros(MESH_XY)
        self.lib.swo_coriolis_fJ2(self._h, _p(f))
        return f

    def sync_weights(self):
        q = [np.zeros(self.N + 1) for _ in range(4)]
        self.lib.swo_sync_weights(self._h, *[_p(a) for a in q])
        return dict(zip(("qb", "qt", "ql", "qr"), q))

    def energy(self, u, v, h):
        u, v, h = self._c(u), self._c(v), self._c(h)
        return self.lib.swo_energy(self._h, _p(u), _p(v), _p(h))

    def mass(self, h):
        h = self._c(h)
        return self.lib.swo_mass(self._h, _p(h))

    def curl(self, u, v):
        """curl_ch_sbp_sat at the cell centres, shape (6, N, N)."""
        u, v = self._c(u), self._c(v)
        out = np.zeros((6, self.N, self.N))
        self.lib.swo_curl(self._h, _p(u), _p(v), _p(out))
        return out

    def integrals_diag(self, u, v, h, grav=9.80616):
        """print_integrals_diag: dict of hmin, hmax, mass, TE, KE, PE, Enstrophy."""
        u, v, h = self._c(u), self._c(v), self._c(h)
        out = np.zeros(7)
        self.lib.swo_integrals_diag(self._h, _p(u), _p(v), _p(h), grav, _p(out))
        return dict(zip(("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"), out.tolist()))

    def integral_tends(self, u, v, h, grav=9.80616, scales=False):
        """print_integral_tends of the operator's tendency: (TE_tend,
        Enstrophy_tend), and with scales=True also the sums of their terms'
        absolute values (the scale of their rounding)."""
        u, v, h = self._c(u), self._c(v), self._c(h)
        out = np.zeros(4)
        self.lib.swo_integral_tends(self._h, _p(u), _p(v), _p(h), grav, _p(out))
        return tuple(float(x) for x in (out if scales else out[:2]))

    def solve(self, u, v, h, dt, rel_tol=1e-10, max_iter=200):
        """operator_swm_lin_t%solve: (u, v, h) = (I - dt L0)^-1 rhs without
        Coriolis, by CG on the Helmholtz problem; returns (u, v, h, iters,
        relres)."""
        u, v, h = self._c(u), self._c(v), self._c(h)
        out = [np.empty_like(u), np.empty_like(v), np.empty_like(h)]
        it, rr = ctypes.c_int(), ctypes.c_double()
        self.lib.swo_solve(self._h, _p(u), _p(v), _p(h), dt, rel_tol, max_iter, *[_p(a) for a in out],
                           ctypes.byref(it), ctypes.byref(rr))
        return out[0], out[1], out[2], it.value, rr.value

    def cfl_diag(self, u, v, h, dt, grav=9.80616):
        """print_CFL_diag: (gravity-wave CFL, advective CFL x, y)."""
        u, v, h = self._c(u), self._c(v), self._c(h)
        out = np.zeros(3)
        self.lib.swo_cfl_diag(self._h, _p(u), _p(v), _p(h), dt, grav, _p(out))
        return tuple(float(x) for x in out)

    def latlon_h(self, h, nlat=None, nlon=None):
        """h on the lat-lon grid as the shallow-water output writes h.dat
        (default 2N+1 latitudes x 4N longitudes): float32 (nlat, nlon)."""
        nlat = 2 * self.N + 1 if nlat is None else nlat
        nlon = 4 * self.N if nlon is None else nlon
        h = self._c(h)
        out = np.zeros((nlat, nlon), np.float32)
        self.lib.swo_latlon_h(self._h, _p(h), nlat, nlon, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    def latlon_uv(self, u, v, nlat=None, nlon=None):
        """The covariant winds on the lat-lon grid as the shallow-water
        output writes u.dat / v.dat: float32 (zonal, meridional), each
        (nlat, nlon)."""
        nlat = 2 * self.N + 1 if nlat is None else nlat
        nlon = 4 * self.N if nlon is None else nlon
        u, v = self._c(u), self._c(v)
        uo, vo = np.zeros((nlat, nlon), np.float32), np.zeros((nlat, nlon), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self.lib.swo_latlon_uv(self._h, _p(u), _p(v), nlat, nlon, uo.ctypes.data_as(fp), vo.ctypes.data_as(fp))
        return uo, vo

    def gaussian_hill(self, lam0, phi0, d0, amp):
        h = self.zeros(MESH_XY)
        self.lib.swo_gaussian_hill(self._h, lam0, phi0, d0, amp, _p(h))
        return h

    def ts2_linear(self, u0, h_mean):
        u, v, h = self.zeros(MESH_U), self.zeros(MESH_V), self.zeros(MESH_XY)
        self.lib.swo_ts2_linear(self._h, u0, h_mean, _p(u), _p(v), _p(h))
        return u, v, h

    def covariant_from_cart(self, vx, vy, vz, mesh):
        vx, vy, vz = self._c(vx), self._c(vy), self._c(vz)
        out = self.zeros(mesh)
        self.lib.swo_covariant_from_cart(self._h, _p(vx), _p(vy), _p(vz), mesh, _p(out))
        return out

    def basis(self, mesh):
        n = self.npts(mesh)
        a1, a2 = np.zeros((n, 3)), np.zeros((n, 3))
        self.lib.swo_basis(self._h, mesh, _p(a1), _p(a2))
        return a1, a2

    def halo_map(self, kind):
        """kind 0: xy scalar depth 1; 1: Ch vector; 2: xy vector depth 1;
        3: tangential components at the duplicated edge points."""
        n = self.lib.swo_halo_map(self._h, kind, *([None] * 9))
        arrs = [np.zeros(n, dtype=np.int32) for _ in range(9)]
        self.lib.swo_halo_map(self._h, kind, *[a.ctypes.data_as(_ip) for a in arrs])
        return dict(zip(("tp", "ti", "tj", "tc", "sp", "si", "sj", "sc", "sign"), arrs))
