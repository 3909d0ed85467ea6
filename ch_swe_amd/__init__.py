"""ch_swe_amd: MI355X-native linear shallow-water RK4 step on the staggered
(Ch) equiangular cubed sphere, SBP orders 21, 42 and 63 (63 on the
multi-kernel path).

Python host mirror of the reference's interface for this path:

    reference (Fortran)                              here
    create_linear_swm_operator(op, grav, cfg, dom)   create_linear_swm_operator(...)
      (operator_swm_factory_mod.f90 L586-685)
    operator_t%apply(vout, vin, domain)              OperatorSwmLin.apply(vout, vin)
      (operator_mod.f90 L11-41, operator_swm_lin_mod.f90 L66-108)
    stvec_t%assign / %update / %create_similar       StvecSwm.assign / .update / .create_similar
      (stvec/stvec_mod.f90 L10-30)
    create_timescheme(ts, v, "rk4")                  create_timescheme("rk4", op)
    rk4_t%step(v0, operator, domain, dt)             Rk4.step(v0, operator, dt)
      (time_schemes/rk4_mod.f90 L20-42)

Every call goes through the C ABI of include/ch_swe.h (libchswe.so, built
in-tree for gfx950).  There is no CPU fallback: if the library is missing or
no device is present, the calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "SweError", "lib", "Config", "OperatorSwmLin", "StvecSwm", "Rk4",
    "create_linear_swm_operator", "create_timescheme", "halo_map", "Ensemble", "Server",
    "FP64", "FP32", "LIB_PATH", "GenericConfig", "parse_config_file",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchswe.so")
FP64, FP32 = 0, 1

# every entry point declared in include/ch_swe.h
EXPORTS = (
    "swe_last_error", "swe_version", "swe_plan_create", "swe_plan_destroy", "swe_plan_npts",
    "swe_plan_stream", "swe_synchronize", "swe_state_create", "swe_state_destroy",
    "swe_state_upload", "swe_state_download", "swe_apply", "swe_assign", "swe_update",
    "swe_rk4_steps", "swe_integrals", "swe_integrals_noalloc", "swe_integrals_diag", "swe_integral_tends", "swe_cfl_diag", "swe_latlon_h", "swe_latlon_uv", "swe_solve", "swe_implicit_euler_steps",
    "swe_crank_nicolson_steps", "swe_state_upload_halo", "swe_state_download_halo", "swe_halo_map", "swe_ensemble_create",
    "swe_ensemble_destroy", "swe_ensemble_rk4", "swe_ensemble_stream", "swe_ensemble_synchronize",
    "swe_plan_timing", "swe_plan_kernel_times", "swe_server_create", "swe_server_destroy",
    "swe_server_submit", "swe_server_tick", "swe_server_wait", "swe_server_take", "swe_server_stats",
    "swe_server_counters",
    "swe_check_tables",
)
# the configuration reader's entry points (include/ch_swe_config.h)
CFG_EXPORTS = ("swe_cfg_parse", "swe_cfg_read", "swe_cfg_destroy", "swe_cfg_has", "swe_cfg_get_real",
               "swe_cfg_get_int", "swe_cfg_get_logical", "swe_cfg_get_str")


class SweError(RuntimeError):
    """A non-zero return code from the C ABI (message from swe_last_error);
    .code holds the SWE_E* value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


# return codes (include/ch_swe.h)
EINVAL, EHIP, ENOMEM, ENOTFINITE = -1, -2, -3, -4


class _CConfig(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("order", ctypes.c_int), ("precision", ctypes.c_int),
                ("coriolis_const", ctypes.c_int), ("radius", ctypes.c_double),
                ("grav", ctypes.c_double), ("h_mean", ctypes.c_double),
                ("omega", ctypes.c_double), ("axis", ctypes.c_double * 3),
                ("flags", ctypes.c_int), ("block_threads", ctypes.c_int),
                ("tile_x0", ctypes.c_int), ("tile_y0", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


def lib():
    """Load libchswe.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with ch_swe_amd/build.py "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.swe_last_error.restype = ctypes.c_char_p
    L.swe_version.restype = ctypes.c_char_p
    L.swe_plan_create.argtypes = [ctypes.POINTER(_CConfig), ctypes.POINTER(_vp)]
    L.swe_plan_destroy.argtypes = [_vp]
    L.swe_plan_npts.argtypes = [_vp, _ip, _ip, _ip]
    L.swe_plan_stream.argtypes = [_vp]
    L.swe_plan_stream.restype = _vp
    L.swe_synchronize.argtypes = [_vp]
    L.swe_state_create.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.swe_state_destroy.argtypes = [_vp]
    L.swe_state_upload.argtypes = [_vp, _dp, _dp, _dp]
    L.swe_state_download.argtypes = [_vp, _dp, _dp, _dp]
    L.swe_apply.argtypes = [_vp, _vp, _vp]
    L.swe_assign.argtypes = [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp]
    L.swe_update.argtypes = [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp]
    L.swe_rk4_steps.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int]
    L.swe_integrals.argtypes = [_vp, _vp, _dp, _dp]
    L.swe_integrals_noalloc.argtypes = [_vp, _vp, _dp, _dp]
    L.swe_integrals_diag.argtypes = [_vp, _vp, ctypes.c_double, _dp]
    L.swe_integral_tends.argtypes = [_vp, _vp, ctypes.c_double, _dp]
    L.swe_latlon_h.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.swe_cfl_diag.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double, _dp]
    L.swe_latlon_uv.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float)]
    L.swe_solve.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, _ip, _dp]
    for name in ("swe_implicit_euler_steps", "swe_crank_nicolson_steps"):
        getattr(L, name).argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, _dp]
    L.swe_state_upload_halo.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp]
    L.swe_state_download_halo.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp]
    L.swe_halo_map.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip] + [_ip] * 9
    L.swe_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
    L.swe_ensemble_destroy.argtypes = [_vp]
    L.swe_ensemble_rk4.argtypes = [_vp, _dp, _ip, ctypes.c_int]
    L.swe_ensemble_stream.argtypes = [_vp]
    L.swe_ensemble_stream.restype = _vp
    L.swe_ensemble_synchronize.argtypes = [_vp]
    L.swe_plan_timing.argtypes = [_vp, ctypes.c_int]
    L.swe_plan_kernel_times.argtypes = [_vp, _dp, _ip, _dp, _ip]
    L.swe_server_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.swe_server_destroy.argtypes = [_vp]
    L.swe_server_submit.argtypes = [_vp, ctypes.POINTER(_CConfig), _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_longlong)]
    L.swe_server_tick.argtypes = [_vp, ctypes.c_int, _ip]
    L.swe_server_wait.argtypes = [_vp]
    L.swe_server_take.argtypes = [_vp, ctypes.c_longlong, _dp, _dp, _dp, _ip]
    L.swe_server_stats.argtypes = [_vp, _ip, _ip, ctypes.POINTER(ctypes.c_size_t)]
    L.swe_server_counters.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 3
    L.swe_cfg_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(_vp)]
    L.swe_cfg_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_vp)]
    L.swe_cfg_destroy.argtypes = [_vp]
    L.swe_cfg_has.argtypes = [_vp, ctypes.c_char_p]
    L.swe_cfg_get_real.argtypes = [_vp, ctypes.c_char_p, _dp]
    L.swe_cfg_get_int.argtypes = [_vp, ctypes.c_char_p, _ip]
    L.swe_cfg_get_logical.argtypes = [_vp, ctypes.c_char_p, _ip]
    L.swe_cfg_get_str.argtypes = [_vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _ip]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().swe_last_error().decode()
        raise SweError(f"ch_swe error {rc}: {msg}", rc)


def _p(a):
    return a.ctypes.data_as(_dp)


# swe_config.flags and swe_ensemble_rk4 flags (include/ch_swe.h)
FLAG_UNFUSED, FLAG_LINEAR_BLOCKS, FLAG_GENERAL_TILES, FLAG_COMPENSATED = 1, 2, 4, 8
ENS_GRAPH = 1


class Config:
    """Operator configuration (the reference's domain/metric/operator config
    keys: N, planet_radius, omega, rotation_axis, coriolis_constant, h_mean,
    grav, SBP order via the operator names)."""

    def __init__(self, N, order=42, precision=FP64, radius=6.371e6, grav=9.81,
                 h_mean=1000.0, omega=7.292e-5, axis=(0.0, 0.0, 1.0),
                 coriolis_const=False, unfused=False, linear_blocks=False, block_threads=0,
                 general_tiles=False, compensated=False, tile_x0=0, tile_y0=0):
        self.N, self.order, self.precision = int(N), int(order), int(precision)
        self.tile_x0, self.tile_y0 = int(tile_x0), int(tile_y0)
        self.compensated = bool(compensated)
        self.general_tiles = bool(general_tiles)
        self.block_threads = int(block_threads)
        self.unfused = bool(unfused)
        self.linear_blocks = bool(linear_blocks)
        self.radius, self.grav, self.h_mean, self.omega = radius, grav, h_mean, omega
        self.axis = tuple(float(a) for a in axis)
        self.coriolis_const = bool(coriolis_const)

    def _c(self):
        c = _CConfig()
        c.N, c.order, c.precision = self.N, self.order, self.precision
        c.coriolis_const = 1 if self.coriolis_const else 0
        c.radius, c.grav, c.h_mean, c.omega = self.radius, self.grav, self.h_mean, self.omega
        c.axis[:] = self.axis
        c.flags = ((FLAG_UNFUSED if self.unfused else 0) | (FLAG_LINEAR_BLOCKS if self.linear_blocks else 0) |
                   (FLAG_GENERAL_TILES if self.general_tiles else 0) | (FLAG_COMPENSATED if self.compensated else 0))
        c.block_threads = self.block_threads
        c.tile_x0, c.tile_y0 = self.tile_x0, self.tile_y0
        return c


class _Plan:
    def __init__(self, cfg):
        L = lib()
        h = _vp()
        _check(L.swe_plan_create(ctypes.byref(cfg._c()), ctypes.byref(h)))
        self.h = h
        self.cfg = cfg

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().swe_plan_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        self.close()


class StvecSwm:
    """Device-resident state (u, v, h): stvec_swm_t
    (models/shallow_water/stvec/stvec_swm_mod.f90 L10-42)."""

    def __init__(self, plan):
        self._plan = plan
        h = _vp()
        _check(lib().swe_state_create(plan.h, ctypes.byref(h)))
        self.h = h
        N = plan.cfg.N
        self.shapes = {"u": (6, N + 1, N), "v": (6, N, N + 1), "h": (6, N + 1, N + 1)}

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().swe_state_destroy(self.h)
            self.h = _vp()

    def create_similar(self):
        return StvecSwm(self._plan)

    def upload(self, u, v, h):
        arrs = []
        for name, a in (("u", u), ("v", v), ("h", h)):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != int(np.prod(self.shapes[name])):
                raise ValueError(f"{name}: expected {self.shapes[name]}, got {a.shape}")
            arrs.append(a)
        _check(lib().swe_state_upload(self.h, *[_p(a) for a in arrs]))
        return self

    def download(self):
        out = [np.empty(self.shapes[k]) for k in ("u", "v", "h")]
        _check(lib().swe_state_download(self.h, *[_p(a) for a in out]))
        return tuple(out)

    def _halo_shapes(self, hw):
        return {k: (6, a + 2 * hw, b + 2 * hw) for k, (_, a, b) in self.shapes.items()}

    def upload_halo(self, hw, u, v, h):
        """Upload from the reference's tile arrays with a halo of width hw
        (p(1-hw:nx+hw, 1-hw:ny+hw) per panel); the halo entries are ignored."""
        shp = self._halo_shapes(hw)
        arrs = []
        for name, a in (("u", u), ("v", v), ("h", h)):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != shp[name]:
                raise ValueError(f"{name}: expected {shp[name]}, got {a.shape}")
            arrs.append(a)
        _check(lib().swe_state_upload_halo(self.h, int(hw), *[_p(a) for a in arrs]))
        return self

    def download_halo(self, hw, u, v, h):
        """Write the interior into the reference's halo'd tile arrays u, v, h
        (float64, C-contiguous, shapes (6, ny + 2hw, nx + 2hw)); their halo
        entries keep their values."""
        shp = self._halo_shapes(hw)
        for name, a in (("u", u), ("v", v), ("h", h)):
            if a.shape != shp[name] or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"{name}: expected a C-contiguous float64 array of shape {shp[name]}")
        _check(lib().swe_state_download_halo(self.h, int(hw), _p(u), _p(v), _p(h)))
        return u, v, h

    # stvec_t%assign (s1*v1 [+ s2*v2]) and %update (+= s1*v1 [+ s2*v2])
    def assign(self, s1, v1, s2=0.0, v2=None):
        _check(lib().swe_assign(self._plan.h, self.h, s1, v1.h, s2, v2.h if v2 is not None else None))

    def update(self, s1, v1, s2=0.0, v2=None):
        _check(lib().swe_update(self._plan.h, self.h, s1, v1.h, s2, v2.h if v2 is not None else None))


def write_latlon_record(path, field, rec):
    """Write one lat-lon field as record rec (1-based) of a direct-access
    float32 file, as latlon_write does (latlon_outputer_mod.f90 L52-60: record
    length nlat * nlon values, longitude fastest)."""
    a = np.ascontiguousarray(field, dtype=np.float32)
    mode = "r+b" if os.path.exists(path) else "wb"
    with open(path, mode) as f:
        f.seek((int(rec) - 1) * a.nbytes)
        f.write(a.tobytes())


class OperatorSwmLin:
    """operator_swm_lin_t: the linear SWE tendency, apply only (explicit
    schemes never call solve)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.plan = _Plan(cfg)

    def create_stvec(self):
        return StvecSwm(self.plan)

    def apply(self, vout, vin):
        _check(lib().swe_apply(self.plan.h, vout.h, vin.h))

    def integrals(self, state):
        m, e = ctypes.c_double(), ctypes.c_double()
        _check(lib().swe_integrals(self.plan.h, state.h, ctypes.byref(m), ctypes.byref(e)))
        return m.value, e.value

    def integrals_diag(self, state, grav=9.80616):
        """print_integrals_diag (swm_output_diag_mod.f90 L156-184): dict of
        hmin, hmax, mass, TE, KE, PE, Enstrophy (grav: the PE's gravity,
        Earth_grav in the reference)."""
        out = np.zeros(7)
        _check(lib().swe_integrals_diag(self.plan.h, state.h, float(grav), _p(out)))
        return dict(zip(("hmin", "hmax", "mass", "TE", "KE", "PE", "Enstrophy"), out.tolist()))

    def solve(self, vout, rhs, dt, rel_tol=1e-10, max_iter=200):
        """operator_swm_lin_t%solve (operator_swm_lin_mod.f90 L110-146):
        vout = (I - dt L0)^-1 rhs without the Coriolis term, CG on the
        Helmholtz problem for h (helm_solver_cfg: rel_tol, max_iter).
        Returns (iterations, final relative residual)."""
        it, rr = ctypes.c_int(), ctypes.c_double()
        _check(lib().swe_solve(self.plan.h, vout.h, rhs.h, float(dt), int(max_iter), float(rel_tol),
                               ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value

    def cfl_diag(self, state, dt, grav=9.80616):
        """print_CFL_diag (swm_output_diag_mod.f90 L110-154): dict of the
        gravity-wave and the two advective CFL numbers."""
        out = np.zeros(3)
        _check(lib().swe_cfl_diag(self.plan.h, state.h, float(dt), float(grav), _p(out)))
        return dict(zip(("gravity", "advective_x", "advective_y"), out.tolist()))

    def latlon_h(self, state, nlat=None, nlon=None):
        """h on the lat-lon grid as write_fields writes h.dat
        (swm_output_diag_mod.f90 L52-73, latlon_outputer_mod.f90 L33-63):
        float32 (nlat, nlon), default 2N+1 latitudes x 4N longitudes."""
        N = self.cfg.N
        nlat = 2 * N + 1 if nlat is None else int(nlat)
        nlon = 4 * N if nlon is None else int(nlon)
        out = np.empty((nlat, nlon), np.float32)
        _check(lib().swe_latlon_h(self.plan.h, state.h, nlat, nlon, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def latlon_uv(self, state, nlat=None, nlon=None):
        """The zonal and meridional winds on the lat-lon grid as write_fields
        writes u.dat / v.dat (swm_output_diag_mod.f90 L75-76,
        latlon_outputer_mod.f90 L65-107): two float32 (nlat, nlon) arrays."""
        N = self.cfg.N
        nlat = 2 * N + 1 if nlat is None else int(nlat)
        nlon = 4 * N if nlon is None else int(nlon)
        uo, vo = np.empty((nlat, nlon), np.float32), np.empty((nlat, nlon), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(lib().swe_latlon_uv(self.plan.h, state.h, nlat, nlon, uo.ctypes.data_as(fp), vo.ctypes.data_as(fp)))
        return uo, vo

    def integral_tends(self, state, grav=9.80616):
        """print_integral_tends (swm_output_diag_mod.f90 L236-342) of this
        operator's tendency at state: dict of TE_tend, Enstrophy_tend."""
        out = np.zeros(2)
        _check(lib().swe_integral_tends(self.plan.h, state.h, float(grav), _p(out)))
        return dict(zip(("TE_tend", "Enstrophy_tend"), out.tolist()))

    def synchronize(self):
        _check(lib().swe_synchronize(self.plan.h))

    @property
    def stream(self):
        return lib().swe_plan_stream(self.plan.h)

    def kernel_timing(self, enable=True):
        """Time every fused stage / records launch (dispatch-timestamped
        start and stop events)."""
        _check(lib().swe_plan_timing(self.plan.h, 1 if enable else 0))

    def kernel_times(self):
        """(stage_ms, n_stage, records_ms, n_records) since the last call."""
        sm, rm = ctypes.c_double(), ctypes.c_double()
        sn, rn = ctypes.c_int(), ctypes.c_int()
        _check(lib().swe_plan_kernel_times(self.plan.h, ctypes.byref(sm), ctypes.byref(sn),
                                           ctypes.byref(rm), ctypes.byref(rn)))
        return sm.value, sn.value, rm.value, rn.value


class Rk4:
    """rk4_t (time_schemes/rk4_mod.f90 L12-42)."""

    def step(self, v0, operator, dt, nsteps=1):
        _check(lib().swe_rk4_steps(operator.plan.h, v0.h, float(dt), int(nsteps)))


def create_linear_swm_operator(N, order=42, precision=FP64, grav=9.81, h_mean=1000.0,
                               radius=6.371e6, omega=7.292e-5, axis=(0.0, 0.0, 1.0),
                               coriolis_const=False, unfused=False, linear_blocks=False, block_threads=0,
                               general_tiles=False, compensated=False, tile_x0=0, tile_y0=0, tuned=True):
    """Factory mirroring create_linear_swm_operator
    (operator_swm_factory_mod.f90 L586-685) for the Ch staggering.
    unfused=True selects the multi-kernel diagnostic path (order 63 always
    runs there: the fused kernels are built for the sbp21 / sbp42 closures, so
    ensembles and the server take sbp21 / sbp42 plans only);
    linear_blocks=True
    the plain block order of the fused stage kernel; general_tiles=True sends
    every tile through its general path; block_threads (0 or 512) its
    workgroup size; compensated=True runs RK4 with compensated summation of
    each step's state update (states carry a compensation field); tile_x0 /
    tile_y0 the first-tile widths of the fused stage kernel's tiling (0: the
    library's split).  tuned=True takes the tiling and block order the
    autotuner stored for (N, order, precision) (ch_swe_amd/autotune.py) when
    none is given; results depend on them only at the level of rounding."""
    if tuned and not unfused and not linear_blocks and tile_x0 == 0 and tile_y0 == 0:
        from . import autotune
        entry = autotune.lookup(N, order, precision)
        if entry:
            tile_x0, tile_y0 = entry["tile_x0"], entry["tile_y0"]
            linear_blocks = bool(entry["linear_blocks"])
    return OperatorSwmLin(Config(N, order, precision, radius, grav, h_mean, omega, axis,
                                 coriolis_const, unfused, linear_blocks, block_threads, general_tiles,
                                 compensated, tile_x0, tile_y0))


class Ensemble:
    """Many independent solves advanced together (swe_ensemble_*): each RK4
    stage of every member is one launch of the fused stage kernel.  members
    are StvecSwm states of fused plans of one precision (plans may differ in
    N and order, members may share a plan)."""

    def __init__(self, members):
        self.members = list(members)  # keeps the states and their plans alive
        arr = (_vp * len(self.members))(*[m.h for m in self.members])
        h = _vp()
        _check(lib().swe_ensemble_create(len(self.members), arr, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().swe_ensemble_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        self.close()

    def rk4(self, dt, nsteps, graph=False):
        """Advance member i by nsteps[i] RK4 steps of dt[i] (scalars broadcast);
        asynchronous on the ensemble's stream, ordered with the plan streams."""
        n = len(self.members)
        dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)).copy()
        ns = np.broadcast_to(np.asarray(nsteps, dtype=np.int32), (n,)).copy()
        _check(lib().swe_ensemble_rk4(self.h, _p(dts), ns.ctypes.data_as(_ip), ENS_GRAPH if graph else 0))

    def synchronize(self):
        _check(lib().swe_ensemble_synchronize(self.h))

    @property
    def stream(self):
        return lib().swe_ensemble_stream(self.h)


class Server:
    """Continuous batching of independent solves under a device-memory
    budget (swe_server_*)."""

    def __init__(self, budget_bytes):
        h = _vp()
        _check(lib().swe_server_create(int(budget_bytes), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().swe_server_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        self.close()

    def submit(self, u, v, h, dt, nsteps, N, order=42, precision=FP64, **kw):
        """Queue a solve; returns its id.  kw: Config keywords (physical
        constants)."""
        cfg = Config(N, order, precision, **kw)
        shapes = ((6, N + 1, N), (6, N, N + 1), (6, N + 1, N + 1))
        arrs = []
        for name, a, shp in zip("uvh", (u, v, h), shapes):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != int(np.prod(shp)):
                raise ValueError(f"{name}: expected {shp} for N={N}, got {a.shape}")
            arrs.append(a)
        rid = ctypes.c_longlong()
        _check(lib().swe_server_submit(self.h, ctypes.byref(cfg._c()), *[_p(a) for a in arrs], float(dt),
                                       int(nsteps), ctypes.byref(rid)))
        self._shapes = getattr(self, "_shapes", {})
        self._shapes[rid.value] = N
        return rid.value

    def tick(self, chunk_steps=25):
        """One scheduling round (never waits for the device): retire finished
        batches, admit, start the next batches; returns completions."""
        n = ctypes.c_int()
        _check(lib().swe_server_tick(self.h, int(chunk_steps), ctypes.byref(n)))
        return n.value

    def wait(self):
        """Block until the batches in flight have finished."""
        _check(lib().swe_server_wait(self.h))

    def take(self, rid):
        """(u, v, h) of a finished request, or None if it is still running.
        A request that failed raises SweError (code ENOTFINITE for a
        non-finite result, whose arrays are then in .result)."""
        N = self._shapes[rid]
        out = [np.empty((6, N + 1, N)), np.empty((6, N, N + 1)), np.empty((6, N + 1, N + 1))]
        ready = ctypes.c_int()
        rc = lib().swe_server_take(self.h, int(rid), *[_p(a) for a in out], ctypes.byref(ready))
        if rc == 0 and not ready.value:
            return None
        del self._shapes[rid]
        if rc != 0:
            err = SweError(f"ch_swe error {rc}: {lib().swe_last_error().decode()}", rc)
            err.result = tuple(out) if rc == ENOTFINITE else None
            raise err
        return tuple(out)

    def stats(self):
        """(queued, active, device bytes in use)."""
        q, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _check(lib().swe_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(b)))
        return q.value, a.value, b.value

    def counters(self):
        """(requests admitted, cached plans evicted, requests failed)."""
        x = [ctypes.c_longlong() for _ in range(3)]
        _check(lib().swe_server_counters(self.h, *[ctypes.byref(v) for v in x]))
        return tuple(v.value for v in x)


class GenericConfig:
    """A parsed configuration in the reference's format, read through
    libchswe.so (include/ch_swe_config.h): parse_config_file
    (src/stuff/config/config_tools_mod.f90 L12-26) and generic_config_t%get
    (abstract_generic_config_mod.f90).  Keys are paths ("domain%metric%omega");
    arrays read back with get_array.  A missing key raises SweError unless a
    default is given; so does an entry of another type."""

    _MISSING = object()

    def __init__(self, text=None, path=None):
        L = lib()
        h = _vp()
        if (text is None) == (path is None):
            raise ValueError("give exactly one of text and path")
        if text is not None:
            _check(L.swe_cfg_parse(text.encode(), ctypes.byref(h)))
        else:
            _check(L.swe_cfg_read(os.fspath(path).encode(), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().swe_cfg_destroy(self._h)
            self._h = None

    __del__ = close

    def has(self, key):
        return lib().swe_cfg_has(self._h, key.encode()) == 1

    def _get(self, key, default, fn, cval, conv):
        if default is not self._MISSING and not self.has(key):
            return default
        _check(fn(self._h, key.encode(), ctypes.byref(cval)))
        return conv(cval.value)

    def get_real(self, key, default=_MISSING):
        return self._get(key, default, lib().swe_cfg_get_real, ctypes.c_double(), float)

    def get_int(self, key, default=_MISSING):
        return self._get(key, default, lib().swe_cfg_get_int, ctypes.c_int(), int)

    def get_logical(self, key, default=_MISSING):
        return self._get(key, default, lib().swe_cfg_get_logical, ctypes.c_int(), bool)

    def get_str(self, key, default=_MISSING):
        if default is not self._MISSING and not self.has(key):
            return default
        n = ctypes.c_int()
        buf = ctypes.create_string_buffer(256)
        _check(lib().swe_cfg_get_str(self._h, key.encode(), buf, len(buf), ctypes.byref(n)))
        if n.value >= len(buf):
            buf = ctypes.create_string_buffer(n.value + 1)
            _check(lib().swe_cfg_get_str(self._h, key.encode(), buf, len(buf), ctypes.byref(n)))
        return buf.value.decode()

    def get_array(self, key, default=_MISSING):
        """A one-level array entry as a list of reals (name%1 ... name%n)."""
        if default is not self._MISSING and not self.has(key + "%length"):
            return default
        n = self.get_int(key + "%length")
        return [self.get_real(f"{key}%{i}") for i in range(1, n + 1)]


def parse_config_file(path):
    """parse_config_file (config_tools_mod.f90 L12-26)."""
    return GenericConfig(path=path)


class ImplicitScheme:
    """implicit_Eul_t (time_schemes/implicit_Eul_mod.f90 L20-30) or
    crank_nikolson_t (crank_nikolson_mod.f90 L20-32) on the operator's solve;
    rel_tol / max_iter: the helm_solver_cfg of the operator.  After a step,
    .iters and .relres hold the largest CG iteration count and final
    relative residual over its steps."""

    def __init__(self, name, rel_tol=1e-10, max_iter=200):
        self.name, self.rel_tol, self.max_iter = name, float(rel_tol), int(max_iter)
        self._fn = "swe_crank_nicolson_steps" if name == "crank_nikolson" else "swe_implicit_euler_steps"
        self.iters, self.relres = 0, 0.0

    def step(self, v0, operator, dt, nsteps=1):
        it, rr = ctypes.c_int(), ctypes.c_double()
        _check(getattr(lib(), self._fn)(operator.plan.h, v0.h, float(dt), int(nsteps), self.max_iter, self.rel_tol,
                                        ctypes.byref(it), ctypes.byref(rr)))
        self.iters, self.relres = it.value, rr.value


def create_timescheme(name, rel_tol=1e-10, max_iter=200):
    """Factory mirroring timescheme_factory_mod.f90 L28-45 for the schemes on
    this path: 'rk4', 'implicit_Eul', 'crank_nikolson' (the last two take the
    operator's Helmholtz solver settings)."""
    if name == "rk4":
        return Rk4()
    if name in ("implicit_Eul", "crank_nikolson"):
        return ImplicitScheme(name, rel_tol, max_iter)
    raise ValueError(f"unknown time scheme '{name}' (this build provides rk4, implicit_Eul, crank_nikolson)")


def check_tables(N, precision=FP32):
    """Host-only self-check of the fused kernels' plan-time tables (see
    swe_check_tables); raises SweError on a mismatch."""
    _check(lib().swe_check_tables(int(N), int(precision)))


def halo_map(N, order, kind):
    """Host-only edge map of the in-device halo gathers (see swe_halo_map)."""
    L = lib()
    n = ctypes.c_int()
    _check(L.swe_halo_map(N, order, kind, ctypes.byref(n), *([None] * 9)))
    arrs = [np.zeros(n.value, dtype=np.int32) for _ in range(9)]
    _check(L.swe_halo_map(N, order, kind, ctypes.byref(n), *[a.ctypes.data_as(_ip) for a in arrs]))
    return dict(zip(("tp", "ti", "tj", "tc", "sp", "si", "sj", "sc", "sign"), arrs))
