"""The configuration reader (include/ch_swe_config.h, ch_swe_amd.GenericConfig)
and the configuration-driven model program (ch_swe_amd/swe_hip_model): the
reference's .cfg format (geosci_config_parser_mod.f90 L21-454,
arithmetic_parser_mod.f90 L13-369) and run_shallow_water_model's checks of
its sections (shallow_water_mod.f90 L29-145).  CPU only: the parser pins are
closed forms of each rule (operator priorities and associativity, integer
arithmetic, the constants of const_mod.f90 L4-11, name lookup order), and the
program is run up to the point where it needs a device."""
import math
import os
import re
import subprocess

import pytest

import ch_swe_amd as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "ch_swe_amd", "swe_hip_model")


def _val(expr):
    c = cs.GenericConfig(text=f"a = {expr}")
    try:
        return "int", c.get_int("a")
    except cs.SweError:
        return "real", c.get_real("a")


def test_library_exports_every_config_symbol():
    hdr = open(os.path.join(ROOT, "include", "ch_swe_config.h")).read()
    declared = set(re.findall(r"\b(swe_cfg_[a-z_]+)\s*\(", hdr))
    assert declared == set(cs.CFG_EXPORTS)
    for name in declared:
        assert hasattr(cs.lib(), name), name


@pytest.mark.parametrize("expr,kind,value", [
    # priorities: add/sub 1 < mul/div 2 < pow 3 < unary minus 10
    ("2+3*4", "int", 14), ("(2+3)*4", "int", 20), ("2*3^2", "int", 18),
    ("-2^2", "int", 4), ("2*-3^2", "int", 18), ("3 + -2", "int", 1),
    # equal priorities pop left to right, pow included
    ("10-4-3", "int", 3), ("100/10/5", "int", 2), ("2^3^2", "int", 64), ("2**3**2", "int", 64),
    # integer arithmetic when both operands are integers (Fortran semantics)
    ("7/2", "int", 3), ("-7/2", "int", -3), ("2^-1", "int", 0), ("7.0/2", "real", 3.5),
    ("2.0^-1", "real", 0.5), ("2^0.5", "real", math.sqrt(2.0)),
    # functions; int() truncates towards zero and keeps integers
    ("int(3.7)", "int", 3), ("int(-3.7)", "int", -3), ("int(5)", "int", 5),
    ("sqrt(16.0)", "real", 4.0), ("sqrt(2)", "real", math.sqrt(2.0)), ("sin(pi/2)", "real", 1.0),
    ("cos(0.0)", "real", 1.0), ("tan(pi/4)", "real", 1.0), ("acos(-1.0)", "real", math.pi),
    ("asin(1.0)*2", "real", math.pi), ("atan(1.0)*4", "real", math.pi), ("exp(1.0)", "real", math.e),
    ("log(exp(2.0))", "real", 2.0),
    # number forms, Fortran exponent letters included
    ("0.6371229d7", "real", 6371229.0), ("1.5D0", "real", 1.5), ("1e-4", "real", 1e-4),
    ("1.E3", "real", 1000.0), (".5", "real", 0.5), ("5.", "real", 5.0),
    # constants (const_mod.f90 L4-11, geosci_config_parser_mod.f90): day is real, the rest int
    ("pi", "real", math.pi), ("Earth_grav", "real", 9.80616), ("Earth_radii", "real", 6371229.0),
    ("Earth_omega", "real", 2 * math.pi / 86164.0), ("day", "real", 86400.0), ("hour", "int", 3600),
    ("min", "int", 60), ("sec", "int", 1), ("1.0*900*sec", "real", 900.0), ("25.0*day", "real", 25 * 86400.0),
    # the linear Gaussian-hill run's mean depth: (2 pi a / 5 days)^2 / g
    ("(2*pi*Earth_radii/(5*day))^2/Earth_grav", "real",
     (2 * math.pi * 6371229.0 / (5 * 86400.0)) ** 2 / 9.80616),
    ("29400/Earth_grav", "real", 29400 / 9.80616),
])
def test_arithmetic(expr, kind, value):
    k, v = _val(expr)
    assert k == kind, (expr, k, v)
    assert v == pytest.approx(value, rel=1e-15, abs=1e-15), expr


def test_values_sections_arrays_and_comments():
    text = """
    # a comment line
    N = 32, name = 'Ch'   # trailing comment
    other = "double quoted"
    flag = .true., off = .false.
    scale = 2.0
    axis = [1.0/sqrt(2.0), 0.0, 1.0/sqrt(2.0)]
    outer = {
        scale = 3.0,
        twice = 2*scale,           # the section's own scale
        root_n = N*2,              # found in the root
        inner = { deeper = scale + 1.0, s = 'x' }   # not the parent's: the root's
    }
    after = scale*N
    """
    c = cs.GenericConfig(text=text)
    assert c.get_int("N") == 32 and c.get_str("name") == "Ch" and c.get_str("other") == "double quoted"
    assert c.get_logical("flag") is True and c.get_logical("off") is False
    assert c.get_array("axis") == pytest.approx([1 / math.sqrt(2), 0.0, 1 / math.sqrt(2)], rel=1e-15)
    assert c.get_int("axis%length") == 3
    assert c.get_real("outer%twice") == 6.0 and c.get_int("outer%root_n") == 64
    assert c.get_real("outer%inner%deeper") == 3.0 and c.get_str("outer%inner%s") == "x"
    assert c.get_real("after") == 64.0
    assert c.has("outer%scale") and not c.has("outer%missing")
    # defaults only for absent keys
    assert c.get_real("missing", 1.5) == 1.5 and c.get_real("scale", 1.5) == 2.0
    assert c.get_str("missing", "None") == "None"


def test_file_reader(tmp_path):
    p = tmp_path / "run.cfg"
    p.write_text("dt = 1.0*600*sec,\nsimulation_time = 2.0*hour\n")
    c = cs.parse_config_file(p)
    assert c.get_real("dt") == 600.0 and c.get_real("simulation_time") == 7200.0


@pytest.mark.parametrize("text,key,method,match", [
    ("a = b + 1", None, None, "unknown variable 'b'"),
    # names resolve in the entry's own section, then the root, then the
    # constants (geosci_config_parser_mod.f90 L410), never a parent section
    ("o = { x = 1, i = { y = x } }", None, None, "unknown variable 'x'"),
    ("a = (1 + 2", None, None, "parenthes"),
    ("a = 1 +", None, None, ""),
    ("a = 'open", None, None, ""),
    ("a = { b = 1", None, None, ""),
    ("a = 1", "b", "get_real", "no entry 'b'"),
    ("a = 1", "a", "get_real", "not a real"),       # the reference's type error for int -> real
    ("a = 1.0", "a", "get_int", "not an integer"),
    ("a = 'x'", "a", "get_logical", "not a logical"),
    ("a = 1", "a", "get_str", ""),
])
def test_errors(text, key, method, match):
    with pytest.raises(cs.SweError, match=match or None) as err:
        c = cs.GenericConfig(text=text)
        getattr(c, method)(key)
    assert err.value.code == cs.EINVAL


# --- the configuration-driven model program -------------------------------

# the paper's linear Gaussian-hill run with the Ch operator block
# (test/scripts/article_experiments/linear_gauss_run.sh and gen_config_linear.sh
# output for N=32, dt=900, Ch, sbp21, corner centre, omega 1e-4)
GAUSS_CFG = """
domain = {
    N = 32,
    Nz = 1,
    topology_type         = 'cube',
    horizontal_staggering = 'Ch',
    metric                = {
            metric_type             = 'shallow_atmosphere_metric',
            metric_2d_type          = 'ecs',
            planet_radius           = Earth_radii,
            omega                   = 1e-4
            coriolis_constant       = .true.
}},
dt = 1.0*900*sec,
simulation_time = 2.0*hour,
tau_write = 1.0*hour,
tau_diagnostics = 1.0*hour,
timescheme_name = 'rk4'
testcase_section = 'Gauss_linear_config'
    Gauss_linear_config = {
        testcase_name ='Gaussian_hill_linear',
        lam0 = 0.25*pi,
        phi0 = asin(1.0/sqrt(3.0)),
        refsol_path='gauss_h_corner_cub192_xy.dat'
    }
swm_operator_section = 'operator_conf'
    operator_conf = {
        swm_operator_type        = 'linear_swm_operator'
        v_components_type        = 'covariant',
        h_mean                   = (2*pi*Earth_radii/(5*day))^2/Earth_grav,
        div_op_name              = 'div_ch_sbp_sat_21_proj',
        grad_op_name             = 'grad_ch_sbp_sat_21',
        coriolis_op_name         = 'coriolis_Ch_sbp21',
        co2contra_op_name        = 'co2contra_ch_sbp21',
        quadrature_name          = 'SBP_C21_quadrature',
        helm_solver_cfg = {
            solver_name = 'cg',
            max_iter = 100,
            rel_tol  = 0.00001
            },
    },
    output_diag_section = 'output_diag_conf_Ch'
    output_diag_conf_Ch = {
        output_curl        = .false.,
        output_div         = .false.
        curl_operator_name = 'curl_ch_sbp_sat_21',
        div_operator_name = 'div_ch_sbp_sat_21_proj'
        co2contra_operator_name = 'co2contra_ch_sbp21',
        h2uv_operator_name = 'interp2d_p2uv_Ch_sbp21',
        uv2q_operator_name = 'interp2d_uv2qvec_Ch_sbp21'
        quadrature_name          = 'SBP_C21_quadrature'
    }
"""


def ts2_cfg(gauss=GAUSS_CFG):
    """run_ts2_linear.sh's configuration: rotated axis, Earth's omega,
    Williamson test 2 in linear mode, h_mean = 29400/g."""
    text = gauss.replace("omega                   = 1e-4\n            coriolis_constant       = .true.",
                         "omega                   = Earth_omega,\n"
                         "            rotation_axis          = [1.0/sqrt(2.0),0.0,1.0/sqrt(2.0)]")
    i, j = text.index("testcase_section"), text.index("swm_operator_section")
    text = (text[:i] + "testcase_section = 'Williamson_test2_config'\n"
            "Williamson_test2_config = {\n    testcase_name ='Williamson_test2',\n    linear_mode   = .true.\n}\n"
            + text[j:])
    return text.replace("(2*pi*Earth_radii/(5*day))^2/Earth_grav", "29400/Earth_grav")


def test_paper_configs_parse():
    c = cs.GenericConfig(text=GAUSS_CFG)
    assert c.get_int("domain%N") == 32 and c.get_str("domain%horizontal_staggering") == "Ch"
    assert c.get_real("domain%metric%planet_radius") == 6371229.0
    assert c.get_logical("domain%metric%coriolis_constant") is True
    assert c.get_real("dt") == 900.0 and c.get_real("tau_write") == 3600.0
    sec = c.get_str("testcase_section")
    assert c.get_real(sec + "%lam0") == pytest.approx(math.pi / 4, rel=1e-15)
    assert c.get_real(sec + "%phi0") == pytest.approx(math.asin(1 / math.sqrt(3)), rel=1e-15)
    op = c.get_str("swm_operator_section")
    assert c.get_str(op + "%div_op_name") == "div_ch_sbp_sat_21_proj"
    assert c.get_int(op + "%helm_solver_cfg%max_iter") == 100
    t = cs.GenericConfig(text=ts2_cfg())
    assert t.get_array("domain%metric%rotation_axis") == pytest.approx([0.5 ** 0.5, 0.0, 0.5 ** 0.5], rel=1e-15)
    assert t.get_logical("Williamson_test2_config%linear_mode") is True
    assert t.get_real("domain%metric%omega") == pytest.approx(2 * math.pi / 86164.0, rel=1e-15)


def _run_model(tmp_path, text, *args):
    if not os.path.exists(MODEL):
        pytest.skip("model program not built (flang absent)")
    cfg = tmp_path / "run.cfg"
    cfg.write_text(text)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    return subprocess.run([MODEL, str(cfg), str(tmp_path / "out.bin"), *args], capture_output=True, text=True,
                          env=env, timeout=120, cwd=tmp_path)


@pytest.mark.parametrize("edit,match", [
    (("'Ch'", "'Ah'"), "horizontal_staggering = 'Ah'"),
    (("div_ch_sbp_sat_21_proj'", "div_ch_sbp_sat_84_proj'"), "div_op_name = 'div_ch_sbp_sat_84_proj'"),
    (("grad_ch_sbp_sat_21'", "grad_ch_sbp_sat_42'"), "grad_op_name: SBP order differs"),
    (("'linear_swm_operator'", "'vector_invariant_swm_operator'"), "swm_operator_type"),
    (("'covariant'", "'contravariant'"), "v_components_type = 'contravariant'"),
    (("timescheme_name = 'rk4'", "timescheme_name = 'ars343'"), "timescheme_name = 'ars343'"),
    (("timescheme_name = 'rk4'", "timescheme_name = 'rk4'\ndiffusion_section = 'diff'"), "diffusion"),
    (("'Gaussian_hill_linear'", "'RH4'"), "testcase_name RH4"),
    (("Nz = 1", "Nz = 10"), "Nz"),
    (("dt = 1.0*900*sec", "dt = 900"), "'dt' is not a real"),
    (("coriolis_constant       = .true.", "rotation_matrix = [0.0,1.0,0.0, 1.0,0.0,0.0, 0.0,0.0,1.0]"),
     "rotation_matrix"),
])
def test_model_rejects_unsupported_configs(tmp_path, edit, match):
    """Each unsupported choice stops the program before any device work, with
    a message naming the key (the reference's parcomm%abort path)."""
    text = GAUSS_CFG.replace(*edit)
    assert text != GAUSS_CFG
    res = _run_model(tmp_path, text)
    assert res.returncode != 0
    assert "ch_swe_amd failure" in res.stdout and match in res.stdout, res.stdout


def test_model_rejects_nonlinear_ts2(tmp_path):
    res = _run_model(tmp_path, ts2_cfg().replace("linear_mode   = .true.", "linear_mode   = .false."))
    assert res.returncode != 0 and "linear_mode" in res.stdout


@pytest.mark.parametrize("text", [GAUSS_CFG, ts2_cfg()], ids=["gauss", "ts2"])
def test_model_accepts_paper_configs(tmp_path, text):
    """Valid paper configurations pass every check and stop only at the
    missing device."""
    res = _run_model(tmp_path, text, "fp32")
    if res.returncode == 0:
        pytest.skip("a device is visible despite HIP_VISIBLE_DEVICES=-1")
    assert "create_domain_hip" in res.stdout and "device" in res.stdout, res.stdout
