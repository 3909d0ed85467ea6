"""The Fortran host layer (ch_swe_amd/fortran: iso_c_binding interfaces and
operator/stvec/timescheme types over the C ABI) and its Gaussian-hill driver,
the counterpart of SHALLOW_WATER_MAIN with testcase Gaussian_hill_linear and
rk4 (models/shallow_water/shallow_water_mod.f90 L29-145)."""
import os
import subprocess

import numpy as np
import pytest

from oracle.swe_oracle import MESH_U, MESH_V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ch_swe_amd", "swe_hip_gaussian")
CENTRES = [(np.pi, 0.0), (np.pi / 4, np.arcsin(1 / np.sqrt(3))),
           (0.27 * np.pi, np.arcsin(1 / np.sqrt(3)) - 0.02 * np.pi)]


def _have_exe():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (flang absent)")


def test_driver_reports_missing_device_cleanly():
    """Without a GPU the driver stops with the C ABI's message (the
    reference's parcomm%abort path), never a crash."""
    _have_exe()
    import ch_swe_amd  # noqa: F401  (library present)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    res = subprocess.run([EXE, "16", "42", "0", "2", "600", "1", "/dev/null"], capture_output=True,
                         text=True, env=env, timeout=120)
    if res.returncode == 0:
        pytest.skip("a device is visible despite HIP_VISIBLE_DEVICES=-1")
    assert "ch_swe_amd failure" in res.stdout and "device" in res.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("centre", [0, 1, 2])
def test_fortran_driver_config1_matches_oracle(oracle_mod, centre, tmp_path):
    """N=32, sbp21, FP64, dt=600 s, 100 steps from the Fortran driver against
    the oracle: 1e-12 relative."""
    _have_exe()
    out = tmp_path / "state.bin"
    print(f"fortran driver centre={centre}", flush=True)
    res = subprocess.run([EXE, "32", "21", "0", "100", "600", str(centre), str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    N = 32
    raw = np.fromfile(out, dtype=np.float64)
    nu, nv = 6 * N * (N + 1), 6 * (N + 1) * N
    u, v, h = raw[:nu], raw[nu:nu + nv], raw[nu + nv:]
    o = oracle_mod.Oracle(N, 21)
    h0 = o.gaussian_hill(*CENTRES[centre], 5e5 / o.radius, 10.0)
    ref = o.rk4(o.zeros(MESH_U), o.zeros(MESH_V), h0, 600.0, 100)
    for g, r in zip((u, v, h), ref):
        r = r.ravel()
        assert np.abs(g - r).max() / np.abs(r).max() < 1e-12
    assert "mass" in res.stdout and "done" in res.stdout
