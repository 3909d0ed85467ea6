"""CPU: the fused kernels compiled for the host and run under
AddressSanitizer (tests/emu/swe_emu.cpp), against the oracle.

Each block's threads run as fibers with emulated barriers and lane shuffles,
blocks in a random order, every device buffer an exact-size heap allocation
(AddressSanitizer reports any access outside it) and the state padding NaN.
The cases cover single-tile panels, uneven last tiles, interior tiles, the
general-tiles and plain-order flags, both SBP orders and precisions, the RK4
stage modes (compensated summation included), the multi-instance (ensemble) kernels, the tile-wise
integrals, the integral diagnostics and tendencies and the lat-lon output.  No GPU is used."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_SRC = os.path.join(ROOT, "tests", "emu", "swe_emu.cpp")
HIP_INCLUDE = "/opt/rocm/include"

# (case, N, order, precision, steps, seed, flags)
CASES = [
    ("apply", 8, 42, 64, 1, 1, 0),      # one tile per panel
    ("apply", 16, 21, 64, 1, 2, 0),     # last FP64 tile one row high
    ("apply", 57, 42, 64, 1, 3, 0),     # interior FP64 tiles
    ("apply", 56, 42, 32, 1, 4, 0),     # interior FP32 tiles
    ("apply", 33, 21, 64, 1, 5, 1),     # every tile on the general path
    ("apply", 40, 42, 32, 1, 6, 2),     # plain block order
    ("rk4", 10, 42, 64, 2, 7, 0),       # stage modes 1-3, two steps
    ("rk4", 50, 21, 32, 1, 8, 0),
    ("rk4", 12, 42, 64, 3, 13, 4),      # compensated last stage (mode 4)
    ("rk4", 40, 21, 32, 2, 14, 4),
    ("integrals", 20, 42, 64, 1, 9, 0),
    ("integrals", 48, 21, 32, 1, 10, 0),
    ("multi", 12, 42, 64, 2, 11, 0),    # two members (N, N + 7), own dt and steps
    ("multi", 9, 21, 32, 2, 12, 0),
    ("diag", 20, 42, 64, 1, 15, 0),     # integral diagnostics (curl SAT halos, centre geometry)
    ("diag", 41, 21, 32, 1, 16, 0),
    ("diag", 57, 42, 64, 1, 17, 0),     # with interior tiles
    ("diag", 32, 21, 64, 1, 18, 0),     # a one-point last tile along x: the centres before it read the SAT halo
    ("tends", 20, 42, 64, 1, 19, 0),    # integral tendencies (two passes over a tile's fields)
    ("tends", 32, 21, 64, 1, 20, 0),
    ("tends", 57, 42, 32, 1, 21, 0),
    ("latlon", 12, 42, 64, 1, 22, 0),   # lat-lon output of h (ECS halo map, cubic interpolation)
    ("latlon", 17, 21, 32, 1, 23, 0),
    ("latlonuv", 12, 42, 64, 1, 24, 0),  # lat-lon winds (colocation, vector halo map, zonal / meridional)
    ("latlonuv", 17, 21, 32, 1, 25, 0),
    ("lines", 12, 63, 64, 1, 26, 0),     # the multi-kernel path's 1-D stencils, SAT included
    ("lines", 13, 63, 32, 1, 27, 0),
    ("lines", 8, 42, 64, 1, 28, 0),
    ("lines", 4, 21, 64, 1, 29, 0),
]


def _have_toolchain():
    return shutil.which("g++") and shutil.which("gcc") and os.path.isdir(HIP_INCLUDE)


@pytest.fixture(scope="module")
def emu_exe(tmp_path_factory):
    if not _have_toolchain():
        pytest.skip("g++/gcc or the HIP headers are missing")
    out = tmp_path_factory.mktemp("emu")
    common = ["-O1", "-g", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-fsanitize=address",
              "-fno-omit-frame-pointer"]
    obj = str(out / "swe_oracle.o")
    subprocess.run(["gcc", "-std=c11"] + common + ["-c", os.path.join(ROOT, "oracle", "swe_oracle.c"), "-o", obj],
                   check=True, capture_output=True, text=True)
    exe = str(out / "swe_emu")
    res = subprocess.run(["g++", "-std=c++17"] + common + ["-Wno-attributes", "-D__HIP_PLATFORM_AMD__",
                          "-I" + HIP_INCLUDE, EMU_SRC, os.path.join(ROOT, "ch_swe_amd", "csrc", "swe_topology.cpp"),
                          os.path.join(ROOT, "ch_swe_amd", "csrc", "swe_latlon.cpp"),
                          obj, "-o", exe, "-lm"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return exe


@pytest.mark.parametrize("case,N,order,prec,steps,seed,flags", CASES)
def test_fused_kernels_under_asan(emu_exe, case, N, order, prec, steps, seed, flags):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
    res = subprocess.run([emu_exe, case, str(N), str(order), str(prec), str(steps), str(seed), str(flags)],
                         capture_output=True, text=True, timeout=600, env=env)
    out = res.stdout + res.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert res.returncode == 0, out[-4000:]
    assert res.stdout.rstrip().endswith("ok"), out[-2000:]
