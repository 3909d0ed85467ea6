"""Build libchswe.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libchswe.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["swe_lib.hip", "swe_topology.cpp", "swe_config.cpp", "swe_latlon.cpp"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("FC_FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_SOURCES = ["swe_hip_capi_mod.f90", "swe_hip_mod.f90", "swe_hip_cfg_mod.f90"]
FORTRAN_PROGRAMS = {"swe_hip_gaussian": "swe_hip_gaussian_main.f90", "swe_hip_model": "swe_hip_model_main.f90"}


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{what} failed")


def build_fortran(verbose=False):
    """The Fortran host layer and driver programs (flang, linked to
    libchswe.so with an $ORIGIN rpath).  Skipped when flang is absent."""
    if not os.path.exists(FC):
        return None
    mod_dir = os.path.join(FORTRAN, "_build")
    os.makedirs(mod_dir, exist_ok=True)
    out = []
    for prog, main in FORTRAN_PROGRAMS.items():
        exe = os.path.join(HERE, prog)
        srcs = [os.path.join(FORTRAN, f) for f in FORTRAN_SOURCES + [main]]
        _run([FC, "-O2", "-module-dir", mod_dir] + srcs +
             ["-L" + HERE, "-lchswe", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"], f"flang {prog}", verbose)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def _asm_cmd(out):
    """Device-only assembly of the kernels, for the ISA check."""
    flags = [f for f in FLAGS if f not in ("-shared", "-Wall", "-Wno-unused-function")]
    return [HIPCC] + flags + ["--cuda-device-only", "-S", "-w", os.path.join(CSRC, "swe_lib.hip"), "-o", out]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    asm_dir = os.path.join(HERE, "_build")
    os.makedirs(asm_dir, exist_ok=True)
    asm = os.path.join(asm_dir, "swe_lib_gfx950.s")
    asm_proc = subprocess.Popen(_asm_cmd(asm), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    try:
        _run(cmd, "hipcc (libchswe.so)", verbose)
    finally:
        asm_out, _ = asm_proc.communicate()
    if asm_proc.returncode != 0:
        sys.stderr.write(asm_out)
        raise RuntimeError("hipcc -S (ISA check) failed")
    if HERE not in sys.path:
        sys.path.insert(0, HERE)
    import isa_check

    found = isa_check.check_file(asm)
    if found:
        lines = [f"{k}: {len(v)} address reads of SGPRs undefined on some path (first at line {v[0][0]})"
                 for k, v in found.items()]
        raise RuntimeError("ISA check failed:\n" + "\n".join(lines))
    os.replace(LIB + ".tmp", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
