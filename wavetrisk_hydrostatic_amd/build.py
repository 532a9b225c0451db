"""Builds the in-tree HIP library libwavetrisk_gpu.so for gfx950 (MI355X).

    python -m wavetrisk_hydrostatic_amd.build

The FP64 kernels are compiled without FMA contraction so that they reproduce
the CPU oracle bit for bit; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libwavetrisk_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("trisk.hip", "elliptic.hip", "ml.hip", "mlp.hip", "diffusion.hip",
                                                     "remap.hip", "vdiff.hip", "grid.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FORTRAN_MODULE = os.path.join(HERE, "fortran", "wavetrisk_gpu.f90")
FORTRAN_CHECK = os.path.join(ROOT, "tests", "fortran", "binding_check.f90")
FORTRAN_BIN = os.path.join(HERE, "fortran", "bin")


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{what} failed with exit code {res.returncode}")


def build_fortran(verbose: bool = False) -> str:
    """the ISO_C_BINDING module and the program that calls the library
    through it (tests/fortran/binding_check.f90), linked against the
    in-tree library with an rpath to it"""
    os.makedirs(FORTRAN_BIN, exist_ok=True)
    obj = os.path.join(FORTRAN_BIN, "wavetrisk_gpu.o")
    exe = os.path.join(FORTRAN_BIN, "binding_check")
    _run([FC, "-O1", "-c", FORTRAN_MODULE, "-module-dir", FORTRAN_BIN, "-o", obj], verbose, "amdflang (module)")
    _run([FC, "-O1", "-I", FORTRAN_BIN, FORTRAN_CHECK, obj, "-L" + HERE, "-lwavetrisk_gpu",
          "-Wl,-rpath,$ORIGIN/../..", "-o", exe + ".tmp"], verbose, "amdflang (binding_check)")
    os.replace(exe + ".tmp", exe)
    return exe


def build(verbose: bool = False) -> str:
    _run([HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES], verbose, "hipcc")
    os.replace(LIB + ".tmp", LIB)
    if os.path.exists(FC):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
