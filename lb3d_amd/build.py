"""Build the HIP library lb3d_amd/liblb3d_gpu.so for gfx950 (and the Fortran
host, when amdflang is present) in place.

The FP64 translation units are compiled with -ffp-contract=off so that the
plain path keeps the reference's rounding order (no fused multiply-add); the
FP32 deviation unit may contract.  Packed FP32 instructions are disabled for
every HIP unit.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblb3d_gpu.so")
PROBE = os.path.join(HERE, "liblb3d_bwprobe.so")  # bench.py's HBM ceiling probe
ARCH = os.environ.get("LB3D_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas"]
UNITS = [
    ("lb3d_gpu.hip", ["-ffp-contract=off"]),
    ("lb3d_batch.hip", ["-ffp-contract=off"]),
    ("lb3d_fp64.hip", ["-ffp-contract=off"]),
    ("lb3d_fp32.hip", ["-ffp-contract=on"]),
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    procs, objs = [], []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + extra
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise SystemExit("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("hipcc failed")
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = PROBE + ".tmp"
    run([HIPCC, os.path.join(CSRC, "lb3d_bwprobe.hip"), "-shared", "-o", tmp] + COMMON)
    os.replace(tmp, PROBE)
    return LIB


def build_fortran():
    """The Fortran host (ISO_C_BINDING) against the library, if amdflang exists."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    fdir = os.path.join(HERE, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if fc is None or not os.path.exists(mk):
        return None
    run(["make", "-s", "-C", fdir, f"FC={fc}"])
    return os.path.join(fdir, "bin", "lb3d")


def build_all():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "fortran":
        build_fortran()
    else:
        build_all()
