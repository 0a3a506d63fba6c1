"""Build libcube_hip.so (the HIP library behind include/cube.h) in-tree.

hipcc compiles every csrc/*.hip for gfx950 in parallel and links one shared
library next to this file.  Flags:
  -ffp-contract=off      no FMA contraction: each expression rounds as the
                         reference's gfortran build on x86-64 does
  -fno-fast-math         IEEE division/sqrt
  -target-feature -packed-fp32-ops
                         no packed FP32 instructions (they are unreliable on
                         the target machines)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcube_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUBE_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FORT = os.path.join(HERE, "fortran")
FC = os.environ.get("CUBE_FC", "/opt/rocm/bin/amdflang")
DRIVER = os.path.join(FORT, "cube_main")


def build_fortran():
    """The Fortran driver (fortran/cube_main.f90 over the ISO_C_BINDING
    module fortran/cube_c.f90 and the time integration fortran/cube_clock.f90),
    linked against libcube_hip.so.  No FMA contraction, as in the HIP
    library."""
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    tmp = DRIVER + ".tmp"
    cmd = [FC, "-O2", "-ffp-contract=off", "-module-dir", mod, os.path.join(FORT, "cube_c.f90"),
           os.path.join(FORT, "cube_clock.f90"), os.path.join(FORT, "cube_main.f90"),
           "-o", tmp, "-L" + HERE, "-lcube_hip", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
