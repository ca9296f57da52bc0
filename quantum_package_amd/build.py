"""Builds libqpx.so (HIP, gfx950) in-tree with hipcc.

Every call recompiles (objects go to quantum_package_amd/_build/); the
shared library lands next to this file so that it travels with the source
tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqpx.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-ffp-contract=fast",
]

SOURCES = ["context.hip", "wavefunction.hip", "sigma.hip", "sigma_sparse.hip", "davidson.hip", "selection.hip", "mlp.hip",
           "psi_update.hip", "s2_complete.hip"]


# the selection kernel keeps the reference's floating-point operation order
# (bit-identical couplings and EN-PT2 terms): no FMA contraction there
PER_FILE = {"selection.hip": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *PER_FILE.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FORTRAN_DRIVER = os.path.join(HERE, "qpx_sigma_driver")


def build_fortran() -> str:
    """The Fortran binding module and its example host program (links libqpx)."""
    fdir = os.path.join(OBJ, "fortran")
    os.makedirs(fdir, exist_ok=True)
    src = os.path.join(HERE, "fortran")
    subprocess.check_call([FLANG, "-O2", "-J", fdir, "-c", os.path.join(src, "qpx_mod.f90"), "-o",
                           os.path.join(fdir, "qpx_mod.o")])
    subprocess.check_call([FLANG, "-O2", "-J", fdir, "-c", os.path.join(src, "qpx_sigma_driver.f90"), "-o",
                           os.path.join(fdir, "driver.o")])
    subprocess.check_call([FLANG, "-o", FORTRAN_DRIVER, os.path.join(fdir, "driver.o"), os.path.join(fdir, "qpx_mod.o"),
                           "-L" + HERE, "-lqpx", "-Wl,-rpath,$ORIGIN"])
    return FORTRAN_DRIVER


def build(jobs: int = 6) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    if os.path.exists(FLANG):
        build_fortran()
    return OUT


if __name__ == "__main__":
    print(build())
