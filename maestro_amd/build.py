"""Build the gfx950 library (and the Fortran drop-in shim) in-tree.

``python -m maestro_amd.build`` compiles:
  * ``maestro_amd/libmaestro_mac.so``   hipcc, --offload-arch=gfx950
  * ``maestro_amd/fortran/libmacproject_hip.so`` and the Fortran drivers
    ``maestro_amd/fortran/macproject_driver`` and ``mac_multigrid_driver``
    (amdflang; skipped when no Fortran compiler is present)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmaestro_mac.so")
FDIR = os.path.join(HERE, "fortran")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # only the explicit fma() calls of the stencils are fused (bitwise-stable
    # single sweeps); packed FP32 instructions stay off
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
SOURCES = ["mac_kernels.hip", "mac_inputs.hip", "mac_solve.hip", "mac_ml.hip", "mac_plan.hip"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def fortran_compiler():
    for c in ("amdflang", "/opt/rocm/bin/amdflang", "gfortran", "flang"):
        p = shutil.which(c) or (c if os.path.exists(c) else None)
        if p:
            return p
    return None


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_lib(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    out = _run(cmd)
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    return LIB


def build_fortran(verbose=False):
    fc = fortran_compiler()
    if fc is None or not os.path.exists(os.path.join(FDIR, "macproject_hip.f90")):
        return None
    mod = os.path.join(FDIR, "macproject_hip.f90")
    drv = os.path.join(FDIR, "macproject_driver.f90")
    shim = os.path.join(FDIR, "libmacproject_hip.so")
    exe = os.path.join(FDIR, "macproject_driver")
    _run([fc, "-O2", "-fPIC", "-shared", "-J", FDIR, mod, "-o", shim, f"-L{HERE}",
          "-lmaestro_mac", f"-Wl,-rpath,{HERE}"], cwd=FDIR)
    for src, out in ((drv, exe), (os.path.join(FDIR, "mac_multigrid_driver.f90"),
                                  os.path.join(FDIR, "mac_multigrid_driver")),
                     (os.path.join(FDIR, "macproject_ml_driver.f90"),
                      os.path.join(FDIR, "macproject_ml_driver"))):
        if os.path.exists(src):
            _run([fc, "-O2", "-J", FDIR, src, "-o", out, f"-L{FDIR}", "-lmacproject_hip",
                  f"-L{HERE}", "-lmaestro_mac", f"-Wl,-rpath,{FDIR}", f"-Wl,-rpath,{HERE}",
                  "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."], cwd=FDIR)
    return shim


def build(verbose=False):
    build_lib(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB, file=sys.stderr)
