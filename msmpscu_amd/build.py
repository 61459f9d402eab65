"""In-tree build of libmsm_hip.so (hipcc, gfx950) and of the Fortran drop-in layer (amdflang).

    python -m msmpscu_amd.build          # library + Fortran module and driver
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "libmsm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return out


SOURCES = ("msm_api.hip", "msm_kernels.hip", "msm_bricks.hip", "msm_tpa.hip", "msm_pair.hip",
           "msm_ldstab.hip", "msm_stage.hip", "msm_thermal.hip", "msm_adjoint.hip", "msm_host.cpp")


def build_library(jobs: int = 8) -> str:
    """Each translation unit compiles to its own object (in parallel; every kernel is launched
    from the unit that defines it, so no relocatable device code is needed), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

    def compile_one(f):
        obj = os.path.join(odir, os.path.splitext(f)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, f), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(SOURCES)))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> list:
    """Fortran drop-in module + test driver, linked against libmsm_hip.so."""
    outs = []
    mod_src = os.path.join(FSRC, "MD_ForceLib_HIP.F90")
    if not os.path.exists(mod_src) or shutil.which(FC) is None and not os.path.exists(FC):
        return outs
    bdir = os.path.join(FSRC, "_fbuild")
    os.makedirs(bdir, exist_ok=True)
    # the layer, then (in dependency order) the modules the example callers use: the reference's
    # type modules as the glue reads them, the CGS potential routines, the force-class glue
    objs = []
    for name in ("MD_ForceLib_HIP", "mdlib_ref_types", "w_fs_cgs", "hip_forceclass_wrappers"):
        src = os.path.join(FSRC, name + ".F90")
        obj = os.path.join(bdir, name + ".o")
        _run([FC, "-O2", "-fPIC", "-c", src, "-I" + bdir, "-module-dir", bdir, "-o", obj])
        objs.append(obj)
    for name in ("fs_w_driver", "cgs_driver", "ref_glue_driver"):
        drv = os.path.join(FSRC, name + ".F90")
        exe = os.path.join(FSRC, name)
        if os.path.exists(drv):
            _run([FC, "-O2", drv, *objs, "-I" + bdir, "-module-dir", bdir, "-L" + HERE,
                  "-lmsm_hip", "-Wl,-rpath," + HERE, "-o", exe])
            outs.append(exe)
    return outs


def build_all() -> None:
    build_library()
    build_fortran()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
