"""Build recipe for the native parts of electricbubbles_amd.

* ``libeb.so``          HIP kernels + C-ABI (hipcc, gfx950 only)
* ``libeb_fortran.so``  ISO_C_BINDING shim exporting the reference names CADAPDR / CADAPMOD /
                        CADAP (ROCm flang), linked against libeb.so

Both land in this package directory (in-tree), so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

HIP_SOURCES = ["eb_sort.hip", "eb_fmm.hip", "eb_spectral.hip", "eb_atree.hip", "eb_gmres.hip", "eb_field.hip", "eb_sim.hip",
               "eb_adjoint.hip", "eb_dropin.cpp"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"),
]
# per-source additions to the flags above
HIP_EXTRA: dict = {}


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_hip(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libeb.so")
    objs, cmds = [], []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, *HIP_FLAGS, *HIP_EXTRA.get(src, []), *lang, "-c", path, "-o", obj])
        objs.append(obj)
    # the translation units are independent: compile them side by side (MAX_JOBS caps it)
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING shim (needs ROCm flang; skipped if absent)."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran", "eb_cadap_shim.f90")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "libeb_fortran.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    obj = os.path.join(objdir, "eb_cadap_shim.o")
    _run([FLANG, "-O2", "-fPIC", "-c", src, "-o", obj, "-module-dir", objdir])
    tmp = out + ".tmp"
    _run([FLANG, "-shared", "-fPIC", obj, "-o", tmp, f"-L{HERE}", "-leb", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_fortran(verbose)


def clean() -> None:
    shutil.rmtree(os.path.join(HERE, "_build"), ignore_errors=True)
    for f in ("libeb.so", "libeb_fortran.so"):
        p = os.path.join(HERE, f)
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    build(verbose=True)
