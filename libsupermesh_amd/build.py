"""Builds the HIP library libsupermesh_amd/lib/libsupermesh_gpu.so for gfx950.

Every translation unit is compiled with -ffp-contract=off (no FMA
contraction: the clippers must round exactly like the reference's Fortran)
and with packed FP32 instructions disabled.  Objects are built in parallel,
then linked into one shared library.  ``python -m libsupermesh_amd.build``.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIBNAME = "libsupermesh_gpu.so"
SOURCES = ["finder.hip", "supermesh.hip", "predictor.hip", "graphs.hip", "capi.hip", "interval.cpp", "advancing_front.cpp"]
ARCH = os.environ.get("LIBSUPERMESH_GPU_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _stale(src, obj):
    """True when obj is missing or older than src or any header it may include."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps += [os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, only_stale=False):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    if only_stale and not _stale(src, obj):
        return obj
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_fortran(lib_path, verbose=True):
    """The Fortran interface module (ISO_C_BINDING over the C ABI) and the
    Fortran driver used by the tests; skipped when amdflang is absent."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran interface not built")
        return None
    src = os.path.join(PKG, "fortran", "libsupermesh.F90")
    mod_dir = os.path.join(LIBDIR, "fortran")
    os.makedirs(mod_dir, exist_ok=True)
    flib = os.path.join(LIBDIR, "libsupermesh_fortran.so")
    cmd = [fc, "-O2", "-fPIC", "-shared", "-module-dir", mod_dir, src, "-o", flib,
           "-L" + LIBDIR, "-lsupermesh_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    test_src = os.path.join(ROOT, "tests", "fortran", "test_intersector_gpu.F90")
    exe = os.path.join(LIBDIR, "test_intersector_gpu")
    test_mod_dir = os.path.join(mod_dir, "test")
    os.makedirs(test_mod_dir, exist_ok=True)
    cmd = [fc, "-O2", "-I", mod_dir, "-module-dir", test_mod_dir, test_src, "-o", exe, "-L" + LIBDIR,
           "-lsupermesh_fortran", "-lsupermesh_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {flib} and {exe}")
    return flib


def build(verbose=True, only_stale=False):
    """Compiles every source (only_stale: only those older than their sources
    or headers) and links the library and the Fortran interface."""
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda src: _compile(src, only_stale), SOURCES))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    build_fortran(out, verbose)
    return out


if __name__ == "__main__":
    build(only_stale="--stale" in sys.argv)
    sys.exit(0)
