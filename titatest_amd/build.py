"""Build libcrt_hip.so (the gfx950 HIP kernels + C-ABI) in-tree with hipcc.

    python -m titatest_amd.build

Each source is compiled with ``hipcc --offload-arch=gfx950`` and linked into
``titatest_amd/libcrt_hip.so``.  Packed FP32 instruction selection is disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) for every HIP source.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libcrt_hip.so"
BUILD = PKG / "_build"
SOURCES = ["crt_factor.hip", "crt_post.hip", "crt_inv.hip", "crt_api.cpp"]
ARCH = os.environ.get("CRT_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.exists(cand) or cand == "hipcc"):
            return cand
    return "hipcc"


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", str(CSRC), "-I", str(PKG.parent / "include"),
           "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_fortran_host(verbose)
    return OUT


FORTRAN_HOST = PKG / "crmod_hip"


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran_host(verbose: bool = False):
    """The Fortran host (iso_c_binding module + CRMod-style driver), linked
    against libcrt_hip.so with an $ORIGIN rpath.  Skipped without a Fortran
    compiler."""
    fc = flang()
    if fc is None:
        return None
    fdir = PKG / "fortran"
    cmd = [fc, "-O2", "-J", str(BUILD), "-o", str(FORTRAN_HOST),
           str(fdir / "crt_hip_mod.f90"), str(fdir / "crmod_hip.f90"),
           "-L", str(PKG), "-lcrt_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(BUILD))
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("Fortran host build failed")
    if verbose:
        print(f"built {FORTRAN_HOST}")
    return FORTRAN_HOST


if __name__ == "__main__":
    build(verbose=True)
