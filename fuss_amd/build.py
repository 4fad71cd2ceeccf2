"""Build the HIP library of the FUSS explicit step for gfx950 (in-tree)."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfuss_gpu.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # source-order IEEE arithmetic (no fused multiply-add contraction), the
    # rounding the reference's expressions are written for
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flang():
    for cand in (os.environ.get("FLANG"), shutil.which("flang"), "/opt/rocm/lib/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    return None


FORTRAN = os.path.join(HERE, "fortran")
DEMO = os.path.join(FORTRAN, "fuss_gpu_demo")


def build_fortran(verbose=False):
    """Compile the iso_c_binding module and the Fortran driver against the
    library (skipped when no Fortran compiler is installed)."""
    fc = flang()
    if fc is None:
        return None
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmd = [fc, "-O2", "-J", bdir, os.path.join(FORTRAN, "fuss_gpu_m.f90"), os.path.join(FORTRAN, "fuss_gpu_demo.f90"),
           "-L", HERE, "-lfuss_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", DEMO]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=bdir)
    return DEMO


def build(verbose=False, jobs=None):
    """Compile each fuss_amd/csrc/*.hip unit (in parallel) and link them into
    fuss_amd/libfuss_gpu.so (always rebuilds)."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    odir = os.path.join(CSRC, "_build")
    os.makedirs(odir, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]

    def compile_one(src):
        obj = os.path.join(odir, os.path.basename(src)[:-4] + ".o")
        cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr}")
        return obj

    jobs = jobs or max(1, min(len(srcs), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
