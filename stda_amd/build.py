"""Build of the HIP library (gfx950), in-tree.

``python -m stda_amd.build`` or :func:`build` compiles the kernels and the
host orchestration with hipcc and links rocBLAS/rocSOLVER; the host LAPACK
used for the small subspace problems is opened at run time.  Two libraries
share the kernel objects and differ in the index type of the C ABI
(``stda_hip_int``): ``libstda_hip.so`` (32-bit iaoat/iconf) and
``libstda_hip_ilp64.so`` (64-bit, for a Fortran driver built with 8-byte
default integers).  Then the ISO_C_BINDING module ``fortran/stda_hip.F90`` and
the Fortran harness ``fortran/tda_harness.F90`` are compiled with ROCm's
flang into ``fortran/tda_harness`` (linked to the 32-bit library) and
``fortran/tda_harness_ilp64`` (``-fdefault-integer-8``, linked to the 64-bit
one), when flang is present.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstda_hip.so")
LIB_ILP64 = os.path.join(HERE, "libstda_hip_ilp64.so")
FORTRAN = os.path.join(ROOT, "fortran")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
FLANG = os.environ.get("FLANG", os.path.join(ROCM, "lib", "llvm", "bin", "flang"))

KERNEL_SOURCES = ["kernels.hip", "gemm.hip", "batch.hip", "wide.hip", "spectrum.hip",
                  "hyper.hip", "transform.hip"]
HOST_SOURCE = "core.cpp"
HIPFLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
            # packed FP32 instructions stay off (hipcc's host pass warns about the feature)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _run(cmd, what, verbose):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed")
    if verbose:
        sys.stderr.write(r.stderr)


def _compile(src, obj, defines=(), verbose=False):
    cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPFLAGS, *defines, "-I", CSRC,
           "-I", os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    _run(cmd, f"hipcc on {src}", verbose)


def _link(objs, out, verbose=False):
    tmp = out + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC",
           *objs, "-o", tmp, "-L", os.path.join(ROCM, "lib"), "-lrocblas", "-lrocsolver",
           "-ldl", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    _run(cmd, f"link of {os.path.basename(out)}", verbose)
    os.replace(tmp, out)


def build_fortran(verbose: bool = False) -> list:
    """The ISO_C_BINDING module and the harness, both integer sizes (needs flang)."""
    if not os.path.exists(FLANG):
        return []
    out = []
    for tag, flags, lib in (("", [], "stda_hip"),
                            ("_ilp64", ["-fdefault-integer-8", "-DSTDA_HIP_ILP64"], "stda_hip_ilp64")):
        moddir = os.path.join(FORTRAN, "_mod" + tag)
        os.makedirs(moddir, exist_ok=True)
        exe = os.path.join(FORTRAN, "tda_harness" + tag)
        cmd = [FLANG, "-O2", *flags, "-module-dir", moddir,
               os.path.join(FORTRAN, "stda_hip.F90"), os.path.join(FORTRAN, "tda_harness.F90"),
               "-o", exe + ".tmp", "-L", HERE, f"-l{lib}", "-Wl,-rpath,$ORIGIN/../stda_amd",
               f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
        _run(cmd, f"flang build of tda_harness{tag}", verbose)
        os.replace(exe + ".tmp", exe)
        shutil.rmtree(moddir, ignore_errors=True)
        out.append(exe)
    return out


def build(verbose: bool = False) -> str:
    objs = []
    for src in KERNEL_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        _compile(src, obj, verbose=verbose)
        objs.append(obj)
    host32 = os.path.join(CSRC, "core.o")
    host64 = os.path.join(CSRC, "core_ilp64.o")
    _compile(HOST_SOURCE, host32, verbose=verbose)
    _compile(HOST_SOURCE, host64, defines=["-DSTDA_HIP_ILP64"], verbose=verbose)
    _link(objs + [host32], LIB, verbose)
    _link(objs + [host64], LIB_ILP64, verbose)
    for o in objs + [host32, host64]:
        os.remove(o)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
