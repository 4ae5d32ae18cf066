"""Build of the HIP library ``libstda_hip.so`` (gfx950), in-tree.

``python -m stda_amd.build`` or :func:`build` compiles the kernels and the
host orchestration with hipcc and links rocBLAS (plain FP64 GEMMs); the host
LAPACK used for the small Rayleigh-Ritz eigensolves is opened at run time.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstda_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["kernels.hip", "batch.hip", "spectrum.hip", "hyper.hip", "transform.hip", "core.cpp"]
HIPFLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
            # packed FP32 instructions stay off (hipcc's host pass warns about the feature)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPFLAGS, "-I", CSRC,
               "-I", os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose:
            sys.stderr.write(r.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC",
           *objs, "-o", tmp, "-L", os.path.join(ROCM, "lib"), "-lrocblas", "-lrocsolver",
           "-ldl", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libstda_hip.so failed")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
