"""Build libfullmag_fdm.so (SONAME libfullmag_fdm.so.0) for gfx950 in-tree.

    python -m fullmag_amd.build

compiles fullmag_amd/csrc/*.{hip,cpp} with hipcc into
fullmag_amd/_build/libfullmag_fdm.so.  The library has no torch or Python
dependency: it is the C-ABI drop-in for the reference's libfullmag_fdm.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB_NAME = "libfullmag_fdm.so"
LIB_PATH = os.path.join(OUT_DIR, LIB_NAME)
SONAME = "libfullmag_fdm.so.0"
# (source, object stem, extra flags): the device code is compiled once per
# storage precision; the FP32 unit also holds the precision-free host helpers
SOURCES = [
    ("kernels.hip", "kernels_f32", ["-DFM_REAL=float", "-DFM_PRIMARY"]),
    ("kernels.hip", "kernels_f64", ["-DFM_REAL=double"]),
    ("api.cpp", "api", []),
    ("newell_host.cpp", "newell_host", []),
]
ARCH = os.environ.get("FULLMAG_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _stale(obj: str, src: str) -> bool:
    """True when obj is missing or older than its source or any header."""
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    deps += [os.path.join(HERE, "..", "include", f) for f in os.listdir(os.path.join(HERE, "..", "include"))]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source and link the shared library; returns its path.
    incremental=True recompiles only objects older than their sources."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
              "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include")]
    procs = []
    for src, stem, extra in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT_DIR, stem + ".o")
        cmd = [hipcc()] + common + extra
        if src.endswith(".hip"):
            # packed FP32 instructions disabled: plain scalar FP32 only
            # kernel arguments preloaded into SGPRs (the argument-record
            # pointer then needs no load at kernel entry)
            cmd += ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-ffp-contract=fast", "-mllvm", "-amdgpu-kernarg-preload-count=16"]
        else:
            # host-only translation units: HIP runtime API, no device code
            cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-Wno-unused-value"]
        cmd += ["-c", path, "-o", obj]
        objs.append(obj)
        if incremental and not _stale(obj, path):
            continue
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        f"-Wl,-soname,{SONAME}", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
