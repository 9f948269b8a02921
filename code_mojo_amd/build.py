"""Build recipe for libbfly_hip.so (hipcc, gfx950) -- runs on a CPU-only host.

`python -m code_mojo_amd.build` (or __graft_entry__.build()) compiles the HIP
kernels and the C-ABI into code_mojo_amd/libbfly_hip.so, in-tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libbfly_hip.so"
ARCH = os.environ.get("BFLY_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "api.hip", "engine.hip", "noalloc.hip", "planner.cpp", "qasm.cpp", "jit.cpp"]

# Packed-FP32 instructions are disabled for every device pass (the host pass
# warns that it does not know the feature; that is harmless).
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the GPU strategy cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile every source to an object, link the shared library."""
    out_dir = PKG / "_build"
    out_dir.mkdir(exist_ok=True)
    # tile_ops.h as a string literal: the generated tile-pass kernels
    # (jit.cpp) hand it to hipRTC
    ops = (CSRC / "tile_ops.h").read_text()
    if ")BFLYSRC\"" in ops:
        raise RuntimeError("tile_ops.h contains the raw-string delimiter")
    (out_dir / "tile_ops_src.inc").write_text('R"BFLYSRC(' + ops + ')BFLYSRC"\n')
    objs = []
    cc = hipcc()
    jobs = []
    for src in SOURCES:
        obj = out_dir / (src + ".o")
        cmd = [cc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = [cc, "-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{out_dir}", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(str(obj))
    # the translation units compile concurrently (one process each)
    failed = []
    for src, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            failed.append(f"compile failed for {src}:\n{out}\n{err}")
    if failed:
        raise RuntimeError("\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp), "-L/opt/rocm/lib", "-lhiprtc",
           "-pthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
