"""Compile libdyc.so (the HIP kernels and the C ABI of include/dyc.h) for gfx950, in-tree."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdyc.so")
SOURCES = ["dyc_kernels.hip", "dyc_bulk.hip", "dyc_table.hip", "dyc_group.hip", "dyc_server.hip", "dyc_wide.hip"]
ARCH = os.environ.get("DYC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: libdyc.so cannot be built")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # packed fp32 instructions stay off in all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


def build(verbose: bool = False) -> str:
    """Compile every source and link libdyc.so; always rebuilds."""
    objs = []
    cc = hipcc()
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_probe(src: str, out: str, verbose: bool = False) -> str:
    """Compile a stand-alone HIP probe library (not part of libdyc) with the library's flags."""
    tmp = out + ".tmp"
    cmd = [hipcc(), *flags(), "-shared", src, "-o", tmp]
    subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
