"""Builds the in-tree native libraries (no JIT caches; the .so files travel with the tree).

    python -m ptilopsis_amd.build

libptl_hip.so   hipcc --offload-arch=gfx950 (HIP kernels + C ABI)
libptl_host.so  g++ (front end + program generator)
oracle          make -C oracle (test infrastructure; includes oracle/_ref when the
                reference checkout is present)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_hip() -> str:
    out = os.path.join(HERE, "libptl_hip.so")
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, os.path.join(CSRC, "ptl_gpu.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    os.replace(tmp, out)
    return out


def build_host() -> str:
    out = os.path.join(HERE, "libptl_host.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", tmp,
           os.path.join(CSRC, "frontend.cpp"), os.path.join(CSRC, "progen.cpp")]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


def build_all() -> None:
    build_host()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", file=sys.stderr)
