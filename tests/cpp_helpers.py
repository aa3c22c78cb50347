"""Compiles small C++ helpers with the host g++ for cross-checks against the
real C++ standard library (std::sort, std::mt19937_64, __builtin_ffs)."""
import os
import shutil
import subprocess
import tempfile

import pytest

_cache = {}


def compile_cpp(name: str, source: str, extra=()) -> str:
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if name in _cache:
        return _cache[name]
    d = tempfile.mkdtemp(prefix="ptl_cpp_")
    src = os.path.join(d, name + ".cpp")
    exe = os.path.join(d, name)
    with open(src, "w") as f:
        f.write(source)
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, src, *extra], check=True, capture_output=True)
    _cache[name] = exe
    return exe
