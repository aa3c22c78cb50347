"""The C ABI library loads and exports every entry point include/ptl.h declares."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_SO = os.path.join(ROOT, "ptilopsis_amd", "libptl_hip.so")


def declared():
    txt = open(os.path.join(ROOT, "include", "ptl.h")).read()
    return sorted(set(re.findall(r"\b(ptl_[a-z_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    names = declared()
    for must in ("ptl_create", "ptl_compile", "ptl_compile_batch", "ptl_destroy", "ptl_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(HIP_SO):
        pytest.skip("libptl_hip.so not built (run __graft_entry__.build())")
    lib = C.CDLL(HIP_SO)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing
    from ptilopsis_amd import abi
    assert set(abi.EXPORTS) == set(declared())


def test_strerror_without_a_device():
    if not os.path.exists(HIP_SO):
        pytest.skip("libptl_hip.so not built")
    lib = C.CDLL(HIP_SO)
    lib.ptl_strerror.restype = C.c_char_p
    assert lib.ptl_strerror(0) == b"ok"
    assert b"DepthTree" in lib.ptl_strerror(-2)


def test_entry_points_refuse_a_null_context_without_a_device():
    # argument checks come before any HIP call, so they hold on a host with no GPU
    if not os.path.exists(HIP_SO):
        pytest.skip("libptl_hip.so not built")
    from ptilopsis_amd import abi
    lib = abi.bind(C.CDLL(HIP_SO))
    keys = (C.c_int32 * 4)(3, 1, 2, 0)
    bounds = (C.c_uint32 * 2)(0, 4)
    out = (C.c_uint32 * 4)()
    arg = next(k for k, v in abi.ERRORS.items() if v == "PTL_ERR_ARG")
    assert lib.ptl_debug_depth_sort(None, keys, bounds, 1, out) == arg
    assert lib.ptl_debug_exclusive_scan(None, out, out, 4) == arg
    assert lib.ptl_debug_overlap(None, 1) == arg
    r = abi.PtlResult()
    assert lib.ptl_compile_noalloc(None, None, None, 0, None, None, None, None, 0, C.byref(r)) == arg
