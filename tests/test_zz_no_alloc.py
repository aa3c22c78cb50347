"""ptl_compile_noalloc through the C ABI: the same result as ptl_compile on the
same tree, on two calls in a row (the second on a smaller input, so the
scratch array holds a larger call's leftovers), plus its refusals: an input
above its node limit, an output buffer too small, an invalid tree.  Device
buffers for the inputs and outputs come from the HIP runtime (the test's
plumbing; the library itself allocates nothing in these calls)."""
import ctypes as C
import sys

import numpy as np
import pytest

from ptilopsis_amd import abi, parse_source, workloads as W
from ptilopsis_amd.native import hip_lib

pytestmark = pytest.mark.gpu
OUT = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")
LIMIT = 1 << 18


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def status(name):
    return next(k for k, v in abi.ERRORS.items() if v == name)


class Hip:
    """The HIP runtime's allocation and copy calls (the library's own runtime,
    already loaded in this process): the test's plumbing for device buffers."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(4, nbytes)) == 0
        self.live.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        # the library works on its own non-blocking stream: the copy must be complete
        assert self.rt.hipDeviceSynchronize() == 0
        return p

    def download(self, p, n):
        out = np.zeros(n, np.uint32)
        if n:
            assert self.rt.hipMemcpy(out.ctypes.data, p, 4 * n, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


@pytest.fixture
def hip():
    h = Hip()
    yield h
    h.free()


class DevTree:
    """The six arrays of a tree in device memory, and its ptl_ast."""

    def __init__(self, hip, tree, n_nodes=None, max_depth=None):
        ptrs = [hip.upload(a) for a in tree.arrays()]
        self.ast = abi.PtlAst(*[p.value for p in ptrs], n_nodes if n_nodes is not None else tree.n,
                              max_depth if max_depth is not None else tree.max_depth)


def compile_noalloc(hip, ctx, dt, words_cap, funcs_cap):
    w = hip.alloc(4 * max(1, words_cap))
    f = [hip.alloc(4 * max(1, funcs_cap)) for _ in range(4)]
    r = abi.PtlResult()
    s = hip_lib().ptl_compile_noalloc(ctx.h, C.byref(dt.ast), w, words_cap, *f, funcs_cap, C.byref(r))
    if s != abi.PTL_OK:
        return s, r, None
    nw, nf = int(r.n_words), int(r.n_funcs)
    out = {"words": hip.download(w, nw)}
    for k, x in zip(OUT[1:], f):
        out[k] = hip.download(x, nf)
    return s, r, out


def assert_same(got, want, what):
    for k in OUT:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert g.shape == w.shape and np.array_equal(g, w), (what, k)


def test_gpu_noalloc_matches_compile_on_two_calls(gpu_ctx, hip):
    big = W.program_of_size(0x5EED0A11, 240_000)
    small = W.single_50k()
    assert small.n < big.n < LIMIT
    for name, tree in (("240k", big), ("50k after 240k", small)):
        say(f"noalloc: {name} ({tree.n} nodes)")
        want = gpu_ctx.compile(tree)
        dt = DevTree(hip, tree)
        s, r, got = compile_noalloc(hip, gpu_ctx, dt, want["n_words"] + 16, want["n_funcs"] + 4)
        assert s == abi.PTL_OK, abi.ERRORS.get(s, s)
        assert (int(r.n_words), int(r.n_funcs)) == (want["n_words"], want["n_funcs"])
        assert (int(r.n_words_gen), int(r.n_words_alloc)) == (want["W0"], want["W1"])
        assert_same(got, want, name)
    # a compile through the context's own arena still works after the calls
    assert_same(gpu_ctx.compile(small), want, "ptl_compile after noalloc")


def test_gpu_noalloc_register_named_by_two_functions(gpu_ctx, hip):
    # a stress program (~172k nodes) with registers named by two functions:
    # the exact serial allocation path
    tree = W.program_of_size((W.SEED_STRESS << 32) + 3, 30000, W.STRESS)
    assert tree.n < LIMIT
    say(f"noalloc: stress program ({tree.n} nodes)")
    gpu_ctx.profile(True)
    want = gpu_ctx.compile(tree)
    names = gpu_ctx.kernel_times()
    gpu_ctx.profile(False)
    assert "k_ra_serial" in names
    s, r, got = compile_noalloc(hip, gpu_ctx, DevTree(hip, tree), want["n_words"], want["n_funcs"])
    assert s == abi.PTL_OK, abi.ERRORS.get(s, s)
    assert_same(got, want, "stress")


def test_gpu_noalloc_refusals(gpu_ctx, hip):
    tree = parse_source("function f(a : int) : int { return a + 1; }")
    want = gpu_ctx.compile(tree)
    dt = DevTree(hip, tree)
    say("noalloc: output too small")
    s, r, _ = compile_noalloc(hip, gpu_ctx, dt, want["n_words"] - 1, want["n_funcs"])
    assert s == status("PTL_ERR_ARG") and int(r.n_words) == want["n_words"]
    say("noalloc: above the node limit (refused before any device work)")
    s, _, _ = compile_noalloc(hip, gpu_ctx, DevTree(hip, tree, n_nodes=LIMIT + 1), 64, 4)
    assert s == status("PTL_ERR_TOO_LARGE")
    say("noalloc: wrong max_depth (checked on the device)")
    s, _, _ = compile_noalloc(hip, gpu_ctx, DevTree(hip, tree, max_depth=tree.max_depth + 1), 64, 4)
    assert s == status("PTL_ERR_INVALID_AST")
    say("noalloc: a valid call after the refusals")
    s, _, got = compile_noalloc(hip, gpu_ctx, dt, want["n_words"], want["n_funcs"])
    assert s == abi.PTL_OK
    assert_same(got, want, "after refusals")
