"""ptl_compile_noalloc through the C ABI: the same result as ptl_compile on the
same tree, on two calls in a row (the second on a smaller input, so the
scratch array holds a larger call's leftovers), plus its refusals: an input
above its node limit, an output buffer too small, an invalid tree.  Device
buffers for the inputs and outputs come from torch (the test's plumbing; the
library itself allocates nothing in these calls)."""
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


class DevTree:
    """The six arrays of a tree in device memory, and its ptl_ast."""

    def __init__(self, tree, n_nodes=None, max_depth=None):
        import torch
        arrs = [np.ascontiguousarray(a) for a in tree.arrays()]
        self.keep = []
        for a in arrs:
            t = torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to("cuda")
            self.keep.append(t)
        self.ast = abi.PtlAst(*[t.data_ptr() for t in self.keep], n_nodes if n_nodes is not None else tree.n,
                              max_depth if max_depth is not None else tree.max_depth)


def compile_noalloc(ctx, dt, words_cap, funcs_cap):
    import torch
    w = torch.zeros(max(1, words_cap), dtype=torch.int32, device="cuda")
    f = [torch.zeros(max(1, funcs_cap), dtype=torch.int32, device="cuda") for _ in range(4)]
    r = abi.PtlResult()
    # the library runs on its own non-blocking stream: torch's uploads and
    # fills must be complete first
    torch.cuda.synchronize()
    s = hip_lib().ptl_compile_noalloc(ctx.h, C.byref(dt.ast), w.data_ptr(), words_cap,
                                      *[x.data_ptr() for x in f], funcs_cap, C.byref(r))
    if s != abi.PTL_OK:
        return s, r, None
    torch.cuda.synchronize()
    nw, nf = int(r.n_words), int(r.n_funcs)
    out = {"words": w[:nw].cpu().numpy().view(np.uint32)}
    for k, x in zip(OUT[1:], f):
        out[k] = x[:nf].cpu().numpy().view(np.uint32)
    return s, r, out


def assert_same(got, want, what):
    for k in OUT:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert g.shape == w.shape and np.array_equal(g, w), (what, k)


def test_gpu_noalloc_matches_compile_on_two_calls(gpu_ctx):
    big = W.program_of_size(0x5EED0A11, 240_000)
    small = W.single_50k()
    assert small.n < big.n < LIMIT
    for name, tree in (("240k", big), ("50k after 240k", small)):
        say(f"noalloc: {name} ({tree.n} nodes)")
        want = gpu_ctx.compile(tree)
        dt = DevTree(tree)
        s, r, got = compile_noalloc(gpu_ctx, dt, want["n_words"] + 16, want["n_funcs"] + 4)
        assert s == abi.PTL_OK, abi.ERRORS.get(s, s)
        assert (int(r.n_words), int(r.n_funcs)) == (want["n_words"], want["n_funcs"])
        assert (int(r.n_words_gen), int(r.n_words_alloc)) == (want["W0"], want["W1"])
        assert_same(got, want, name)
    # a compile through the context's own arena still works after the calls
    assert_same(gpu_ctx.compile(small), want, "ptl_compile after noalloc")


def test_gpu_noalloc_register_named_by_two_functions(gpu_ctx):
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
    s, r, got = compile_noalloc(gpu_ctx, DevTree(tree), want["n_words"], want["n_funcs"])
    assert s == abi.PTL_OK, abi.ERRORS.get(s, s)
    assert_same(got, want, "stress")


def test_gpu_noalloc_refusals(gpu_ctx):
    tree = parse_source("function f(a : int) : int { return a + 1; }")
    want = gpu_ctx.compile(tree)
    dt = DevTree(tree)
    say("noalloc: output too small")
    s, r, _ = compile_noalloc(gpu_ctx, dt, want["n_words"] - 1, want["n_funcs"])
    assert s == status("PTL_ERR_ARG") and int(r.n_words) == want["n_words"]
    say("noalloc: above the node limit (refused before any device work)")
    s, _, _ = compile_noalloc(gpu_ctx, DevTree(tree, n_nodes=LIMIT + 1), 64, 4)
    assert s == status("PTL_ERR_TOO_LARGE")
    say("noalloc: wrong max_depth (checked on the device)")
    s, _, _ = compile_noalloc(gpu_ctx, DevTree(tree, max_depth=tree.max_depth + 1), 64, 4)
    assert s == status("PTL_ERR_INVALID_AST")
    say("noalloc: a valid call after the refusals")
    s, _, got = compile_noalloc(gpu_ctx, dt, want["n_words"], want["n_funcs"])
    assert s == abi.PTL_OK
    assert_same(got, want, "after refusals")
