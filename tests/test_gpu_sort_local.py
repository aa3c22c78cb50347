"""The device depth sort on chosen key patterns (ptl_debug_depth_sort) against the
oracle's libstdc++ introsort restatement, range by range: sizes on both sides of
the 2048-element ranges finished in LDS and of the insertion-sort threshold,
all-equal and two-valued keys (closed-form placement, stops on both sides),
sorted, reversed and periodic inputs, and many ranges of one batch."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def check(ctx, keys, sizes):
    bounds = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    say("depth sort", len(sizes), "ranges", int(bounds[-1]), "keys")
    got = ctx.depth_sort(keys, bounds)
    for p in range(len(sizes)):
        b, e = int(bounds[p]), int(bounds[p + 1])
        want = oracle.introsort(keys[b:e]) + b
        bad = np.nonzero(got[b:e] != want)[0]
        assert bad.size == 0, f"range {p} (size {e - b}): {bad.size} positions differ, first at {bad[0]}"


SIZES = [1, 16, 17, 18, 100, 1000, 2047, 2048, 2049, 2050, 4095, 4096, 4097, 6000, 33000]


@pytest.mark.parametrize("pattern", ["depths", "equal", "two", "ascending", "descending", "sawtooth16", "organ"])
def test_gpu_depth_sort_patterns(gpu_ctx, pattern):
    rng = np.random.default_rng(sum(map(ord, pattern)))
    parts = []
    for n in SIZES:
        i = np.arange(n)
        k = {"depths": rng.integers(0, 70, n),
             "equal": np.full(n, 7),
             "two": rng.integers(0, 2, n),
             "ascending": i,
             "descending": n - i,
             "sawtooth16": i % 16,
             "organ": np.minimum(i, n - 1 - i)}[pattern]
        parts.append(k.astype(np.int32))
    check(gpu_ctx, np.concatenate(parts), SIZES)


def test_gpu_depth_sort_many_ranges(gpu_ctx):
    # 3000 ranges of 17..5000 keys (most finished in LDS), depth-like keys
    rng = np.random.default_rng(0x5EED0501)
    sizes = rng.integers(17, 5000, 3000)
    keys = rng.integers(0, 70, int(sizes.sum())).astype(np.int32)
    check(gpu_ctx, keys, sizes)


def test_gpu_depth_sort_refuses_bad_bounds(gpu_ctx):
    from ptilopsis_amd.generator import PtlError
    with pytest.raises(PtlError):
        gpu_ctx.depth_sort(np.zeros(4, np.int32), [0, 2, 2, 4])  # an empty range
