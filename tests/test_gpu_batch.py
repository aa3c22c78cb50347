"""GPU parity of the batched path (ptl_compile_batch and the resident batch
slot): every program of a batch against the oracle run on that program alone,
bit-exact, and against the same library compiling it alone.  Batches mix tiny
programs (one insertion-sort leaf), programs with equal-depth writers (the
depth sort), programs of many lane-0 groups and stress programs whose
registers are named by two functions (the exact serial allocation), so every
per-program table of the pipeline is exercised next to other programs."""
import sys

import numpy as np
import pytest

import oracle
from ptilopsis_amd import parse_source, workloads as W
from ptilopsis_amd.generator import PtlError

pytestmark = pytest.mark.gpu
OUT = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def assert_same(gpu, ref, what=""):
    for k in OUT:
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        assert g.shape == r.shape, (what, k, g.shape, r.shape)
        bad = np.nonzero(g != r)[0]
        assert bad.size == 0, f"{what} {k}: {bad.size} differ, first at {bad[0]}: {g[bad[0]]:#x} != {r[bad[0]]:#x}"


def mixed_batch():
    trees = [parse_source("function f() : void { }"),
             parse_source("function f(a : int) : int { return a; }")]
    for s in range(5):
        trees.append(parse_source(W.program_source(300 + s, 3 + 4 * s)[0]))
    trees.append(W.program_of_size((W.SEED_STRESS << 32) + 3, 64000, W.STRESS))  # shared register names
    trees.append(W.single_50k())                                                   # equal-depth writers
    trees.append(parse_source(W.program_source(31, 300, min_stmts=1, max_stmts=3)[0]))  # 38 lane-0 groups
    trees.append(parse_source(W.program_source(201, 8, base=dict(W.BASE, float_frac=0.9))[0]))
    trees.append(parse_source("function f(a : float, b : int) : float { while a < 2.0 { a = a + 1.0; } return a; }"))
    return trees


def check_batch(res, trees, alone=None):
    assert len(res) == len(trees)
    for p, (r, t) in enumerate(zip(res, trees)):
        assert_same(r, oracle.compile_tree(t), f"program {p} ({t.n} nodes)")
        if alone is not None:
            a = alone[p]
            assert (r["W0"], r["W1"], r["n_words"], r["n_funcs"]) == (a["W0"], a["W1"], a["n_words"], a["n_funcs"]), p


def test_gpu_batch_mixed(gpu_ctx):
    trees = mixed_batch()
    say("mixed batch", len(trees), "programs,", sum(t.n for t in trees), "nodes")
    gpu_ctx.profile(True)
    res = gpu_ctx.compile_batch(trees)
    names = gpu_ctx.kernel_times()
    gpu_ctx.profile(False)
    assert "k_ra_serial" in names and "k_sort_root" in names
    say("alone")
    alone = [gpu_ctx.compile(t) for t in trees]
    check_batch(res, trees, alone)


def test_gpu_batch_order_and_resident(gpu_ctx):
    trees = mixed_batch()[::-1]
    say("reversed mixed batch, resident slot 7")
    gpu_ctx.upload_batch(7, trees)
    a = gpu_ctx.compile_resident_batch(7)
    b = gpu_ctx.compile_resident_batch(7)
    check_batch(a, trees)
    for x, y in zip(a, b):
        assert_same(x, y, "repeat")
    with pytest.raises(PtlError) as e:  # a batch slot is not a single tree
        gpu_ctx.compile_resident(7)
    assert e.value.status == -1


def test_gpu_batch_sample_of_4096(gpu_ctx):
    # the first 48 programs of the 4096-program workload (log-normal sizes)
    sizes = W.request_sizes(4096, W.SEED_BATCH, 8000, 1.5, 1000, 1_000_000)[:48]
    trees = [W.program_of_size((W.SEED_BATCH << 32) + p, int(s)) for p, s in enumerate(sizes)]
    say("batch sample", len(trees), "programs,", sum(t.n for t in trees), "nodes")
    check_batch(gpu_ctx.compile_batch(trees), trees)


def test_gpu_batch_sample_of_stress(gpu_ctx):
    trees = [W.program_of_size((W.SEED_STRESS << 32) + p, 64_000, W.STRESS) for p in range(6)]
    say("stress sample", len(trees), "programs,", sum(t.n for t in trees), "nodes")
    check_batch(gpu_ctx.compile_batch(trees), trees)


def test_gpu_batch_many_tiny(gpu_ctx):
    # 2000 programs of one or two small functions: thousands of one-wave
    # lane-0 workgroups and first-level sort ranges
    trees = [parse_source(W.program_source(5000 + p, 1 + p % 2, min_stmts=1, max_stmts=6)[0]) for p in range(2000)]
    say("many tiny programs", len(trees), sum(t.n for t in trees))
    res = gpu_ctx.compile_batch(trees)
    for p in range(0, 2000, 97):
        assert_same(res[p], oracle.compile_tree(trees[p]), f"program {p}")
    words = np.concatenate([r["words"] for r in res])
    ref = np.concatenate([oracle.compile_tree(t)["words"] for t in trees])
    assert np.array_equal(words, ref)


def test_gpu_batch_errors(gpu_ctx):
    good = parse_source("function f(a : int) : int { return a; }")
    empty = parse_source("")
    with pytest.raises(PtlError) as e:
        gpu_ctx.compile_batch([good, empty])
    assert e.value.status == -3
    t = W.single_50k()
    bad = W.Tree(t.node_types, t.result_types, t.parents.copy(), t.depth, t.child_idx, t.node_data, t.max_depth)
    bad.parents[10] = 5
    with pytest.raises(PtlError) as e:
        gpu_ctx.compile_batch([good, bad])
    assert e.value.status == -2
    # the context still works after a refused batch
    assert_same(gpu_ctx.compile_batch([good])[0], oracle.compile_tree(good))
