"""GPU parity at the benchmark configurations' full sizes.

large_16m     one program of ~17M nodes and 20k functions, compiled by the plain
              pipeline and compared word for word with the oracle.
batch_4096    all 4096 programs (~100M nodes) in one ptl_compile_batch; every
              program compared with the oracle run on that program alone.
stress_1024   all 1024 regalloc/jump stress programs (~65M nodes) in one batch;
              every program compared with the oracle alone.

The oracle runs on a pool of host threads (it is thread-safe per compile and
ctypes releases the GIL), one program per task.  Besides the oracle, every
result is checked for properties fixed by the output format: the function
table tiles the words in order, sizes are end - start, and every function
ends with `ret` (jalr x0, 0(ra)).
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from ptilopsis_amd import workloads as W

pytestmark = pytest.mark.gpu
OUT = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")
RET = 0x00008067  # jalr x0, 0(x1)
THREADS = max(1, min(16, os.cpu_count() or 1) - 1)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def first_difference(gpu, ref):
    for k in OUT:
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        if g.shape == r.shape and np.array_equal(g, r):
            continue
        m = min(g.shape[0], r.shape[0])
        idx = np.nonzero(g[:m] != r[:m])[0]
        i = int(idx[0]) if idx.size else m
        return {"field": k, "index": i, "got": int(g[i]) if i < g.shape[0] else -1,
                "want": int(r[i]) if i < r.shape[0] else -1, "got_len": g.shape[0], "want_len": r.shape[0]}
    return None


def check_format(r, what):
    """The function table tiles the words and every function returns last."""
    w, s, e, z = (np.asarray(r[k]) for k in ("words", "func_starts", "func_ends", "func_sizes"))
    assert s.shape[0] == r["n_funcs"] > 0, what
    assert s[0] == 0 and e[-1] == w.shape[0], what
    assert np.array_equal(s[1:], e[:-1]) and np.array_equal(z, e - s), what
    assert np.all(w[e - 1] == RET), what


def against_oracle(trees, res):
    """Oracle on every program (host thread pool); returns the mismatches and
    the oracle's wall time."""
    t0 = time.time()

    def one(p):
        return p, first_difference(res[p], oracle.compile_tree(trees[p]))

    with ThreadPoolExecutor(THREADS) as ex:
        bad = [(p, d) for p, d in ex.map(one, range(len(trees))) if d is not None]
    return bad, time.time() - t0


def run_batch(gpu_ctx, name, trees):
    n = sum(t.n for t in trees)
    say(f"{name}: {len(trees)} programs, {n} nodes; one compile_batch")
    t0 = time.time()
    res = gpu_ctx.compile_batch(trees)
    say(f"{name}: compiled in {time.time() - t0:.2f} s (with copies)")
    for p, r in enumerate(res):
        check_format(r, f"{name} program {p}")
    bad, el = against_oracle(trees, res)
    say(f"{name}: oracle on all {len(trees)} programs, {THREADS} threads, {el:.1f} s; mismatches {len(bad)}")
    assert not bad, bad[:4]


def test_gpu_large_16m_whole(gpu_ctx):
    t0 = time.time()
    tree = W.large_16m()
    say(f"large_16m: {tree.n} nodes (built in {time.time() - t0:.1f} s)")
    r = gpu_ctx.compile(tree)
    check_format(r, "large_16m")
    assert r["n_funcs"] == 20000
    t0 = time.time()
    want = oracle.compile_tree(tree)
    say(f"large_16m: oracle {time.time() - t0:.1f} s, {r['n_words']} words")
    d = first_difference(r, want)
    assert d is None, d


def test_gpu_batch_4096_whole(gpu_ctx):
    t0 = time.time()
    trees = W.batch_4096()
    say(f"batch_4096 built in {time.time() - t0:.1f} s")
    assert len(trees) == 4096
    run_batch(gpu_ctx, "batch_4096", trees)


def test_gpu_stress_1024_whole(gpu_ctx):
    t0 = time.time()
    trees = W.stress_1024()
    say(f"stress_1024 built in {time.time() - t0:.1f} s")
    assert len(trees) == 1024
    run_batch(gpu_ctx, "stress_1024", trees)
