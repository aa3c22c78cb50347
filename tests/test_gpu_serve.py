"""GPU tests of serving (ptl_server_*): each request's result is the same
alone and in the mix, requests that arrive while batches run join later
batches, and admission refuses what can never fit."""
import sys
import threading
import time

import numpy as np
import pytest

import oracle
from ptilopsis_amd import parse_source, workloads as W
from ptilopsis_amd.generator import PtlError, Server

pytestmark = pytest.mark.gpu
OUT = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def assert_same(gpu, ref, what=""):
    for k in OUT:
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        assert g.shape == r.shape and np.array_equal(g, r), (what, k)


def mixed_requests():
    reqs = [parse_source("function f(a : int) : int { return a; }")]
    reqs += W.serve_pool(24)
    reqs.append(W.program_of_size((W.SEED_STRESS << 32) + 3, 64000, W.STRESS))  # shared register names
    reqs.append(W.single_50k())
    return reqs


def test_gpu_serve_mix_equals_alone(gpu_ctx):
    reqs = mixed_requests()
    say("serving mix of", len(reqs), "requests,", sum(t.n for t in reqs), "nodes")
    alone = [gpu_ctx.compile(t) for t in reqs]
    srv = Server(0, max_batch=64, workers=2)
    try:
        tickets = [srv.submit(t) for t in reqs]  # all queued at once: they share batches
        got = [srv.wait(k) for k in tickets]
        st = srv.stats()
    finally:
        srv.close()
    say("stats", st)
    assert st["requests_done"] == len(reqs) and st["largest_batch"] > 1
    assert max(tm["batch_programs"] for _, tm in got) > 1
    for p, ((r, tm), a, t) in enumerate(zip(got, alone, reqs)):
        assert_same(r, a, f"request {p}")
    for p in (0, 1, 5, len(reqs) - 2):
        assert_same(got[p][0], oracle.compile_tree(reqs[p]), f"request {p} vs oracle")


def test_gpu_serve_continuous_arrivals(gpu_ctx):
    pool = W.serve_pool(48)
    alone = [gpu_ctx.compile(t)["words"] for t in pool]
    arr = W.arrivals(2000, 0.4)
    say("continuous arrivals:", len(arr), "requests over 0.4 s")
    srv = Server(0, max_batch=256, workers=2)
    tickets = [None] * len(arr)

    def submitter():
        t0 = time.perf_counter()
        for i, ta in enumerate(arr):
            while time.perf_counter() - t0 < ta:
                time.sleep(0.0001)
            tickets[i] = srv.submit(pool[i % len(pool)])

    try:
        th = threading.Thread(target=submitter)
        th.start()
        for i in range(len(arr)):
            while tickets[i] is None:
                time.sleep(0.0001)
            r, tm = srv.wait(tickets[i])
            assert np.array_equal(r["words"], alone[i % len(pool)]), i
        th.join()
        st = srv.stats()
    finally:
        srv.close()
    say("stats", st)
    assert st["requests_done"] == len(arr) and st["batches"] > 1


def test_gpu_serve_admission():
    small = parse_source("function f(a : int) : int { return a; }")
    big = W.program_of_size(0x5EED0010, 20_000)
    srv = Server(0, hbm_budget=4 << 20, max_batch=8, workers=1)  # 4 MiB: ~7.7k nodes
    try:
        with pytest.raises(PtlError) as e:
            srv.submit(big)
        assert e.value.status == -7
        r, _ = srv.wait(srv.submit(small))
        assert_same(r, oracle.compile_tree(small))
    finally:
        srv.close()
