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


def first_difference(gpu, ref):
    """None when every output array is equal, else the first differing field,
    index and values (-1: past the end)."""
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


def assert_same(gpu, ref, what=""):
    d = first_difference(gpu, ref)
    assert d is None, (what, d)


def check_served(got, alone, tickets, nodes):
    """Every served result equals its alone compile; on a mismatch the message
    names the request, its ticket, batch, worker, the batch's members and the
    first differing word."""
    members = {}
    for i, (_, tm) in enumerate(got):
        members.setdefault(tm["batch_id"], []).append((tickets[i], i))
    bad = []
    for i, ((r, tm), a) in enumerate(zip(got, alone)):
        d = first_difference(r, a)
        if d is not None:
            d.update(request=i, ticket=tickets[i], nodes=nodes[i], batch_id=tm["batch_id"], worker=tm["worker"],
                     batch_programs=tm["batch_programs"], batch_members=members[tm["batch_id"]])
            say("MISMATCH", d)
            bad.append(d)
    assert not bad, bad


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
    check_served(got, alone, tickets, [t.n for t in reqs])
    for p in (0, 1, 5, len(reqs) - 2):
        assert_same(got[p][0], oracle.compile_tree(reqs[p]), f"request {p} vs oracle")


def test_gpu_serve_continuous_arrivals(gpu_ctx):
    pool = W.serve_pool(48)
    alone = [gpu_ctx.compile(t) for t in pool]
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

    got = []
    try:
        th = threading.Thread(target=submitter)
        th.start()
        for i in range(len(arr)):
            while tickets[i] is None:
                time.sleep(0.0001)
            got.append(srv.wait(tickets[i]))
        th.join()
        st = srv.stats()
    finally:
        srv.close()
    say("stats", st)
    check_served(got, [alone[i % len(pool)] for i in range(len(arr))], tickets,
                 [pool[i % len(pool)].n for i in range(len(arr))])
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


def test_gpu_serve_records_and_arrays_mixed(gpu_ctx):
    """Requests given as parser records (tree built on the device inside the
    batch) and as DepthTree arrays share batches; each result equals the alone
    compile of the flattened tree.  A request with malformed records fails
    alone: its batch is rerun one request at a time."""
    from ptilopsis_amd import parse_arena
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in np.clip(rng.lognormal(np.log(4000), 1.5, 30), 200, 80_000)]
    srcs = [W.source_of_size((W.SEED_REQUESTS << 32) + 5000 + i, s) for i, s in enumerate(sizes)]
    trees = [parse_source(s) for s in srcs]
    recs = [parse_arena(s) for s in srcs]
    alone = [gpu_ctx.compile(t) for t in trees]
    bad = parse_arena(srcs[0])
    bad.kids = bad.kids.copy()
    bad.kids[3] = bad.kids[4]  # one node listed twice
    srv = Server(0, max_batch=64, workers=2)
    try:
        # all queued at once, every other one as records: they share batches
        tickets = [srv.submit_records(recs[i]) if i % 2 == 0 else srv.submit(trees[i]) for i in range(len(srcs))]
        got = [srv.wait(k) for k in tickets]
        assert max(tm["batch_programs"] for _, tm in got) > 1
        check_served(got, alone, tickets, [t.n for t in trees])
        # the malformed request among good ones
        some = list(range(0, 12))
        t2 = [srv.submit_records(recs[i]) for i in some[:6]] + [srv.submit_records(bad)] + \
             [srv.submit(trees[i]) for i in some[6:]]
        with pytest.raises(PtlError) as e:
            srv.wait(t2[6])
        assert e.value.status == -2
        got2 = [srv.wait(k) for k in t2[:6] + t2[7:]]
        check_served(got2, [alone[i] for i in some], t2[:6] + t2[7:], [trees[i].n for i in some])
        st = srv.stats()
    finally:
        srv.close()
    say("stats", st)
