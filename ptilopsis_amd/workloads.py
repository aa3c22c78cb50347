"""Named synthetic workloads (seeded), one per benchmark configuration.

Every workload is Pareas-language source from the seeded generator
(``csrc/progen.cpp``) parsed by the front end (``csrc/frontend.cpp``), so the
code generator sees exactly what the reference front end would hand it.

single_50k     one program, ~50k nodes, 64 functions of 0-4 int/float args,
               bodies of 5-40 statements (assign 50 %, if/else 20 %, while 15 %,
               return 10 %, call 5 %), geometric expression depth (mean 4).
large_16m      one program, ~16M nodes, 20k functions.  Per-function sizes are
               drawn log-uniform over [100, 20k] and then scaled by one common
               factor so the total is 16M (E[logU(100, 20k)] * 20k would be ~75M,
               so the three stated numbers cannot all hold; node and function
               counts are kept, the range shrinks to ~[21, 4.3k]).
batch_4096     4096 programs, log-normal sizes (median 8k, sigma 1.5) clipped to
               [1k, 1M], ~100M nodes in total.
stress_1024    1024 programs of ~64k nodes (each within [48k, 80k]): 6-16 mixed
               args, wide balanced expressions of depth 10-14, chains of if/while
               nested 8-16 deep, ~30 % floats.  Functions (of ~1.5k-node targets)
               are added until the program reaches ~64k nodes.
serve_stream   request sizes log-normal (median 4k, sigma 2.0) clipped to [200, 4M].
"""
from __future__ import annotations

import math

import numpy as np

from .native import Tree, generate_source, parse_source

SEED_SINGLE = 0x5EED0001
SEED_LARGE = 0x5EED0002
SEED_BATCH = 0x5EED0003
SEED_STRESS = 0x5EED0004
SEED_ARRIVALS = 0x5EED0005
SEED_REQUESTS = 0x5EED0006

BASE = dict(
    min_args=0, max_args=4, min_stmts=5, max_stmts=40, nested_min=1, nested_max=2, max_nest=2,
    p_assign=0.50, p_if=0.20, p_while=0.15, p_return=0.10,
    p_var_leaf=0.5, expr_p=0.25, p_leaf_operand=0.75, p_arith=0.55, p_compare=0.30, float_frac=0.5,
    balanced=0, balanced_min=0, balanced_max=0,
    chain_min=0, chain_max=0, p_chain=0.0, total_target=0,
)

STRESS = dict(BASE, min_args=6, max_args=16, nested_min=1, nested_max=2, max_nest=16,
              p_assign=0.45, p_if=0.22, p_while=0.18, p_return=0.10, float_frac=0.3,
              balanced=1, balanced_min=10, balanced_max=14)


# float_frac 0.08: float operands of int comparisons make whole float subtrees,
# so ~30 % of the nodes end up float-typed
STRESS_NESTED = dict(STRESS, chain_min=8, chain_max=16, p_chain=0.35, float_frac=0.08)


def _params(seed: int, n_functions: int, base=BASE, **kw) -> dict:
    p = dict(base)
    p.update(kw)
    p["seed"] = seed
    p["n_functions"] = n_functions
    return p


def program_source(seed: int, n_functions: int, targets=None, base=BASE, **kw) -> tuple[str, int]:
    return generate_source(_params(seed, n_functions, base, **kw), targets)


def single_50k(seed: int = SEED_SINGLE) -> Tree:
    """One ~50k-node program of 64 functions (regenerated with a derived seed
    until the node count lands in [45k, 55k])."""
    for attempt in range(64):
        src, n = program_source(seed + (attempt << 40), 64)
        if 45_000 <= n <= 55_000:
            return parse_source(src)
    raise RuntimeError("could not hit the 45k-55k node window")


def sized_source(seed: int, total_nodes: int, n_functions: int | None = None, base=BASE,
                 lo: float = 100.0, hi: float = 20_000.0) -> str:
    """Source of one program whose per-function node targets are log-uniform
    over [lo, hi] scaled to sum to total_nodes."""
    rng = np.random.default_rng(seed)
    if n_functions is None:
        n_functions = max(1, int(round(total_nodes / ((hi - lo) / math.log(hi / lo)))))
    t = np.exp(rng.uniform(math.log(lo), math.log(hi), n_functions))
    t *= total_nodes / t.sum()
    targets = np.maximum(8, np.round(t)).astype(np.uint32)
    src, _ = program_source(seed, n_functions, targets, base)
    return src


def sized_program(seed: int, total_nodes: int, n_functions: int | None = None, base=BASE,
                  lo: float = 100.0, hi: float = 20_000.0) -> Tree:
    return parse_source(sized_source(seed, total_nodes, n_functions, base, lo, hi))


def large_16m_source(seed: int = SEED_LARGE, total_nodes: int = 16_000_000, n_functions: int = 20_000) -> str:
    return sized_source(seed, total_nodes, n_functions)


def large_16m(seed: int = SEED_LARGE, total_nodes: int = 16_000_000, n_functions: int = 20_000) -> Tree:
    return parse_source(large_16m_source(seed, total_nodes, n_functions))


def request_sizes(n: int, seed: int, median: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(math.log(median), sigma, n)), lo, hi).astype(np.int64)


def source_of_size(seed: int, nodes: int, base=BASE) -> str:
    """Source of one program of about `nodes` nodes, functions of ~800 nodes each."""
    nf = max(1, nodes // 800)
    return sized_source(seed, nodes, nf, base, lo=100.0, hi=4000.0)


def program_of_size(seed: int, nodes: int, base=BASE) -> Tree:
    return parse_source(source_of_size(seed, nodes, base))


def batch_4096(n_programs: int = 4096, seed: int = SEED_BATCH) -> list[Tree]:
    sizes = request_sizes(n_programs, seed, 8000, 1.5, 1000, 1_000_000)
    return [program_of_size((seed << 32) + p, int(s)) for p, s in enumerate(sizes)]


def stress_program(seed: int, nodes: int = 64_000) -> Tree:
    """Functions of ~1.5k-node targets until the program holds 90 % of `nodes`
    nodes; the last function's overshoot brings the mean to about `nodes`.  A
    draw outside [0.75, 1.25] x `nodes` (one deep chain of wide expressions can
    add ~100k nodes) is regenerated with a derived seed."""
    cap = nodes // 64 + 16
    for attempt in range(64):
        src, n = program_source(seed + (attempt << 40), cap, np.full(cap, 1500, np.uint32), STRESS_NESTED,
                                total_target=int(nodes * 0.9))
        if 0.75 * nodes <= n <= 1.25 * nodes:
            return parse_source(src)
    raise RuntimeError("could not hit the stress program's node window")


def stress_1024(n_programs: int = 1024, seed: int = SEED_STRESS, nodes: int = 64_000) -> list[Tree]:
    return [stress_program((seed << 32) + p, nodes) for p in range(n_programs)]


def serve_pool(n: int = 384, seed: int = SEED_REQUESTS, threads: int = 1) -> list[Tree]:
    """Programs for the serving stream: request r draws its size from the
    log-normal (median 4k, sigma 2.0, clipped to [200, 4M]) and its program from
    seed SEED_REQUESTS * 2^32 + r.  A stream of more requests than the pool
    reuses them in order (request r -> program r mod n).  With threads > 1 the
    programs are generated on a pool of host threads (the generator and the front
    end run outside the interpreter lock); the result is the same."""
    sizes = request_sizes(n, seed, 4000, 2.0, 200, 4_000_000)
    one = lambda r: program_of_size((seed << 32) + r, int(sizes[r]))
    if threads <= 1:
        return [one(r) for r in range(n)]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(n)))


def serve_pool_records(n: int = 384, seed: int = SEED_REQUESTS, threads: int = 1):
    """The same programs as serve_pool(n, seed) as parser records (each
    flattens to serve_pool's tree of the same index)."""
    from .native import parse_arena
    sizes = request_sizes(n, seed, 4000, 2.0, 200, 4_000_000)
    one = lambda r: parse_arena(source_of_size((seed << 32) + r, int(sizes[r])))
    if threads <= 1:
        return [one(r) for r in range(n)]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(n)))


def arrivals(rate: float, seconds: float, seed: int = SEED_ARRIVALS) -> np.ndarray:
    """Poisson arrival times (s) at `rate` requests/s over `seconds`."""
    rng = np.random.default_rng(seed + int(rate))
    gaps = rng.exponential(1.0 / rate, int(rate * seconds * 1.2) + 16)
    t = np.cumsum(gaps)
    return t[t < seconds]
