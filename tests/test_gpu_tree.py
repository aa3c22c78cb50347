"""Device-side tree construction (ptl_tree_build, ptl_upload_arena,
ptl_compile_arena) against the DepthTree that the reference's front end writes
(tests/golden/trees, made by pareas-ast's DepthTree, depthtree.cpp:11-100), the
host front end's flattening of the same source, and a restatement of
depthtree.cpp's traversal for synthetic trees of extreme shapes; the compile
results against ptl_compile / ptl_compile_batch of the host-flattened trees.
Malformed records are refused with a status, never a fault."""
import glob
import os

import numpy as np
import pytest

from ptilopsis_amd import Arena, parse_arena, parse_source, workloads as W
from ptilopsis_amd.generator import PtlError

from test_tree_records import FIELDS, flatten

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "trees", "*.npz")))


def same_tree(a, b, what=""):
    for f in FIELDS:
        x, y = getattr(a, f) if not isinstance(a, dict) else a[f], getattr(b, f) if not isinstance(b, dict) else b[f]
        assert x.shape == y.shape, (what, f)
        if not np.array_equal(x, y):
            i = int(np.nonzero(x != y)[0][0])
            raise AssertionError(f"{what} {f}[{i}]: {x[i]} != {y[i]}")


def same_result(a, b, what=""):
    for k in ("words", "func_ids", "func_starts", "func_ends", "func_sizes"):
        assert np.array_equal(a[k], b[k]), (what, k)


def test_gpu_tree_golden_fixtures(gpu_ctx):
    """The reference front end's own DepthTrees, one by one and as one forest."""
    arenas, trees = [], []
    for path in FIXTURES:
        z = np.load(path)
        ar = parse_arena(z["source"].tobytes())
        t = gpu_ctx.tree_build(ar)[0]
        for f in FIELDS:
            assert np.array_equal(getattr(t, f), z[f]), (os.path.basename(path), f)
        assert t.max_depth == int(z["max_depth"][0])
        arenas.append(ar)
        trees.append(t)
    forest = gpu_ctx.tree_build(Arena.concat(arenas))
    assert len(forest) == len(trees)
    for p, (a, b) in enumerate(zip(forest, trees)):
        same_tree(a, b, f"forest program {p}")


def test_gpu_tree_single_50k_and_compile(gpu_ctx):
    src = None
    for attempt in range(64):
        s, n = W.program_source(W.SEED_SINGLE + (attempt << 40), 64)
        if 45_000 <= n <= 55_000:
            src = s
            break
    ar = parse_arena(src)
    host = parse_source(src)
    dev = gpu_ctx.tree_build(ar)[0]
    same_tree(dev, host, "single_50k")
    assert dev.max_depth == host.max_depth
    want = gpu_ctx.compile(host)
    got = gpu_ctx.compile_arena(ar)
    assert len(got) == 1
    same_result(got[0], want, "compile_arena")
    gpu_ctx.upload_arena(3, ar)
    same_result(gpu_ctx.compile_resident(3), want, "upload_arena")


def test_gpu_tree_large_16m(gpu_ctx):
    """About 34M tour events: two ruler levels above the tour."""
    src = W.large_16m_source()
    print("large_16m source", len(src), flush=True)
    ar = parse_arena(src)
    host = parse_source(src)
    del src
    dev = gpu_ctx.tree_build(ar)[0]
    same_tree(dev, host, "large_16m")
    assert dev.max_depth == host.max_depth


def test_gpu_tree_forest_compiles_like_batch(gpu_ctx):
    """A forest of uneven programs builds to the batch's concatenated trees and
    compiles to the same results as ptl_compile_batch of the host-flattened ones."""
    rng = np.random.default_rng(41)
    sizes = [int(x) for x in np.clip(rng.lognormal(np.log(4000), 1.2, 40), 300, 60000)]
    srcs = [W.source_of_size((W.SEED_REQUESTS << 32) + 900 + i, s) for i, s in enumerate(sizes)]
    arenas = [parse_arena(s) for s in srcs]
    hosts = [parse_source(s) for s in srcs]
    forest = Arena.concat(arenas)
    devs = gpu_ctx.tree_build(forest)
    for p, (a, b) in enumerate(zip(devs, hosts)):
        same_tree(a, b, f"program {p}")
        assert a.max_depth == b.max_depth
    want = gpu_ctx.compile_batch(hosts)
    got = gpu_ctx.compile_arena(forest)
    for p in range(len(hosts)):
        same_result(got[p], want[p], f"program {p}")
    gpu_ctx.upload_arena(4, forest)
    res = gpu_ctx.compile_resident_batch(4)
    for p in range(len(hosts)):
        same_result(res[p], want[p], f"resident program {p}")


def synthetic(shape: str, n: int, seed: int) -> Arena:
    """A tree of the given shape over records in shuffled order, typed so that
    it passes the pipeline's input checks (root STATEMENT_LIST, first node in
    post-order FUNC_DECL_DUMMY)."""
    rng = np.random.default_rng(seed)
    if shape == "chain":
        parent = np.arange(-1, n - 1)
    elif shape == "star":
        parent = np.zeros(n, np.int64); parent[0] = -1
    elif shape == "random":
        parent = np.concatenate([[-1], rng.integers(0, np.arange(1, n))])
    elif shape == "caterpillar":
        spine = n // 4
        parent = np.concatenate([[-1], np.arange(0, spine - 1), rng.integers(0, spine, n - spine)])
    else:
        raise ValueError(shape)
    perm = rng.permutation(n)  # record index of logical node i
    children = [[] for _ in range(n)]
    for i in range(1, n):
        children[int(parent[i])].append(i)
    for ch in children:
        rng.shuffle(ch)
    first = np.zeros(n, np.uint32); cnt = np.zeros(n, np.uint32)
    kids = []
    for i in range(n):
        first[perm[i]] = len(kids); cnt[perm[i]] = len(children[i])
        kids.extend(int(perm[c]) for c in children[i])
    ty = rng.integers(2, 42, n).astype(np.uint8)
    ar = Arena(ty, rng.integers(0, 6, n).astype(np.uint8), rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
               first, cnt, np.array(kids, np.uint32), np.array([perm[0]], np.uint32))
    # leftmost leaf -> FUNC_DECL_DUMMY, root -> STATEMENT_LIST
    v = int(perm[0])
    while ar.n_children[v]:
        v = int(ar.kids[int(ar.first_child[v])])
    ar.node_types[v] = 42
    ar.node_types[int(perm[0])] = 1
    return ar


@pytest.mark.parametrize("shape,n", [("chain", 200_000), ("star", 150_000), ("random", 300_000),
                                     ("caterpillar", 100_000), ("random", 3), ("chain", 2)])
def test_gpu_tree_synthetic_shapes(gpu_ctx, shape, n):
    ar = synthetic(shape, n, 1000 + n)
    want = flatten(ar, int(ar.roots[0]))
    got = gpu_ctx.tree_build(ar)[0]
    same_tree(got, want, shape)
    assert got.max_depth == int(want["depth"].max())


def test_gpu_tree_programs_in_root_order(gpu_ctx):
    """The forest's programs come out in the order of its roots, whatever the
    order of their records."""
    a = synthetic("random", 4000, 21)
    b = synthetic("caterpillar", 3000, 22)
    f = Arena.concat([a, b])
    f.roots = f.roots[::-1].copy()
    got = gpu_ctx.tree_build(f)
    assert len(got) == 2
    same_tree(got[0], flatten(b, int(b.roots[0])), "first root")
    same_tree(got[1], flatten(a, int(a.roots[0])), "second root")


def remove_from_parent(ar: Arena, c: int) -> Arena:
    """The records with c taken out of its parent's child list."""
    for u in range(ar.n):
        b, k = int(ar.first_child[u]), int(ar.n_children[u])
        lst = [int(x) for x in ar.kids[b:b + k]]
        if c in lst:
            lst.remove(c)
            ar.kids = np.concatenate([ar.kids, np.array(lst, np.uint32)])
            ar.first_child[u] = ar.kids.shape[0] - len(lst)
            ar.n_children[u] = len(lst)
            return ar
    raise ValueError("no parent")


def broken(kind: str) -> Arena:
    ar = synthetic("random", 5000, 7)
    ar = Arena(*[x.copy() if isinstance(x, np.ndarray) else x for x in
                 (ar.node_types, ar.result_types, ar.node_data, ar.first_child, ar.n_children, ar.kids, ar.roots)])
    root = int(ar.roots[0])
    inner = [v for v in range(ar.n) if ar.n_children[v] >= 2 and v != root]
    v = inner[0]
    b = int(ar.first_child[v])
    if kind == "kid_out_of_range":
        ar.kids[b] = ar.n + 5
    elif kind == "child_twice":
        ar.kids[b + 1] = ar.kids[b]
    elif kind == "list_past_end":
        ar.first_child[v] = ar.kids.shape[0] - 1
    elif kind == "unreached":
        # v's last child dropped from its list: that subtree hangs from nothing
        ar.n_children[v] -= 1
    elif kind == "cycle":
        # an inner node c leaves its parent's list, and c's first child g lists
        # c as its only child: c and g are each other's parent, detached from
        # the root
        c = [u for u in range(ar.n) if ar.n_children[u] >= 1 and u != root][0]
        ar = remove_from_parent(ar, c)
        g = int(ar.kids[int(ar.first_child[c])])
        ar.kids = np.concatenate([ar.kids, np.array([c], np.uint32)])
        ar.first_child[g] = ar.kids.shape[0] - 1
        ar.n_children[g] = 1
    elif kind == "root_is_child":
        ar.kids[b] = root
    elif kind == "bad_type":
        ar.node_types[v] = 44
    elif kind == "root_repeated":
        ar.roots = np.array([root, root], np.uint32)
    elif kind == "empty_program":
        ar = Arena(np.array([1], np.uint8), np.array([0], np.uint8), np.zeros(1, np.uint32), np.zeros(1, np.uint32),
                   np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(1, np.uint32))
    return ar


@pytest.mark.parametrize("kind,code", [("kid_out_of_range", -2), ("child_twice", -2), ("list_past_end", -2),
                                       ("unreached", -2), ("cycle", -2), ("root_is_child", -2), ("bad_type", -2),
                                       ("root_repeated", -2), ("empty_program", -3)])
def test_gpu_tree_refuses_malformed_records(gpu_ctx, kind, code):
    ar = broken(kind)
    with pytest.raises(PtlError) as e:
        gpu_ctx.tree_build(ar)
    assert e.value.status == code, kind
    with pytest.raises(PtlError):
        gpu_ctx.compile_arena(ar)
    # the context still works
    good = synthetic("random", 1000, 3)
    same_tree(gpu_ctx.tree_build(good)[0], flatten(good, int(good.roots[0])), "after " + kind)
