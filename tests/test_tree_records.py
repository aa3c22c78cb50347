"""Parser node records (ptl_parse_arena), the input of the device-side tree
construction, against the DepthTree fixtures written by the reference's own
pareas-ast front end: flattened the way DepthTree::construct does
(subprojects/pareas-ast/src/codegen/depthtree.cpp:52-100: post-order with the
children in order, depth from the root, child_idx the position among the
parent's children, root parent and child_idx -1), the records give exactly the
fixture's arrays."""
import glob
import os

import numpy as np
import pytest

from ptilopsis_amd import Arena, parse_arena, parse_source, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "trees", "*.npz")))
FIELDS = ("node_types", "result_types", "parents", "depth", "child_idx", "node_data")


def flatten(ar: Arena, root: int):
    """depthtree.cpp:52-100 on the records, with an explicit stack."""
    order = []          # (node, depth, child position) in post-order
    stack = [(root, 0, -1, 0)]  # node, depth, child position, next child
    while stack:
        v, d, ci, nxt = stack[-1]
        if nxt < int(ar.n_children[v]):
            stack[-1] = (v, d, ci, nxt + 1)
            c = int(ar.kids[int(ar.first_child[v]) + nxt])
            stack.append((c, d + 1, nxt, 0))
        else:
            order.append((v, d, ci))
            stack.pop()
    index = {v: i for i, (v, _, _) in enumerate(order)}
    n = len(order)
    out = {f: np.zeros(n, dt) for f, dt in zip(FIELDS, (np.uint8, np.uint8, np.int32, np.int32, np.int32, np.uint32))}
    out["parents"][:] = -1
    for i, (v, d, ci) in enumerate(order):
        out["node_types"][i] = ar.node_types[v]
        out["result_types"][i] = ar.result_types[v]
        out["depth"][i] = d
        out["child_idx"][i] = ci
        out["node_data"][i] = ar.node_data[v]
        b = int(ar.first_child[v])
        for k in range(int(ar.n_children[v])):
            out["parents"][index[int(ar.kids[b + k])]] = i
    return out


def check_records(ar: Arena):
    """Every record in the root's tree, listed once as a child."""
    n = ar.n
    seen = np.zeros(n, np.int64)
    for v in range(n):
        b, k = int(ar.first_child[v]), int(ar.n_children[v])
        for c in ar.kids[b:b + k]:
            seen[int(c)] += 1
    r = int(ar.roots[0])
    assert seen[r] == 0
    assert np.all(np.delete(seen, r) == 1)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_records_flatten_to_reference_fixture(path):
    z = np.load(path)
    ar = parse_arena(z["source"].tobytes())
    check_records(ar)
    out = flatten(ar, int(ar.roots[0]))
    for f in FIELDS:
        assert np.array_equal(out[f], z[f]), f


@pytest.mark.parametrize("seed", [11, 12])
def test_records_match_flattened_parse(seed):
    src, n = W.program_source(seed, 12)
    ar = parse_arena(src)
    t = parse_source(src)
    assert ar.n == t.n == n
    check_records(ar)
    out = flatten(ar, int(ar.roots[0]))
    for f in FIELDS:
        assert np.array_equal(out[f], getattr(t, f)), f
    assert ar.n_functions == t.n_functions


def test_records_parse_errors():
    with pytest.raises(ValueError):
        parse_arena("func f() -> int { return x; }")


def test_concat_shifts_indices():
    a = parse_arena(W.program_source(3, 2)[0])
    b = parse_arena(W.program_source(4, 3)[0])
    f = Arena.concat([a, b])
    assert f.n == a.n + b.n
    assert list(f.roots) == [int(a.roots[0]), a.n + int(b.roots[0])]
    ob = flatten(b, int(b.roots[0]))
    of = flatten(f, int(f.roots[1]))
    for k in FIELDS:
        assert np.array_equal(ob[k], of[k]), k


@pytest.mark.parametrize("path", FIXTURES[:6], ids=[os.path.basename(p)[:-4] for p in FIXTURES[:6]])
def test_host_flatten_of_records(path):
    """ptl_flatten_records (the CPU baseline of the device build) gives the
    reference fixture's arrays too."""
    from ptilopsis_amd.native import flatten_records
    z = np.load(path)
    t = flatten_records(parse_arena(z["source"].tobytes()))
    for f in FIELDS:
        assert np.array_equal(getattr(t, f), z[f]), f
    assert t.max_depth == int(z["max_depth"][0])
