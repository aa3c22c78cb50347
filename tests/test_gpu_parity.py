"""GPU parity: libptl_hip.so against the oracle, bit-exact, through the C ABI.

Small and medium programs are compared word for word (and stage by stage for
the corpus); large ones through size-independent properties.  Each case prints
what it runs before touching the GPU."""
import glob
import os
import sys

import numpy as np
import pytest

import oracle
from ptilopsis_amd import parse_source, workloads as W
from ptilopsis_amd.generator import Context, PtlError

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CORPUS = sorted(glob.glob(os.path.join(HERE, "golden", "corpus", "*")))
OUT = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def assert_same(gpu, ref):
    for k in OUT:
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        assert g.shape == r.shape, (k, g.shape, r.shape)
        bad = np.nonzero(g != r)[0]
        assert bad.size == 0, f"{k}: {bad.size} differ, first at {bad[0]}: {g[bad[0]]:#x} != {r[bad[0]]:#x}"


def test_gpu_scan_kat(gpu_ctx):
    import mt19937_64 as MT
    x = MT.uniform_0_4095(0xC0FFEEA4DBEEF, 1 << 20)
    say("scan KAT 2^20")
    ex = gpu_ctx.exclusive_scan(x)
    ref = np.cumsum(x.astype(np.uint64))
    assert ex[0] == 0 and np.array_equal(ex[1:].astype(np.uint64), ref[:-1])


# tile edges of the single-pass scan (4096 elements a tile) and of its
# look-back window (64 tiles a poll)
@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 4095, 4096, 4097, 64 * 4096 + 1, 129 * 4096 - 1, 1 << 22])
def test_gpu_scan_sizes(gpu_ctx, n):
    x = np.random.default_rng(n).integers(0, 1000, n).astype(np.uint32)
    ex = gpu_ctx.exclusive_scan(x)
    assert np.array_equal(ex, (np.cumsum(x.astype(np.uint64)) - x).astype(np.uint32))


STAGE_ARRAYS = {1: ("types", "rtypes", "node_data"), 2: ("locs", "func_ids", "func_starts", "func_ends", "func_sizes"),
                3: ("instr", "rd", "rs1", "rs2", "jt"), 4: ("used",), 5: ("sym_sw", "instr", "rd", "rs1", "rs2", "jt")}


def norm(name, a):
    a = np.asarray(a, np.uint64)
    if name in ("used", "sym_sw"):
        return (a != 0).astype(np.uint64)
    if name in ("types", "rtypes"):
        return a & 0xFF
    return a


@pytest.mark.parametrize("path", CORPUS, ids=[os.path.basename(p) for p in CORPUS])
def test_gpu_corpus_stage_by_stage(gpu_ctx, path):
    t = parse_source(open(path, "rb").read())
    say("corpus", os.path.basename(path), t.n)
    o = oracle.Oracle(t)
    try:
        for st in range(1, 6):
            o.run(st)
            gpu_ctx.stop_after(st)
            gpu_ctx.compile(t)
            for nm in STAGE_ARRAYS[st]:
                assert np.array_equal(norm(nm, gpu_ctx.dump(nm)), norm(nm, o.get(nm))), (st, nm)
    finally:
        gpu_ctx.stop_after(0)
    o.run(7)
    assert_same(gpu_ctx.compile(t), o.result())


GEN_CASES = [("base", s, 4 + s) for s in range(6)] + [("stress", 100 + s, 2) for s in range(3)] + \
            [("manyfloat", 200 + s, 8) for s in range(2)]


@pytest.mark.parametrize("kind,seed,nf", GEN_CASES, ids=[f"{k}-{s}" for k, s, _ in GEN_CASES])
def test_gpu_generated(gpu_ctx, kind, seed, nf):
    base = {"base": W.BASE, "stress": W.STRESS, "manyfloat": dict(W.BASE, float_frac=0.9)}[kind]
    src, n = W.program_source(seed, nf, base=base)
    t = parse_source(src)
    say("generated", kind, seed, t.n)
    assert_same(gpu_ctx.compile(t), oracle.compile_tree(t))


def test_gpu_single_50k(gpu_ctx):
    t = W.single_50k()
    say("single_50k", t.n)
    r = gpu_ctx.compile(t)
    assert r["tie_slots"] > 0  # exercises the introsort ordering of equal-depth writers
    assert_same(r, oracle.compile_tree(t))


@pytest.mark.parametrize("which", ["single_50k", "1M"])
def test_gpu_sort_permutation(gpu_ctx, which):
    # the whole depth-sort permutation (libstdc++ introsort order, including the
    # closed-form placement of all-equal ranges) against the oracle's restatement
    t = W.single_50k() if which == "single_50k" else W.program_of_size(0x5EED0007, 1_000_000)
    say("sort permutation", which, t.n)
    r = gpu_ctx.compile(t)
    assert r["tie_slots"] > 0  # the permutation is only computed when equal-depth writers collide
    o = oracle.Oracle(t)
    o.run(3)
    want = o.get("sorted")
    got = gpu_ctx.dump("sorted")
    assert got.shape == want.shape
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} positions differ, first at {bad[0]}"


def test_gpu_register_named_by_two_functions(gpu_ctx):
    # calls with many arguments: an fmv.x.w at the start of one function names
    # a register that the end of the previous function defines; the program is
    # allocated in the reference's exact step/group order (k_ra_serial)
    t = W.program_of_size((W.SEED_STRESS << 32) + 3, 64000, W.STRESS)
    say("stress program with shared register names", t.n)
    gpu_ctx.profile(True)
    r = gpu_ctx.compile(t)
    names = gpu_ctx.kernel_times()
    gpu_ctx.profile(False)
    assert "k_ra_serial" in names
    assert_same(r, oracle.compile_tree(t))


def test_gpu_one_million_nodes(gpu_ctx):
    t = W.program_of_size(0x5EED0007, 1_000_000)
    say("1M-node program", t.n)
    assert_same(gpu_ctx.compile(t), oracle.compile_tree(t))


@pytest.mark.parametrize("which", ["single_50k", "300k"])
def test_gpu_walk_and_lockstep_overlap(gpu_ctx, which):
    # one program: the walk publishes its lookup rows while the lockstep runs;
    # run one after the other, the output is the same, and both equal the oracle
    t = W.single_50k() if which == "single_50k" else W.program_of_size(0x5EED0008, 300_000)
    want = oracle.compile_tree(t)
    got = {}
    try:
        for on in (0, 1, 2):
            say("walk/lockstep overlap", on, which, t.n)
            gpu_ctx.set_overlap(on)
            gpu_ctx.profile(True)
            got[on] = gpu_ctx.compile(t)
            names = gpu_ctx.kernel_times()
            gpu_ctx.profile(False)
            assert "k_ra_walk" in names and "k_ra_lane0" in names
            # 2: the lockstep's wait gives up and the allocation is redone one
            # kernel after the other (the table copies then merged)
            assert ("k_walk_merge" in names) == (on != 1)
            if on == 2:
                assert names["k_ra_lane0"][1] == 2 and names["k_ra_walk"][1] == 2
    finally:
        gpu_ctx.set_overlap(1)
        gpu_ctx.profile(False)
    for on in (0, 1, 2):
        assert_same(got[on], want)


@pytest.mark.parametrize("which", ["single_50k", "stress"])
def test_gpu_tuning_knobs_do_not_change_output(gpu_ctx, which):
    # every launch parameter the autotuner may pick gives the same words
    t = W.single_50k() if which == "single_50k" else W.program_of_size((W.SEED_STRESS << 32) + 5, 64000, W.STRESS)
    want = oracle.compile_tree(t)
    b = 0 if t.n < (64 << 10) else 1
    knobs = ("block.preprocess", "block.isn_cnt", "block.isn_gen", "block.optimize", "block.emit", "block.fix_jumps")
    old = {k: gpu_ctx.tune_get(k, b) for k in knobs + ("walk_ring",)}
    try:
        for bs, ring in ((64, 8), (128, 4), (64, 4)):
            say("tuning", which, bs, ring)
            for k in knobs:
                gpu_ctx.tune_set(k, b, bs)
            gpu_ctx.tune_set("walk_ring", b, ring)
            assert_same(gpu_ctx.compile(t), want)
    finally:
        for k, v in old.items():
            gpu_ctx.tune_set(k, b, v)
    with pytest.raises(PtlError):
        gpu_ctx.tune_set("block.emit", 0, 96)
    with pytest.raises(PtlError):
        gpu_ctx.tune_set("walk_ring", 4, 4)


@pytest.mark.parametrize("nf", [40000, 66000])
def test_gpu_lockstep_largest_classes(gpu_ctx, nf):
    # 40000 functions: 5000 groups, the 8192-group class (k_ra_lane0<8, 8192>);
    # 66000: 8250 groups, past it (k_ra_lane0_generic)
    src, _ = W.program_source(77, nf, min_stmts=1, max_stmts=1)
    t = parse_source(src)
    say("lockstep class of", nf, "functions", t.n)
    gpu_ctx.profile(True)
    try:
        r = gpu_ctx.compile(t)
        names = gpu_ctx.kernel_times()
    finally:
        gpu_ctx.profile(False)
    assert ("k_ra_lane0_generic" in names) == (nf > 65536)
    assert_same(r, oracle.compile_tree(t))


def test_gpu_many_small_functions(gpu_ctx):
    src, _ = W.program_source(31, 300, min_stmts=1, max_stmts=3)
    t = parse_source(src)
    say("300 tiny functions", t.n)
    assert_same(gpu_ctx.compile(t), oracle.compile_tree(t))


def test_gpu_single_function_and_single_statement(gpu_ctx):
    for src in ("function f() : void { }", "function f(a : int) : int { return a; }",
                "function f(a : float, b : int) : float { while a < 2.0 { a = a + 1.0; } return a; }"):
        t = parse_source(src)
        assert_same(gpu_ctx.compile(t), oracle.compile_tree(t))


def test_gpu_errors(gpu_ctx):
    empty = parse_source("")
    with pytest.raises(PtlError) as e:
        gpu_ctx.compile(empty)
    assert e.value.status == -3
    t = W.single_50k()
    bad = W.Tree(t.node_types, t.result_types, t.parents.copy(), t.depth, t.child_idx, t.node_data, t.max_depth)
    bad.parents[10] = 5  # a child before its parent breaks post-order
    with pytest.raises(PtlError) as e:
        gpu_ctx.compile(bad)
    assert e.value.status == -2


def test_gpu_resident_deterministic(gpu_ctx):
    t = W.program_of_size(0x5EED0008, 300_000)
    a = gpu_ctx.compile(t)
    gpu_ctx.upload(3, t)
    b = gpu_ctx.compile_resident(3)
    c = gpu_ctx.compile_resident(3)
    assert_same(a, b)
    assert_same(b, c)


def test_gpu_large_properties(gpu_ctx):
    # size-independent properties at 4M nodes: the function table tiles the
    # output, functions open with the prologue stores and close with ret
    t = W.sized_program(0x5EED0009, 4_000_000, 5000)
    say("4M-node program", t.n)
    r = gpu_ctx.compile(t)
    w, st, en, sz = r["words"], r["func_starts"], r["func_ends"], r["func_sizes"]
    assert st[0] == 0 and en[-1] == len(w) and np.array_equal(st[1:], en[:-1]) and np.array_equal(sz, en - st)
    assert np.all(w[st] == 0xFE112E23) and np.all(w[st + 1] == 0xFE812C23) and np.all(w[en - 1] == 0x8067)
    assert np.array_equal(np.sort(r["func_ids"]), np.arange(len(st)))
    r2 = gpu_ctx.compile(t)
    assert_same(r, r2)


def small_program_cases():
    # (name, tree, whether k_ra_small takes it: 9 to 64 functions, states within 40 KB)
    return [
        ("one_group", W.program_of_size(0x5EED0B01, 3000), False),                   # k_ra_group
        ("two_groups", W.program_of_size(0x5EED0B04, 9000), True),
        ("eight_groups", parse_source(W.program_source(0x5EED0B02, 64, min_stmts=2, max_stmts=6)[0]), True),
        ("65_functions", parse_source(W.program_source(0x5EED0B03, 65, min_stmts=2, max_stmts=6)[0]), False),
        ("cross_named", W.program_of_size((W.SEED_STRESS << 32) + 35, 9000, W.STRESS), True),  # serial path
        ("large_state", W.stress_program((W.SEED_STRESS << 32) + 19, 64000), False),  # 96 KB of states
        ("stress_small", W.program_of_size((W.SEED_STRESS << 32) + 7, 6000, W.STRESS), False),  # 7 functions
    ]


@pytest.mark.parametrize("small", [1, 0])
def test_gpu_small_programs_both_paths(gpu_ctx, small):
    # with the small_programs knob at 1 (default 0), programs of at most 64
    # functions go to k_ra_small (one wave each); both paths give the oracle's words
    cases = small_program_cases()
    old = [gpu_ctx.tune_get("small_programs", b) for b in range(4)]
    try:
        for b in range(4):
            gpu_ctx.tune_set("small_programs", b, small)
        for name, t, takes in cases:
            say("small programs", small, name, t.n, "nodes")
            gpu_ctx.profile(True)
            got = gpu_ctx.compile(t)
            names = gpu_ctx.kernel_times()
            gpu_ctx.profile(False)
            assert ("k_ra_small" in names) == (small == 1 and takes), (name, sorted(names))
            if name == "cross_named":
                assert "k_ra_serial" in names
            assert_same(got, oracle.compile_tree(t))
        say("small programs", small, "batch of all")
        res = gpu_ctx.compile_batch([t for _, t, _ in cases])
        for (name, t, _), r in zip(cases, res):
            assert_same(r, oracle.compile_tree(t))
    finally:
        for b in range(4):
            gpu_ctx.tune_set("small_programs", b, old[b])
