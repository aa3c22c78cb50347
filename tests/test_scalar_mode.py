"""The oracle's scalar-semantics mode (rv_generator_st, src/codegen/rv_generator.cpp:227-1736;
oracle/ptl_scalar.inc) against the AVX2 restatement, to show which
differences between the reference's two generators there are.

Each catalogued divergence can be switched to the
AVX2 behaviour in the scalar mode.  With every switch on, each stage fed the
same input (the AVX2 restatement's previous stages) gives exactly the AVX2
restatement's output, on generated programs of three shapes and on targeted
ones; for regalloc on programs of one function, because D5 (lookups read
region 0 of a step snapshot shared by all functions) is the one divergence
that is a different model rather than a switch.  Each switch is also shown to
matter on some program, except D10 (equivalent) and D6 (needs two swapped
operands, which without D5's cross-function marks no program here produces).
Switching them also found a divergence the catalogue lacks, D17: the two
stack-count stores of preprocess's argument-list pass (rv_generator_avx.cpp:236-259)
are masked by the previous node's type alone, so in an 8-node vector holding an
argument list every node after a stack call argument, or after a float call
argument in a float register, takes the count as its node data -- a literal or
a variable reference among them changes the code.

The scalar restatement itself is pinned by the stage-by-stage equality (every
line outside the switches is shared behaviour) and by answers worked out by
hand from rv_generator.cpp at the switch points (test_scalar_known_answers)."""
import numpy as np
import pytest

import oracle
import rvdecode
from ptilopsis_amd import parse_source, workloads as W

STAGE_ARRAYS = {
    1: ["types", "rtypes", "node_data"],
    2: ["sizes", "locs", "func_ids", "func_starts", "func_ends", "func_sizes"],
    3: ["instr", "rd", "rs1", "rs2", "jt"],
    4: ["used"],
    5: ["instr", "rd", "rs1", "rs2", "jt", "func_starts", "func_ends", "func_sizes"],
    6: ["instr", "rd", "rs1", "rs2", "jt", "func_starts", "func_ends", "func_sizes"],
    7: ["instr"],
}

NINE_FLOATS = ("function g(" + ", ".join(f"p{k} : float" for k in range(9)) + ") : int { return 0; } "
               "function f(x : float) : int { return g(x, x, x, x, x, x, x, x, x); }")
CORPUS = {
    "single_64fn": lambda: W.program_source(W.SEED_SINGLE, 64)[0],
    "small_4fn": lambda: W.program_source(5, 4)[0],
    "stress_7fn": lambda: W.source_of_size((W.SEED_STRESS << 32) + 3, 6000, W.STRESS),
    "one_fn": lambda: W.program_source(9, 1)[0],
    "one_fn_stress": lambda: W.program_source(11, 1, base=W.STRESS)[0],
    "nine_float_args": lambda: NINE_FLOATS,
    "casts": lambda: "function f(a : int, x : float, y : float) : float { return (a + x @int + (y * x) @int) @float; }",
    "unbraced_return": lambda: "function f(x : int) : int { if x return 1; return 0; }",
    "float_args_then_int": lambda: ("function g(a : float, b : int, c : float, d : int) : int { return b + d; } "
                                    "function f(x : float, y : int) : int { return g(x, y, x, 7) + g(1.5, y, x, y); }"),
}
_trees = {}


def tree(name):
    if name not in _trees:
        _trees[name] = parse_source(CORPUS[name]())
    return _trees[name]


def stage_diff(t, k, avx):
    """Arrays that differ after stage k when both semantics get the AVX2
    restatement's stages 1..k-1 as input."""
    a, s = oracle.Oracle(t), oracle.Oracle(t)
    for j in range(1, k):
        a.run_stage(j)
        s.run_stage(j)
    a.run_stage(k)
    s.run_stage(k, scalar=True, avx=avx)
    out = []
    for nm in STAGE_ARRAYS[k]:
        x, y = a.get(nm), s.get(nm)
        if nm == "used":  # the AVX2 generator keeps a u32 mask, the scalar one a bool
            x, y = x != 0, y != 0
        if x.shape != y.shape or not np.array_equal(x, y):
            out.append(nm)
    a.close()
    s.close()
    return out


@pytest.mark.parametrize("name", list(CORPUS))
def test_scalar_with_catalogued_switches_equals_avx_stage_by_stage(name):
    t = tree(name)
    for k in range(1, 8):
        if k == 5 and t.n_functions > 1:
            continue  # D5, below
        assert stage_diff(t, k, oracle.SC_ALL) == [], (name, k)
    if t.n_functions == 1:
        a = oracle.compile_tree(t)
        s = oracle.compile_tree(t, scalar=True, avx=oracle.SC_ALL)
        for k in ("words", "func_ids", "func_starts", "func_ends", "func_sizes"):
            assert np.array_equal(a[k], s[k]), (name, k)


@pytest.mark.parametrize("name", ["small_4fn", "stress_7fn"])
def test_d5_is_what_remains_with_several_functions(name):
    assert stage_diff(tree(name), 5, oracle.SC_ALL) != []


def scalar_output(t, avx, stage=7, name="instr"):
    o = oracle.Oracle(t)
    o.run_scalar(stage, avx)
    r = o.get(name)
    o.close()
    return r


@pytest.mark.parametrize("d", ["D1", "D2", "D3", "D4", "D9", "D11", "D12", "D13", "D14", "D15", "D17"])
def test_each_catalogued_divergence_is_observable(d):
    """With every switch but one on, some program's output changes; D14 only
    changes the (internal) jump targets after fix_jumps."""
    bit = oracle.SC[d]
    stage, name = (6, "jt") if d == "D14" else (7, "instr")
    hit = [nm for nm in CORPUS
           if not np.array_equal(scalar_output(tree(nm), oracle.SC_ALL, stage, name),
                                 scalar_output(tree(nm), oracle.SC_ALL & ~bit, stage, name))]
    assert hit, d


@pytest.mark.parametrize("d", ["D6", "D10"])
def test_divergences_this_corpus_cannot_show(d):
    bit = oracle.SC[d]
    for nm in CORPUS:
        assert np.array_equal(scalar_output(tree(nm), oracle.SC_ALL), scalar_output(tree(nm), oracle.SC_ALL & ~bit)), nm


def test_d17_overwrites_the_node_after_a_float_call_argument():
    """In the AVX2 output (the canonical one) the node after a float call
    argument in a float register takes the stack count, here 0, when its 8-node
    vector holds an argument list; found by switching D10 alone."""
    found = 0
    for nm in ("single_64fn", "float_args_then_int"):
        t = tree(nm)
        a = oracle.Oracle(t)
        a.run_stage(1)
        nd_avx = a.get("node_data")
        a.close()
        s = oracle.Oracle(t)
        s.run_stage(1, scalar=True, avx=oracle.SC_ALL & ~oracle.SC["D17"])
        nd_sc = s.get("node_data")
        s.close()
        ty = t.node_types
        for i in np.nonzero(nd_avx != nd_sc)[0]:
            # pareas FUNC_CALL_ARG (11) before it, not an argument list (12) itself
            assert ty[i - 1] == 11 and ty[i] != 12, (nm, i)
            assert nd_avx[i] == 0
            found += 1
    assert found


def test_scalar_known_answers():
    # D1 (rv_generator.cpp:695, :69-73): lui takes (nd - sext(nd & 0xFFF)) & 0xFFFFF000,
    # so 75845 = 0x12845 (bit 11 set) loads 0x13000 and adds 0x845 - 0x1000 = -1979
    w = oracle.compile_tree(parse_source("function f() : int { return 75845; }"), scalar=True)["words"]
    lui = [k for k, x in enumerate(w) if (int(x) & 0x7F) == 0x37 and ((int(x) >> 7) & 31) == 5]
    assert len(lui) == 1 and int(w[lui[0]]) >> 12 == 0x13
    assert rvdecode.fields(int(w[lui[0] + 1]))["imm_i"] == -1979
    # D12 (rv_generator.cpp:746): a while's back edge goes to registers[3n+0], the
    # loop head, which lies before the jump (the AVX2 one jumps to itself)
    w = oracle.compile_tree(parse_source("function f(a : int) : int { while a < 10 { a = a + 1; } return a; }"),
                            scalar=True)["words"]
    pairs = [k for k in range(len(w) - 1)
             if (int(w[k]) & 0x7F) == 0x17 and ((int(w[k]) >> 7) & 31) == 1 and (int(w[k + 1]) & 0x7F) == 0x67
             and ((int(w[k + 1]) >> 7) & 31) == 0 and int(w[k + 1]) != 0x00008067]
    back = [k for k in pairs if rvdecode.sext((int(w[k]) & 0xFFFFF000) + rvdecode.fields(int(w[k + 1]))["imm_i"], 32) < 0]
    assert len(back) == 1
    # D9 (rv_generator.cpp:247-281): an int argument after one float is in a0
    w = oracle.compile_tree(parse_source("function f(x : float, a : int) : int { return a; }"), scalar=True)["words"]
    stores = [(rvdecode.fields(int(x))["rs2"], int(x) & 0x7F) for x in w
              if (int(x) & 0x7F) in (0x23, 0x27) and rvdecode.fields(int(x))["rs1"] == 5]
    assert stores == [(10, 0x27), (10, 0x23)]  # fsw fa0; sw a0
    # D13 (rv_generator.cpp:1488-1602): a function using a callee-saved register
    # saves and restores it, and its frame grows by 4 per saved register
    t = tree("one_fn_stress")
    o = oracle.Oracle(t)
    o.run_scalar(7)
    w = o.get("instr")
    o.close()
    frame = (int(w[3]) >> 20) & 0xFFF  # addi s0, s0, frame (start + 3)
    assert frame > 8 and frame % 4 == 0
    saves = [x for x in w[6:6 + (frame - 8) // 4]]
    assert all((int(x) & 0x707F) in (0x2023, 0x2027) and ((int(x) >> 15) & 31) == 8 for x in saves)
