"""Known answers for the oracle's behaviours where the reference's AVX2
generator (the canonical semantics) departs from the obvious reading or from
its own scalar generator.  Each expected value is derived by hand from the
cited reference lines and the RISC-V encodings (the ISA manual), not from the
oracle's code, and each is chosen so that the plausible misreading (usually
the scalar generator's behaviour) fails it.

Reference: src/codegen/rv_generator_avx.cpp (AVX2 at one thread),
include/codegen/rv_nodetype.hpp (tables), subprojects/pareas-ast (front end).
"""
import numpy as np
import pytest

import oracle
import rvdecode
from ptilopsis_amd import parse_source, workloads as W

# pareas NodeType (pareas-ast include/codegen/astnode.hpp:11-56) and DataType
# (datatype.hpp:7-14) values
LESS_EXPR, DEREF_EXPR, FLOAT = 28, 37, 3


def compile_src(src):
    return oracle.compile_tree(parse_source(src))


def f(w):
    return rvdecode.fields(int(w))


def find(words, pred):
    return [k for k, w in enumerate(words) if pred(int(w))]


def b_target(words, k):
    """Target index of the branch at k (B-type immediate, ISA manual 2.5)."""
    w = int(words[k])
    imm = ((w >> 31) << 12) | (((w >> 7) & 1) << 11) | (((w >> 25) & 0x3F) << 5) | (((w >> 8) & 0xF) << 1)
    return k + rvdecode.sext(imm, 13) // 4


def jump_target(words, k):
    """Target index of the auipc+jalr pair starting at k."""
    a, j = int(words[k]), f(words[k + 1])
    off = (a & 0xFFFFF000) + j["imm_i"]
    off = off - (1 << 32) if off >= (1 << 31) else off
    return k + off // 4


def is_auipc_ra(w): return (w & 0x7F) == 0x17 and ((w >> 7) & 31) == 1
def is_jalr_x0(w): return (w & 0x7F) == 0x67 and ((w >> 7) & 31) == 0
def is_beq(w): return (w & 0x7F) == 0x63 and ((w >> 12) & 7) == 0


# ---------------------------------------------------------------- D1
@pytest.mark.parametrize("value,upper,lower", [
    (75845, 0x12, 0x845),     # 0x12845: bit 11 set
    (2049, 0x0, 0x801),       # 0x801
    (4095, 0x0, 0xFFF),
    (1234, 0x0, 0x4D2),       # bit 11 clear: both readings agree
])
def test_d1_literal_lui_takes_the_upper_bits_unadjusted(value, upper, lower):
    # rv_generator_avx.cpp:766-768: lui |= nd & 0xFFFFF000, addi |= (nd & 0xFFF) << 20.
    # The scalar generator (rv_generator.cpp:695, :69-73) adds 1 to the upper
    # part when bit 11 is set, so that lui + addi rebuild the value; the AVX2
    # one does not.
    r = compile_src(f"function f() : int {{ return {value}; }}")
    w = r["words"]
    lui = find(w, lambda x: (x & 0x7F) == 0x37 and ((x >> 7) & 31) == 5)  # lui t0 (s0's frame luis excluded)
    assert len(lui) == 1
    k = lui[0]
    assert int(w[k]) >> 12 == upper
    addi = f(w[k + 1])
    assert addi["op"] == 0x13 and addi["f3"] == 0 and addi["rd"] == 5 and addi["rs1"] == 5
    assert int(w[k + 1]) >> 20 == lower


def test_d1_float_literal_bit_pattern():
    # a float literal travels as its IEEE bit pattern (lexer.cpp:73-77): 0.1f is
    # 0x3DCCCCCD, bit 11 set; then fmv.w.x moves it to a float register
    r = compile_src("function f() : float { return 0.1; }")
    w = r["words"]
    k = find(w, lambda x: (x & 0x7F) == 0x37 and ((x >> 7) & 31) == 5)[0]
    assert int(w[k]) >> 12 == 0x3DCCC and int(w[k + 1]) >> 20 == 0xCCD
    assert rvdecode.mnemonic(int(w[k + 2])) == "fmv.w.x"


# ---------------------------------------------------------------- control flow
def test_d12_while_back_edge_jumps_to_itself():
    # while = [WHILE_DUMMY, cond, body] (parser.cpp:331-336), condition child 1
    # (rv_nodetype.hpp:56-60).  Its words (rv_nodetype.hpp:764-771): a branch at
    # its own location and a jalr placed at registers[3n+2] (rv_generator_avx.cpp:720-732),
    # which the body (child 2, a non-condition child) sets to its end, loc + size
    # (:703-708).  Jump targets (rv_nodetype.hpp:1064, :785-797): the branch takes
    # calc 4 = registers[3n+2] + 1 (past the jalr, the loop exit) and the jalr
    # calc 5 = registers[3n+2], its own slot.  So the back edge jumps to itself
    # (after fix_jumps' auipc+jalr expansion: auipc ra, 0 / jalr x0, 0(ra)); the
    # scalar generator jumps to the loop head (registers[3n+0]).
    r = compile_src("function f(a : int) : int { while a < 10 { a = a + 1; } return a; }")
    w = r["words"]
    br = find(w, is_beq)
    assert len(br) == 1
    exit_ = b_target(w, br[0])
    k = exit_ - 2  # the expanded jalr sits right before the exit
    assert is_auipc_ra(int(w[k])) and is_jalr_x0(int(w[k + 1]))
    assert int(w[k]) >> 12 == 0 and f(w[k + 1])["imm_i"] == 0 and f(w[k + 1])["rs1"] == 1
    assert jump_target(w, k) == k
    assert f(w[br[0]])["rs2"] == 0  # beq cond, x0: leave when the condition is 0


def test_if_else_branch_and_jump_targets():
    # if_else = [cond, then, else] (parser.cpp:318-325), condition child 0.  Its
    # branch goes to registers[3n+1] + 1 (calc 2) and its jalr, placed at
    # registers[3n+1] (the end of the then-body), to registers[3n+2] (calc 3, the
    # end of the else-body) (rv_nodetype.hpp:1065; rv_generator_avx.cpp:720-732, 785-797).
    src = "function f(a : int) : int { if a < 10 { a = 3; } else { a = 4; } return a; }"
    w = compile_src(src)["words"]
    br = find(w, is_beq)
    assert len(br) == 1
    els = b_target(w, br[0])
    k = els - 2  # the jump pair ends the then-body
    assert is_auipc_ra(int(w[k])) and is_jalr_x0(int(w[k + 1]))
    # then-body: addi (address of a), lui+addi (3), sw; else-body the same with 4
    then_body = w[br[0] + 1:k]
    else_body = w[els:jump_target(w, k)]
    assert len(then_body) == 4 and len(else_body) == 4
    assert int(then_body[2]) >> 20 == 3 and int(else_body[2]) >> 20 == 4
    assert rvdecode.mnemonic(int(then_body[3])) == "sw" and rvdecode.mnemonic(int(else_body[3])) == "sw"


def test_if_branch_target():
    # if = [cond, then]: one branch to registers[3n+1] (calc 1), the end of the
    # then-body (rv_nodetype.hpp:1063)
    w = compile_src("function f(a : int) : int { if a < 10 { a = 3; } return a; }")["words"]
    br = find(w, is_beq)
    assert len(br) == 1
    t = b_target(w, br[0])
    assert rvdecode.mnemonic(int(w[t - 1])) == "sw" and int(w[t - 2]) >> 20 == 3  # after "a = 3"
    assert rvdecode.mnemonic(int(w[t])) == "addi" and f(w[t + 1])["op"] == 0x03    # "return a" loads a


# ---------------------------------------------------------------- arguments
def arg_stores(words):
    """(register field, op) of the parameter stores: sw/fsw whose base is the
    address just formed by addi t0, s0, -k (rv_nodetype.hpp:825-831)."""
    out = []
    for k in range(1, len(words)):
        a, s = f(words[k - 1]), f(words[k])
        if a["op"] == 0x13 and a["rs1"] == 8 and s["op"] in (0x23, 0x27) and s["rs1"] == a["rd"]:
            out.append((s["rs2"], s["op"]))
    return out


def test_d9_int_argument_after_a_float_takes_the_next_register():
    # rv_generator_avx.cpp:187-205: reg_offset = int index + (float args before
    # it, minus 8 past 8); an int argument after one float goes to a1 (x11),
    # not a0 as in the scalar generator (rv_generator.cpp:247-281) and the ABI
    r = compile_src("function f(x : float, a : int) : int { return a; }")
    assert arg_stores(r["words"]) == [(10, 0x27), (11, 0x23)]  # fsw fa0; sw a1


def test_params_past_the_eighth_int():
    # rv_generator_avx.cpp:196-203: on_stack = (reg_offset > 7) & int_index is ANDed
    # with the index's value, then only its bit 1 selects the stack type (| 0b10),
    # and 8 & on_stack is subtracted.  Index 8 (0b1000): no bit 1 -> register
    # 8 - 8 = 0 -> a0.  Index 9 (0b1001): a1.  Index 10 (0b1010): bit 1 -> on the
    # stack, slot 10 - 8 = 2, whose word carries 4 * 2 in bits 31:20
    # (rv_nodetype.hpp:922, calc 4: (nd * 4) << 20).
    params = ", ".join(f"p{k} : int" for k in range(11))
    r = compile_src(f"function f({params}) : int {{ return p10; }}")
    st = arg_stores(r["words"])
    assert [reg for reg, _ in st] == [10, 11, 12, 13, 14, 15, 16, 17, 10, 11]
    w = r["words"]
    # the stack parameter: its two words follow the 10th store (sizes: rv_nodetype.hpp:256)
    k = [k for k in range(1, len(w)) if f(w[k - 1])["op"] == 0x13 and f(w[k - 1])["rs1"] == 8
         and f(w[k])["op"] in (0x23, 0x27)][-1] + 1
    assert int(w[k + 1]) >> 20 == 8


def test_d4_float_to_int_conversion_takes_a_float_destination():
    # rv_generator_avx.cpp:1209-1212: rd is allocated from the float file when
    # the word is FCVT.S.W or any OP-FP word; fcvt.w.s (0xC0007053) is OP-FP.
    # With GCC's ffz (include/utils.hpp:25-45, low 32 bits only) a float
    # destination always overflows to f5 (x37), field 5.  In "a + x @int", a is
    # live in t0 (x5) when the conversion runs, so an int destination would be
    # x6; the add then reads rs1 = 5 (a) and rs2 = 5 (the conversion).
    w = compile_src("function f(a : int, x : float) : int { return a + x @int; }")["words"]
    k = find(w, lambda x: (x & 0xFE00707F) == 0xC0007053)
    assert len(k) == 1
    cv = f(w[k[0]])
    assert cv["rd"] == 5
    add = f(w[k[0] + 1])
    assert rvdecode.mnemonic(int(w[k[0] + 1])) == "add" and add["rs1"] == 5 and add["rs2"] == 5
    # the flw before it is a LOAD-FP word, not OP-FP: an int-class destination (x6 free)
    assert f(w[k[0] - 1])["op"] == 0x07 and f(w[k[0] - 1])["rd"] == 6 == cv["rs1"]


@pytest.mark.parametrize("pad", range(4))
def test_d11_float_comparison_type_crosses_only_vector_boundaries(pad):
    # rv_generator_avx.cpp:270-305: for each 8-node vector in order, the result
    # types are loaded once, and every FLOAT node whose parent is a comparison
    # makes the parent FLOAT (scalar stores).  An inner comparison of floats
    # turns FLOAT while its operands' vector is processed; the outer comparison
    # turns FLOAT only if that happened in an earlier vector than the inner
    # comparison's own (its loaded type is stale otherwise).  The scalar
    # generator propagates the whole chain (rv_generator.cpp:304-313).
    # An outer "==" that is FLOAT is one feq.s; an INT one is sub + sltiu
    # (sizes rv_nodetype.hpp:246).
    body = "".join(f"a{k} : int = 1; " for k in range(pad))
    t = parse_source("function f(x : float, y : float) : int { " + body + "return (x < y) == 0; }")
    ty, rt, par = t.arrays()[0], t.arrays()[1], t.arrays()[2]
    lt = int(np.nonzero(ty == LESS_EXPR)[0][0])
    kids = np.nonzero(par == lt)[0]
    assert all(rt[c] == FLOAT for c in kids)
    outer_float = any(c // 8 < lt // 8 for c in kids)
    w = oracle.compile_tree(t)["words"]
    m = [rvdecode.mnemonic(int(x)) for x in w]
    assert m.count("flt.s") == 1
    assert ("feq.s" in m) == outer_float
    assert ("sltiu" in m) == (not outer_float)
    assert outer_float == (pad % 2 == 0)  # the layouts of this program alternate


def test_d2_float_argument_past_the_eighth_passes_a_location():
    # A 9th float argument becomes func_call_arg_float_as_int (0xF5,
    # rv_generator_avx.cpp:156-185).  isn_gen's non-condition test
    # ((ptype & 0xD0) == 0xD0 and child_idx != (ptype & 1), :636-640) also matches
    # 0xF5, so its child hands the parent slot loc + size instead of its
    # register (:703-708, :713).  The argument move (fmv.x.w) then reads a
    # physical register numbered by that location, and the argument's own
    # words, whose register nobody reads, are removed by optimize.
    params = ", ".join(f"p{k} : float" for k in range(9))
    src = f"function g({params}) : int {{ return 0; }} function f(x : float) : int {{ return g(x, x, x, x, x, x, x, x, x); }}"
    t = parse_source(src)
    o = oracle.Oracle(t)
    o.run(7)
    w = o.get("instr")
    fs = o.get("func_starts")
    body = w[fs[1]:]
    m = [rvdecode.mnemonic(int(x)) for x in body]
    assert m.count("flw") == 8  # 9 argument loads in the source, one removed
    mv = [k for k, x in enumerate(m) if x == "fmv.x.w"]
    assert len(mv) == 1
    rs1 = f(body[mv[0]])["rs1"]
    # that register: the location after the 9th argument's last word, in isn_gen's word indices
    o2 = oracle.Oracle(t)
    o2.run(2)
    locs = o2.get("locs")
    ty, par, ci = t.arrays()[0], t.arrays()[2], t.arrays()[4]
    call_args = [i for i in range(t.n) if ty[i] == 11]  # FUNC_CALL_ARG
    ninth = [a for a in call_args if ci[a] == 8]
    assert len(ninth) == 1
    child = int(np.nonzero(par == ninth[0])[0][0])
    assert ty[child] == DEREF_EXPR
    assert rs1 == (int(locs[child]) + 1) & 31  # a deref is one word (rv_nodetype.hpp:229)


# ---------------------------------------------------------------- regalloc structure
def words_without_jumps(words, s, e):
    """A function's words with every auipc+jalr pair blanked: calls (jalr rd =
    ra) reach other functions, and return jumps go to the function's epilogue,
    whose offset moves when the function's own words change."""
    w = [int(x) for x in words[s:e]]
    out = list(w)
    for k in range(len(w) - 1):
        if is_auipc_ra(w[k]) and (w[k + 1] & 0x7F) == 0x67:
            out[k] = out[k + 1] = None
    return out


def test_functions_outside_lane_0_are_independent_of_other_bodies():
    # rv_generator_avx.cpp:1041-1418 walks 8 functions per vector; only lane 0
    # writes region 0 of register_state (:1332-1354), and every swap lookup
    # reads region 0 (:1290-1312).  A function outside lane 0 (index not 0 mod 8)
    # therefore allocates from its own instructions alone: changing the body of
    # another function (here function 1, keeping its declarations, so symbol
    # ids stay) leaves its words unchanged apart from call offsets and return
    # jumps (both cross function boundaries).  Two other effects are global and
    # are kept out: preprocess's float propagation (D11) depends on each node's
    # 8-node vector, counted from node 0 of the program, so the change adds a
    # multiple of 8 nodes; and a word slot shared by two writers of equal depth
    # goes to the later one in the depth sort over all nodes (libstdc++
    # introsort, not stable, rv_generator_avx.cpp:539-541), so the program has no
    # while loops or && / || (the nodes that write such shared slots) and the
    # oracle reports no such slot.
    base, _ = W.program_source(73, 12, p_while=0.0, p_if=0.35, p_arith=0.7, p_compare=0.3)
    funcs = base.split("function ")[1:]
    assert len(funcs) == 12
    f1 = funcs[1].rstrip()
    assert f1.endswith("}")
    changed = f1[:-1] + " if 0 { } if 1 { } 5; }"  # 2 x (if, literal, list) + (expression, literal)
    srcs = ["".join("function " + (b1 if k == 1 else fb) for k, fb in enumerate(funcs))
            for b1 in (funcs[1], changed + "\n")]
    ta, tb = (parse_source(x) for x in srcs)
    assert tb.n - ta.n == 8
    res = []
    for t in (ta, tb):
        assert not np.isin(t.arrays()[0], (24, 25, 41)).any()  # LAND_EXPR, LOR_EXPR, WHILE_DUMMY
        o = oracle.Oracle(t)
        o.run(7)
        assert o.ties == 0
        res.append(o.result())
    a, b = res
    assert a["func_sizes"][1] != b["func_sizes"][1]
    same = 0
    for k in range(12):
        if k % 8 == 0 or k == 1:
            continue
        wa = words_without_jumps(a["words"], a["func_starts"][k], a["func_ends"][k])
        wb = words_without_jumps(b["words"], b["func_starts"][k], b["func_ends"][k])
        assert wa == wb, k
        same += 1
    assert same == 9


def test_d5_lane_0_register_marked_by_a_lookup_from_a_higher_group():
    # Functions f0 (group 0, lane 0) and f8 (group 1, lane 0), both
    # "return 1 + 2;", with seven empty functions between them (lanes 1-7 of
    # group 0; their words name no virtual register).  Both bodies are, from
    # step 6: lui v1; addi v2 <- v1; lui v3; addi v4 <- v3; add v5 <- v2, v4;
    # add a0 <- v5 (v = word + 64; f0's v4 is 73).  Rules (rv_generator_avx.cpp):
    # a read of a register live in the reader's lifetime looks up region 0 of
    # the step-start snapshot and marks the virtual found there (:1244-1318);
    # groups run in order within a step, each marking before its commit resets
    # the virtuals it names (:1319-1398); a swapped operand is read through a
    # temp, 5 for rs1 and 6 for rs2, and only the temp leaves the reader's
    # lifetime (:1166-1199); every lane-0 commit writes region 0 (:1332).
    #   step 6: f0 then f8 write region0[5]; f8's v1 wins.
    #   step 7: f0 reads x5 and marks f8's v1; f8's read then goes through temp
    #           5, which is the same register: v2 -> x5 in both.
    #   step 8: both lui -> x6; region0[6] = f8's v3.
    #   step 9: f0 reads x6 and marks f8's v3.  f8 reads it through temp 5, so
    #           its lifetime keeps x6 and its addi takes x5: region0[5] = f8's
    #           v4, region0[6] keeps f0's v4 (73).
    #   step 10: f0's add reads x5 (marks f8's v4) and x6 (its own 73, then
    #           resets it).  f8's add finds its v4 marked, reads it through
    #           temp 6, which is live in its lifetime: the lookup of region0[6]
    #           marks f0's 73 after f0's reset.  f0 never names 73 again.
    # So 73 ends swapped: emit (:1586-1650) stores it after its addi
    # (sw, rd_stack_offset 0: imm = (0 - 1) * -4 = 4 in the reference's split
    # encoding, register 6 in bits 19:15) and loads it into temp 6 before the add
    # (lw x6, (0 - 1) * 4(x0)).  Compiled alone, f0 spills nothing.  f8 ends
    # with nothing swapped (its own reads reset every mark on its registers),
    # but its commits kept the temps as registers: v3 (read through temp 5 at
    # step 9) is x5 and v4 (read through temp 6 at step 10) is x6, so its
    # second lui writes t0 and its addi reads t0 (clobbering v2: the
    # reference's output as it is).
    empties = "".join(f"function e{k}() : void {{ }} " for k in range(1, 8))
    two = oracle.compile_tree(parse_source("function f0() : int { return 1 + 2; } " + empties +
                                           "function f8() : int { return 1 + 2; }"))
    alone = [int(x) for x in compile_src("function f0() : int { return 1 + 2; }")["words"]]
    k = alone.index(0x00230313)  # addi t1, t1, 2
    assert alone[k + 1] == 0x006282B3  # add t0, t0, t1
    spill_store = 0x2023 | ((4 & 0x1F) << 7) | ((4 & 0xFE) << 19) | (6 << 15)
    spill_load = ((-4 & 0xFFF) << 20) | 0x2003 | (6 << 7)
    assert (spill_store, spill_load) == (0x00232223, 0xFFC02303)
    want = alone[:k + 1] + [spill_store, spill_load] + alone[k + 1:]
    s, e = two["func_starts"], two["func_ends"]
    assert [int(x) for x in two["words"][s[0]:e[0]]] == want
    lui_t0, addi_t1_t0 = 0x37 | (5 << 7), 0x13 | (6 << 7) | (5 << 15) | (2 << 20)
    assert alone[k - 1] == 0x37 | (6 << 7)  # lui t1 alone
    want8 = alone[:k - 1] + [lui_t0, addi_t1_t0] + alone[k + 1:]
    assert [int(x) for x in two["words"][s[8]:e[8]]] == want8


@pytest.mark.parametrize("seed", [100, 103, 107])
def test_d6_both_operands_spilled_leave_one_load_and_a_hole(seed):
    # Emit (rv_generator_avx.cpp:1552-1650): a swapped rs1 loads at base
    # (lw t0), a swapped rs2 at base + (rs2_load_offset > 0), evaluated while
    # rs2_load_offset is still -1, i.e. also at base (lw t1, written after the
    # rs1 load, so it wins); the word itself goes to base + 2.  So a word whose
    # operands are both swapped is preceded by lw t1 and a zero word (which
    # fix_jumps deletes), and reads t0 (f5 for OP-FP words) without a load;
    # the load itself writes x6 even for a float operand.  The scalar generator
    # (rv_generator.cpp:1370-1376) puts the rs2 load at base + 1.
    t = parse_source(W.program_source(seed, 10)[0])
    o = oracle.Oracle(t)
    o.run(4)
    rs1, rs2 = o.get("rs1").astype(np.int64), o.get("rs2").astype(np.int64)
    used = o.get("used")
    o.run(5)
    sw, off = o.get("sym_sw"), o.get("ra_offsets")
    # the allocated words (registers still in their own arrays; postprocess ORs them in)
    w1, rd1, a1, a2 = (o.get(k) for k in ("instr", "rd", "rs1", "rs2"))
    both = [i for i in range(len(rs1)) if used[i] and rs1[i] > 63 and rs2[i] > 63 and sw[rs1[i] - 64] and sw[rs2[i] - 64]]
    assert both
    for i in both:
        b = int(off[i])
        load = f(w1[b])
        assert load["op"] == 0x03 and load["f3"] == 2 and load["rs1"] == 0 and rd1[b] == 6, i
        assert int(w1[b + 1]) == 0, i
        fp = (int(w1[b + 2]) & 0x7F) == 0x53  # OP-FP words read float temps f5 / f6 (37 / 38)
        assert (a1[b + 2], a2[b + 2]) == ((37, 38) if fp else (5, 6)), i
    # every zero word is gone from the final output
    o.run(7)
    assert (o.get("instr") != 0).all()
