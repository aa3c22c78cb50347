"""The instruction-selection spec: the shipped packed form (rv_spec.h) against
the oracle's per-type tables (two separate transcriptions of
include/codegen/rv_nodetype.hpp), and every template decoded by an independent
decoder against the mnemonic the reference's comments name."""
import os
import subprocess

import pytest

from cpp_helpers import compile_cpp
import rvdecode

HERE = os.path.dirname(os.path.abspath(__file__))

INT, FLOAT, INT_REF, FLOAT_REF, VOID, INVALID = 2, 3, 4, 5, 1, 0


@pytest.fixture(scope="module")
def rows():
    src = open(os.path.join(HERE, "spec_dump.cpp")).read().replace('"../', '"' + os.path.dirname(HERE) + "/")
    exe = compile_cpp("spec_dump", src)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    table, pmap = {}, []
    for line in out:
        p = line.split()
        if p[0] == "map":
            pmap.append((int(p[2]), int(p[3])))
            continue
        code, o, r = int(p[0]), int(p[1]), int(p[2])
        table[(code, o, r)] = ([int(x) for x in p[3:13]], [int(x) for x in p[13:23]])
    return table, pmap


def test_pareas_map_agrees(rows):
    _, pmap = rows
    assert len(pmap) == 44 and all(a == b for a, b in pmap)


def test_packed_spec_matches_oracle_tables(rows):
    table, _ = rows
    bad = []
    for (code, o, r), (orc, prod) in table.items():
        has, word, gout, op1, op2, hout, pai, ic, jt, size = orc
        phas, pword, pgout, pop1, pop2, phout, ppai, pic, pjt, psize = prod
        if has != phas:
            bad.append(((code, o, r), "has_instr", has, phas)); continue
        if size != psize:
            bad.append(((code, o, r), "size", size, psize))
        if o == 0 and pai != ppai:  # read for every node at offset 0
            bad.append(((code, o, r), "parent_arg", pai, ppai))
        if not has:
            continue
        for nm, a, b in (("word", word, pword), ("out", gout, pgout), ("op1", op1, pop1), ("op2", op2, pop2),
                         ("has_output", hout, phout), ("parent_arg", pai, ppai), ("iconst", ic, pic), ("jt", jt, pjt)):
            if a != b:
                bad.append(((hex(code), o, r), nm, a, b))
    assert not bad, bad[:20]


EXPECT = {
    # (type code, offset): {data type: mnemonic}; from the comments in rv_nodetype.hpp:563-858
    (0x04, 0): {INT: "add", FLOAT: "fadd.s"}, (0x05, 0): {INT: "sub", FLOAT: "fsub.s"},
    (0x06, 0): {INT: "mul", FLOAT: "fmul.s"}, (0x07, 0): {INT: "div", FLOAT: "fdiv.s"},
    (0x08, 0): {INT: "rem"}, (0x09, 0): {INT: "and"}, (0x0A, 0): {INT: "or"}, (0x0B, 0): {INT: "xor"},
    (0x0C, 0): {INT: "sll"}, (0x0D, 0): {INT: "sra"}, (0x0E, 0): {INT: "srl"},
    (0x11, 0): {INT: "xori"}, (0x12, 0): {INT: "sltiu"}, (0x13, 0): {INT: "sub", FLOAT: "fsgnjn.s"},
    (0x14, 0): {INT: "lui", FLOAT: "lui"}, (0x14, 1): {INT: "addi", FLOAT: "addi"}, (0x14, 2): {FLOAT: "fmv.w.x"},
    (0x15, 0): {INT: "fcvt.w.s", FLOAT: "fcvt.s.w"}, (0x16, 0): {INT: "lw", FLOAT: "flw"},
    (0x17, 0): {INT: "sw", FLOAT: "fsw"}, (0x17, 1): {INT: "add", FLOAT: "fsgnj.s"},
    (0x18, 0): {r: "addi" for r in range(6)}, (0x19, 0): {r: "addi" for r in range(6)},
    (0x1B, 0): {r: "sw" for r in range(6)}, (0x1B, 1): {r: "sw" for r in range(6)},
    (0x1B, 2): {r: "sub" for r in range(6)}, (0x1B, 3): {r: "add" for r in range(6)},
    (0x1C, 0): {VOID: "add", INT: "add", FLOAT: "fsgnj.s"}, (0x1C, 1): {r: "jalr" for r in range(6)},
    (0x1D, 0): {r: "add" for r in range(6)}, (0x1D, 1): {r: "lw" for r in range(6)},
    (0x1D, 2): {r: "lw" for r in range(6)}, (0x1D, 3): {r: "jalr" for r in range(6)},
    (0x1F, 0): {r: "jalr" for r in range(6)}, (0x1F, 2): {INT: "add", FLOAT: "fsgnj.s"},
    (0x20, 0): {r: "andi" for r in range(6)}, (0x20, 1): {r: "ori" for r in range(6)},
    (0xD0, 0): {r: "beq" for r in range(6)}, (0xD1, 0): {r: "beq" for r in range(6)},
    (0xD1, 1): {r: "jalr" for r in range(6)}, (0xD2, 0): {r: "beq" for r in range(6)},
    (0xD2, 1): {r: "jalr" for r in range(6)},
    (0xE0, 0): {INT: "sub", FLOAT: "feq.s"}, (0xE0, 1): {INT: "sltiu"},
    (0xE1, 0): {INT: "sub", FLOAT: "feq.s"}, (0xE1, 1): {INT: "sltu", FLOAT: "sltiu"},
    (0xE2, 0): {INT: "slt", FLOAT: "flt.s"}, (0xE3, 0): {INT: "slt", FLOAT: "flt.s"},
    (0xE4, 0): {INT: "slt", FLOAT: "fle.s"}, (0xE4, 1): {INT: "sltiu"},
    (0xE5, 0): {INT: "slt", FLOAT: "fle.s"}, (0xE5, 1): {INT: "sltiu"},
    (0xF0, 0): {INT_REF: "sw", FLOAT_REF: "fsw"}, (0xF1, 0): {r: "fmv.x.w" for r in range(6)},
    (0xF2, 0): {INT: "sw", FLOAT: "fsw"}, (0xF4, 0): {INT: "add", FLOAT: "fsgnj.s"},
    (0xF5, 0): {r: "fmv.x.w" for r in range(6)}, (0xF6, 0): {INT: "sw", FLOAT: "fsw"},
}

# register fields the comments spell out (rd, rs1, rs2, immediate)
EXPECT_FIELDS = {
    (0x1B, 0): dict(rs1=2, rs2=1, imm_s=-4), (0x1B, 1): dict(rs1=2, rs2=8, imm_s=-8),
    (0x1B, 2): dict(rd=2, rs1=2, rs2=8), (0x1B, 3): dict(rd=8, rs1=8, rs2=2),
    (0x1D, 0): dict(rd=2, rs1=2, rs2=8), (0x1D, 1): dict(rd=1, rs1=2, imm_i=-4),
    (0x1D, 2): dict(rd=8, rs1=8, imm_i=-8), (0x1D, 3): dict(rd=0, rs1=1, imm_i=0),
    (0x1F, 0): dict(rd=1, rs1=0, imm_i=0), (0x20, 0): dict(rd=2, rs1=2), (0x20, 1): dict(rd=2, rs1=2),
    (0x18, 0): dict(rs1=8), (0x19, 0): dict(rs1=8), (0x11, 0): dict(imm_i=-1), (0x12, 0): dict(imm_i=1),
}


def test_templates_decode_to_the_named_instructions(rows):
    table, _ = rows
    bad = []
    for (code, o), per in EXPECT.items():
        for r, mn in per.items():
            orc, prod = table[(code, o, r)]
            assert orc[0] == 1 and prod[0] == 1, (hex(code), o, r)
            got = rvdecode.mnemonic(prod[1])
            if got != mn:
                bad.append((hex(code), o, r, mn, got, hex(prod[1])))
            exp_f = EXPECT_FIELDS.get((code, o), {})
            f = rvdecode.fields(prod[1])
            for k, v in exp_f.items():
                if f[k] != v:
                    bad.append((hex(code), o, r, k, v, f[k]))
    assert not bad, bad


def test_unnamed_emitting_slots_are_the_placeholder(rows):
    # every emitting slot not listed above carries the reference's 0b1110011 placeholder
    table, _ = rows
    for (code, o, r), (orc, prod) in table.items():
        if not prod[0] or r in EXPECT.get((code, o), {}):
            continue
        if (code, o, r) in ((0x1C, 0, INT_REF), (0x1C, 0, FLOAT_REF)):
            assert prod[1] == 0  # the reference initialises only four of six entries
            continue
        assert prod[1] in (0x73, 0x37, 0x13, 0x67, 0x63) or rvdecode.mnemonic(prod[1]) == "ecall", (hex(code), o, r)
