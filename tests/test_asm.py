"""The oracle's to_asm restatement (rv_generator::to_asm, src/codegen/rv_generator.cpp:114-120,
through rvdisasm::instruction, src/disassembler.cpp:453-462), pinned three ways:

* known answers: words built here from the RISC-V unprivileged manual's encodings, where the
  reference prints what the manual says (so a misread field or register name fails);
* the reference's own quirks, each derived by hand from the lines of src/disassembler.cpp cited
  next to it (switch fall-throughs, the fsw register, the branch and jump immediates);
* every word of generated programs against the independent decoder in rvdecode.py.
"""
import numpy as np
import pytest

import oracle
import rvdecode
from ptilopsis_amd import parse_source, workloads as W

X = {n: i for i, n in enumerate("zero ra sp gp tp t0 t1 t2 s0 s1 a0 a1 a2 a3 a4 a5 a6 a7 s2 s3 s4 s5 s6 s7 "
                                "s8 s9 s10 s11 t3 t4 t5 t6".split())}


def R(f7, rs2, rs1, f3, rd, op):
    return (f7 << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def I(imm, rs1, f3, rd, op):
    return ((imm & 0xFFF) << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def S(imm, rs2, rs1, f3, op):
    return (((imm >> 5) & 0x7F) << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | ((imm & 0x1F) << 7) | op


def B(imm, rs2, rs1, f3):
    return ((((imm >> 12) & 1) << 31) | (((imm >> 5) & 0x3F) << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12)
            | (((imm >> 1) & 0xF) << 8) | (((imm >> 11) & 1) << 7) | 0x63)


def U(imm20, rd, op):
    return ((imm20 & 0xFFFFF) << 12) | (rd << 7) | op


def J(imm, rd):
    return ((((imm >> 20) & 1) << 31) | (((imm >> 1) & 0x3FF) << 21) | (((imm >> 11) & 1) << 20)
            | (((imm >> 12) & 0xFF) << 12) | (rd << 7) | 0x6F)


a0, a1, a2, a3, sp, ra, t0, s0 = X["a0"], X["a1"], X["a2"], X["a3"], X["sp"], X["ra"], X["t0"], X["s0"]

# (word, line) where the reference's text is the manual's reading
KNOWN = [
    (I(-5, a1, 0, a0, 0x13), "addi a0, a1, -5"),
    (I(12, sp, 2, t0, 0x03), "lw t0, 12(sp)"),
    (I(0, ra, 0, 0, 0x67), "jalr zero, 0(ra)"),
    (S(-4, ra, sp, 2, 0x23), "sw ra, -4(sp)"),
    (S(-8, s0, sp, 2, 0x23), "sw s0, -8(sp)"),
    (R(0, a2, a1, 0, a0, 0x33), "add a0, a1, a2"),
    (R(0x20, a2, a1, 0, a0, 0x33), "sub a0, a1, a2"),
    (R(1, a1, a0, 0, a0, 0x33), "mul a0, a0, a1"),
    (R(1, a1, a0, 4, a0, 0x33), "div a0, a0, a1"),
    (R(1, a1, a0, 6, a0, 0x33), "rem a0, a0, a1"),
    (R(0x20, a2, a1, 5, a0, 0x33), "sra a0, a1, a2"),
    (R(0, a2, a1, 2, a0, 0x33), "slt a0, a1, a2"),
    (U(0x12345, s0, 0x37), "lui s0, 305418240"),
    (U(0xFFFFF, t0, 0x37), "lui t0, -4096"),
    (U(0, ra, 0x17), "auipc ra, 0"),
    (B(8, a1, a0, 0), "beq a0, a1, 0x8"),
    (B(100, a1, a0, 4), "blt a0, a1, 0x64"),
    (B(0, a1, a0, 1), "bne a0, a1, 0"),  # std::showbase prints zero bare
    (J(2048, ra), "jal ra, 2048"),
    (J(-4, 0), "jal zero, -4"),
    (R(0x00, 12, 11, 7, 10, 0x53), "fadd.s fa0, fa1, fa2"),
    (R(0x08, 12, 11, 7, 10, 0x53), "fmul.s fa0, fa1, fa2"),
    (R(0x0C, 12, 11, 7, 10, 0x53), "fdiv.s fa0, fa1, fa2"),
    (R(0x10, 12, 11, 0, 10, 0x53), "fsgnj.s fa0, fa1, fa2"),
    (R(0x78, 0, a0, 0, 0, 0x53), "fmv.w.x ft0, a0"),
    (R(0x70, 0, 0, 0, a0, 0x53), "fmv.x.w a0, ft0"),
    (R(0x60, 0, 10, 1, a0, 0x53), "fcvt.w.s a0, fa0"),
    (R(0x68, 0, a0, 7, 10, 0x53), "fcvt.s.w fa0, a0"),
    (I(8, sp, 2, 10, 0x07), "flw fa0, 8(sp)"),
]

# (word, line) where the reference departs from the manual; derived from the cited lines
QUIRKS = [
    # format_args prints the whole 12-bit I immediate, funct7 bits included (:317-331)
    (I(0x400 | 3, a0, 5, a0, 0x13), "srai a0, a0, 1027"),
    # an fsw names its register by the rs1 field (:357-361)
    (S(4, 11, sp, 2, 0x27), "fsw ft2, 4(sp)"),
    # BRANCH funct3 2 has no name: the case falls into LOAD (:92-110)
    (B(8, a1, a0, 2), "lw a0, a1, 0x8"),
    # LOAD funct3 3 falls through STORE into OP-IMM (:103-126); I format, load style
    (I(0, sp, 3, a0, 0x03), "sltiu a0, 0(sp)"),
    # SYSTEM other than ecall/ebreak falls into flw (:177-182); I format, plain
    (I(0x300, a1, 1, a0, 0x73), "flw a0, a1, 768"),
    (0x00000073, "ecall zero, zero, 0"),
    # OP funct7 0x20 with funct3 1 falls down the funct3 cases to sra (:131-175)
    (R(0x20, a2, a1, 1, a0, 0x33), "sra a0, a1, a2"),
    # ... and with funct3 6 out of OP into fence
    (R(0x20, a2, a1, 6, a0, 0x33), "fence a0, a1, a2"),
    # OP-IMM funct3 5 with another funct7 falls into ori (:122-128)
    (I(0x200 | 1, a0, 5, a0, 0x13), "ori a0, a0, 513"),
    # the B immediate skips instr[30] and sign-fills bits 31:11 (:366-372): -8 prints as -1032
    (B(-8, a1, a0, 0), "beq a0, a1, 0xfffffffffffffbf8"),
    # the J immediate sign-fills bits 31:19 (:384-392)
    (J(-(1 << 20) + 4096, 0), "jal zero, -520192"),
    # compare results are printed as float registers (:270-278 default case)
    (R(0x50, 11, 10, 0, a0, 0x53), "fle.s fa0, fa0, fa1"),
    (R(0x50, 11, 10, 2, a0, 0x53), "feq.s fa0, fa0, fa1"),
    # fclass.s takes the default three-float format
    (R(0x70, 0, 10, 1, a0, 0x53), "fclass.s fa0, fa0, ft0"),
    # OP-FP funct7 without a case: "unknown", three float registers
    (R(0x7F, 0, 0, 0, 0, 0x53), "unknown ft0, ft0, ft0"),
    # fused multiply-adds print integer names (:395-403)
    ((a3 << 27) | R(0, a2, a1, 0, a0, 0x43), "fmadd.s a0, a1, a2, a3"),
    (U(0x80000, a0, 0x37), "lui a0, -2147483648"),
    (0, "(zero) "),
    (0xFFFFFFFF, "unknown "),
]


@pytest.mark.parametrize("word,line", KNOWN + QUIRKS, ids=[f"{w:08x}" for w, _ in KNOWN + QUIRKS])
def test_asm_line(word, line):
    assert oracle.to_asm(np.array([word], np.uint32)).decode() == line + "\n"


def test_asm_lines_concatenate():
    words = np.array([w for w, _ in KNOWN + QUIRKS], np.uint32)
    want = "".join(line + "\n" for _, line in KNOWN + QUIRKS)
    assert oracle.to_asm(words).decode() == want
    assert oracle.to_asm(np.zeros(0, np.uint32)) == b""


@pytest.mark.parametrize("seed,nf,base", [(3, 6, "base"), (101, 2, "stress"), (201, 8, "float")])
def test_asm_against_decoder(seed, nf, base):
    b = {"base": W.BASE, "stress": W.STRESS, "float": dict(W.BASE, float_frac=0.9)}[base]
    src, _ = W.program_source(seed, nf, base=b)
    words = oracle.compile_tree(parse_source(src))["words"]
    lines = oracle.to_asm(words).decode().split("\n")
    assert lines[-1] == "" and len(lines) == len(words) + 1
    seen = set()
    for w, line in zip(words.tolist(), lines):
        name, _, args = line.partition(" ")
        want = rvdecode.mnemonic(w)
        if want.startswith("?") or want == "system":  # the decoder names only the bare ecall word
            continue
        assert name == want, (hex(w), line, want)
        seen.add(name)
        f = rvdecode.fields(w)
        if w & 0x7F == 0x33:
            assert args == ", ".join(rvdecode_name(f[k]) for k in ("rd", "rs1", "rs2")), line
        elif w & 0x7F in (0x03, 0x67):
            assert args == f"{rvdecode_name(f['rd'])}, {f['imm_i']}({rvdecode_name(f['rs1'])})", line
        elif w & 0x7F == 0x23:
            assert args == f"{rvdecode_name(f['rs2'])}, {f['imm_s']}({rvdecode_name(f['rs1'])})", line
        elif w & 0x7F == 0x13 and name not in ("slli", "srli", "srai"):
            assert args == f"{rvdecode_name(f['rd'])}, {rvdecode_name(f['rs1'])}, {f['imm_i']}", line
    assert {"sw", "lw", "addi", "jalr"} <= seen


def rvdecode_name(r):
    return list(X)[r]
