"""Independent RISC-V (RV32IM + F subset) decoder for the tests: word -> mnemonic
and fields, written from the RISC-V unprivileged ISA manual's encoding tables."""


def sext(v, bits):
    return v - (1 << bits) if v & (1 << (bits - 1)) else v


def fields(w):
    return dict(op=w & 0x7F, rd=(w >> 7) & 31, f3=(w >> 12) & 7, rs1=(w >> 15) & 31, rs2=(w >> 20) & 31,
                f7=w >> 25, imm_i=sext(w >> 20, 12), imm_s=sext(((w >> 25) << 5) | ((w >> 7) & 31), 12),
                imm_u=w & 0xFFFFF000)


R_OPS = {(0x00, 0): "add", (0x20, 0): "sub", (0x00, 1): "sll", (0x00, 2): "slt", (0x00, 3): "sltu",
         (0x00, 4): "xor", (0x00, 5): "srl", (0x20, 5): "sra", (0x00, 6): "or", (0x00, 7): "and",
         (0x01, 0): "mul", (0x01, 4): "div", (0x01, 6): "rem"}
I_OPS = {0: "addi", 2: "slti", 3: "sltiu", 4: "xori", 6: "ori", 7: "andi"}
B_OPS = {0: "beq", 1: "bne", 4: "blt", 5: "bge", 6: "bltu", 7: "bgeu"}


def mnemonic(w):
    f = fields(w)
    op, f3, f7 = f["op"], f["f3"], f["f7"]
    if op == 0x33:
        return R_OPS.get((f7, f3), "?op")
    if op == 0x13:
        return I_OPS.get(f3, "?opimm")
    if op == 0x37:
        return "lui"
    if op == 0x17:
        return "auipc"
    if op == 0x03:
        return {2: "lw"}.get(f3, "?load")
    if op == 0x07:
        return {2: "flw"}.get(f3, "?fload")
    if op == 0x23:
        return {2: "sw"}.get(f3, "?store")
    if op == 0x27:
        return {2: "fsw"}.get(f3, "?fstore")
    if op == 0x67:
        return "jalr"
    if op == 0x6F:
        return "jal"
    if op == 0x63:
        return B_OPS.get(f3, "?branch")
    if op == 0x73:
        return "ecall" if w == 0x73 else "system"
    if op == 0x53:
        if f7 == 0x00: return "fadd.s"
        if f7 == 0x04: return "fsub.s"
        if f7 == 0x08: return "fmul.s"
        if f7 == 0x0C: return "fdiv.s"
        if f7 == 0x10: return {0: "fsgnj.s", 1: "fsgnjn.s", 2: "fsgnjx.s"}.get(f3, "?fsgnj")
        if f7 == 0x50: return {0: "fle.s", 1: "flt.s", 2: "feq.s"}.get(f3, "?fcmp")
        if f7 == 0x60: return "fcvt.w.s"
        if f7 == 0x68: return "fcvt.s.w"
        if f7 == 0x70: return "fmv.x.w"
        if f7 == 0x78: return "fmv.w.x"
        return "?fp"
    return "?"
