"""The listing of rv_generator::print(os, true), i.e. rvdisasm::disassemble
(src/disassembler.cpp:411-451), restated in the oracle: pinned by answers
worked out by hand from that code (header, label lines, column widths, the
colour escapes, absolute branch targets, labels printed in hex after word 0
because format_args restores the stream's hex flags), and against the oracle's
to_asm (pinned elsewhere) on random words with the colours stripped."""
import re

import numpy as np

import oracle

E, W, I, N, R = "\033[38;5;114m", "\033[37m", "\033[38;5;114m", "\033[38;5;186m", "\033[38;5;117m"


def test_listing_known_answer():
    words = [0x00500513, 0x00000463, 0x008000EF]  # addi a0, zero, 5; beq zero, zero, +8; jal ra, 8
    want = (f"{E}3{W} instructions, starting at {I}0x00{W}\n"
            + " " * 21 + "func0:\n"
            + f"{E}0{W}   00:  {E}00{W} {E}50{W} {E}05{W} {E}13{W}   {N}addi     {W} {R}a0{W}, {R}zero{W}, {I}5{W}\n"
            + f"{E}1{W}   04:  {E}00{W} {E}00{W} {E}04{W} {E}63{W}   {N}beq      {W} {R}zero{W}, {R}zero{W}, {I}0xc{W}\n"
            + " " * 21 + "func1:\n"
            + f"{E}2{W}   08:  {E}00{W} {E}80{W} {E}00{W} {E}ef{W}   {N}jal      {W} {R}ra{W}, {I}8{W}\n")
    assert oracle.listing(words, [0, 2]).decode() == want


def test_listing_labels_hex_after_word_zero():
    got = oracle.listing([0x00000013] * 12, list(range(12))).decode().split("\n")
    labels = [ln.strip() for ln in got if ln.strip().startswith("func")]
    assert labels == ["func0:", "func1:", "func2:", "func3:", "func4:", "func5:", "func6:", "func7:", "func8:",
                      "func9:", "funca:", "funcb:"]
    # 12 words: two decimal digits for the index, two hex digits for the address
    assert got[1] == " " * (2 + 3 + 2 + 2 + 12 + 1) + "func0:"
    assert got[-2].startswith(f"{E}11{W}   2c:  ")


def test_listing_zero_branch_target_is_bare():
    # beq zero, zero with offset -1028 (bit 10 of the offset is never taken,
    # :366-372) at byte 1028: target 0, which showbase prints as a bare "0"
    w = 0xBE000EE3
    lst = oracle.listing([0x00000013] * 257 + [w], []).decode().split("\n")
    assert lst[-2].endswith(f"{R}zero{W}, {R}zero{W}, {I}0{W}")
    assert oracle.to_asm([w]).decode() == "beq zero, zero, 0xfffffffffffffbfc\n"


STRIP = re.compile("\033\\[[0-9;]*m")


def test_listing_matches_to_asm_on_random_words():
    rng = np.random.default_rng(7)
    words = rng.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    opcodes = np.array([0x33, 0x53, 0x67, 0x03, 0x13, 0x0F, 0x73, 0x07, 0x23, 0x27, 0x63, 0x37, 0x17, 0x6F, 0x43],
                       np.uint32)
    words[:2500] = (words[:2500] & ~np.uint32(0x7F)) | opcodes[rng.integers(0, len(opcodes), 2500)]
    starts = np.sort(rng.choice(3000, 40, replace=False)).astype(np.uint32)
    lst = STRIP.sub("", oracle.listing(words, starts).decode()).split("\n")[1:-1]
    plain = oracle.to_asm(words).decode().split("\n")[:-1]
    body = [ln for ln in lst if not ln.strip().startswith("func")]
    assert len(body) == 3000 and len(lst) == 3000 + 40
    for i, (ln, pl) in enumerate(zip(body, plain)):
        idx, rest = ln[:4], ln[4:]
        assert int(idx) == i
        addr, rest = rest[3:].split(": ", 1)
        assert int(addr, 16) == 4 * i and len(addr) == 4
        assert rest[:12] == " " + " ".join(f"{(int(words[i]) >> s) & 0xFF:02x}" for s in (24, 16, 8, 0))
        assert rest[12:15] == "   " and rest[24] == " "
        name, args = rest[15:24].rstrip(), rest[25:]
        pname, _, pargs = pl.partition(" ")
        assert name == pname, i
        if (int(words[i]) & 0x7F) == 0x63:  # branch: absolute target
            t = args.rsplit(", ", 1)[1]
            pt = pargs.rsplit(", ", 1)[1]
            off = int(pt, 16) if pt != "0" else 0
            tgt = int(t, 16) if t != "0" else 0
            assert tgt == (4 * i + off) % (1 << 64), i
            assert args.rsplit(", ", 1)[0] == pargs.rsplit(", ", 1)[0]
        else:
            assert args == pargs, i
