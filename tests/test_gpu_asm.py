"""GPU text output (ptl_result_asm / ptl_disassemble) against the oracle's to_asm
restatement, byte for byte, through the C ABI."""
import sys

import numpy as np
import pytest

import oracle
from ptilopsis_amd import parse_source, workloads as W
from ptilopsis_amd.generator import PtlError

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def first_diff(a: bytes, b: bytes) -> str:
    k = next((i for i in range(min(len(a), len(b))) if a[i] != b[i]), min(len(a), len(b)))
    return f"lengths {len(a)} / {len(b)}, first difference at byte {k}: {a[k - 40:k + 40]!r} / {b[k - 40:k + 40]!r}"


def test_gpu_asm_random_words(gpu_ctx):
    # every opcode / funct3 / funct7 combination the formatter distinguishes, and random words;
    # 70k words span several scan tiles and threads whose lines share 4-byte words
    rng = np.random.default_rng(0x5EED0A5A)
    ops = np.array([0x37, 0x17, 0x6F, 0x67, 0x63, 0x03, 0x23, 0x13, 0x33, 0x0F, 0x73, 0x07, 0x27, 0x43, 0x47,
                    0x4B, 0x4F, 0x53, 0x7F, 0x00], np.uint32)
    w = rng.integers(0, 1 << 32, 60000, dtype=np.uint64).astype(np.uint32)
    w[:40000] = (w[:40000] & ~np.uint32(0x7F)) | ops[rng.integers(0, len(ops), 40000)]
    extra = np.array([0, 0xFFFFFFFF, 0x73, 0x80000037, 0xFE112E23, 0x8067], np.uint32)
    w = np.concatenate([w, extra, np.zeros(10000, np.uint32)])
    say("random words", len(w))
    got = gpu_ctx.disassemble(w)
    want = oracle.to_asm(w)
    assert got == want, first_diff(got, want)


@pytest.mark.parametrize("n", [0, 1, 5, 4095, 4096, 4097])
def test_gpu_asm_sizes(gpu_ctx, n):
    w = np.random.default_rng(n).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    assert gpu_ctx.disassemble(w) == oracle.to_asm(w)


def test_gpu_result_asm_single_and_batch(gpu_ctx):
    t = W.single_50k()
    say("single_50k to_asm", t.n)
    r = gpu_ctx.compile(t)
    got = gpu_ctx.result_asm(0)
    want = oracle.to_asm(r["words"])
    assert got == want, first_diff(got, want)
    with pytest.raises(PtlError) as e:
        gpu_ctx.result_asm(1)
    assert e.value.status == -1
    trees = [parse_source(W.program_source(40 + k, 3 + k)[0]) for k in range(3)]
    res = gpu_ctx.compile_batch(trees)
    for p in range(3):
        assert gpu_ctx.result_asm(p) == oracle.to_asm(res[p]["words"])
    # formatting a result twice gives the same text (the arena is given back)
    assert gpu_ctx.result_asm(2) == oracle.to_asm(res[2]["words"])


def test_gpu_result_asm_after_disassemble_refused(gpu_ctx):
    gpu_ctx.compile(parse_source(W.program_source(7, 2)[0]))
    gpu_ctx.disassemble(np.array([0x8067], np.uint32))
    with pytest.raises(PtlError) as e:
        gpu_ctx.result_asm(0)
    assert e.value.status == -1


def test_gpu_result_asm_1m(gpu_ctx):
    t = W.program_of_size(0x5EED0007, 1_000_000)
    say("1M-node program to_asm", t.n)
    r = gpu_ctx.compile(t)
    got = gpu_ctx.result_asm(0)
    want = oracle.to_asm(r["words"])
    assert got == want, first_diff(got, want)
