"""The host mirror of the reference interface, ptilopsis_amd.generator.RvGeneratorHip,
on the GPU: process(profile=True) prints the reference's -p output on stdout
(src/codegen/rv_generator.cpp:150-162: a line of the seven stage names with the
last one repeated, then a line of the seven times in ns with the last one
repeated) and its stage table on stderr; get_instructions(), the function-table
members, to_binary and to_asm equal the oracle's output."""
import io

import numpy as np
import pytest

import oracle
from ptilopsis_amd import abi, workloads as W
from ptilopsis_amd.generator import RvGeneratorHip

pytestmark = pytest.mark.gpu


def test_gpu_rv_generator_hip_process_profile(gpu_ctx, capsys):
    tree = W.single_50k()
    gen = RvGeneratorHip(tree, gpu_ctx)
    gen.process(profile=True)
    out, err = capsys.readouterr()
    lines = out.splitlines()
    assert len(lines) == 2
    names, times = lines[0].split(","), lines[1].split(",")
    assert names == list(abi.STAGES) + [abi.STAGES[-1]]
    assert len(times) == 8 and times[-1] == times[-2] and all(t.isdigit() for t in times)
    assert sum(int(t) for t in times[:7]) > 0
    assert err.splitlines()[0].startswith("Stage") and "Total" in err
    want = oracle.compile_tree(tree)
    assert np.array_equal(gen.get_instructions(), want["words"])
    assert np.array_equal(gen.function_ids, want["func_ids"])
    assert np.array_equal(gen.func_starts, want["func_starts"])
    assert np.array_equal(gen.func_ends, want["func_ends"])
    assert np.array_equal(gen.function_sizes, want["func_sizes"])
    b = io.BytesIO()
    gen.to_binary(b)
    assert b.getvalue() == want["words"].astype("<u4").tobytes()
    s = io.StringIO()
    gen.to_asm(s)
    assert s.getvalue().encode() == oracle.to_asm(want["words"])


def test_gpu_rv_generator_hip_without_profile(gpu_ctx, capsys):
    tree = W.program_of_size(0x5EED0C01, 5000)
    gen = RvGeneratorHip(tree, gpu_ctx)
    gen.process(profile=False)
    out, err = capsys.readouterr()
    assert out == "" and "Total" in err
    assert np.array_equal(gen.get_instructions(), oracle.compile_tree(tree)["words"])
