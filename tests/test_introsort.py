"""The oracle's restatement of libstdc++ std::sort (the isn_gen depth sort,
rv_generator_avx.cpp:539-541) against the host's real std::sort, on key
patterns with many ties: the order of equal keys must match exactly."""
import subprocess

import numpy as np
import pytest

import oracle
from cpp_helpers import compile_cpp
from ptilopsis_amd import workloads as W, parse_source

SORT_SRC = r'''
#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <vector>
int main(int argc, char **argv) {
    FILE *f = fopen(argv[1], "rb");
    uint64_t n; fread(&n, 8, 1, f);
    std::vector<int32_t> key(n); fread(key.data(), 4, n, f); fclose(f);
    std::vector<uint32_t> idx(n);
    for (uint64_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    f = fopen(argv[2], "wb"); fwrite(idx.data(), 4, n, f); fclose(f);
}'''


def libstdcxx_sort(keys, tmp_path):
    exe = compile_cpp("sort", SORT_SRC)
    inp, out = tmp_path / "k.bin", tmp_path / "p.bin"
    with open(inp, "wb") as f:
        f.write(np.uint64(len(keys)).tobytes())
        f.write(np.ascontiguousarray(keys, np.int32).tobytes())
    subprocess.run([exe, str(inp), str(out)], check=True)
    return np.fromfile(out, np.uint32)


def key_patterns():
    rng = np.random.default_rng(7)
    yield "tiny", np.array([3, 1, 2, 1, 3], np.int32)
    for n in [16, 17, 33, 100, 1000, 5000]:
        yield f"rand{n}", rng.integers(0, 8, n).astype(np.int32)
    yield "equal", np.zeros(4000, np.int32)
    yield "sorted", np.sort(rng.integers(0, 50, 3000)).astype(np.int32)
    yield "reverse", np.sort(rng.integers(0, 50, 3000))[::-1].astype(np.int32).copy()
    yield "organ", np.concatenate([np.arange(1500), np.arange(1500)[::-1]]).astype(np.int32)
    src, _ = W.program_source(4242, 12)
    yield "tree_depths", parse_source(src).depth.astype(np.int32)


@pytest.mark.parametrize("name,keys", list(key_patterns()), ids=[p[0] for p in key_patterns()])
def test_introsort_matches_libstdcxx(name, keys, tmp_path):
    want = libstdcxx_sort(keys, tmp_path)
    got = oracle.introsort(keys)
    assert np.array_equal(got, want)


def test_introsort_large_tree(tmp_path):
    t = W.single_50k()
    assert np.array_equal(oracle.introsort(t.depth), libstdcxx_sort(t.depth, tmp_path))
