"""Known-answer tests: the reference's only one (the scan primitive,
benchmarks/prefix_sum.cpp:35-40,105-124,215-226) and the GCC-specific ffz."""
import ctypes as C
import subprocess

import numpy as np

import oracle
from cpp_helpers import compile_cpp
import mt19937_64 as MT

SEED = 0xC0FFEEA4DBEEF
N = 1024 * 1024


def test_mt19937_64_and_distribution_match_libstdcxx():
    exe = compile_cpp("mt", r'''
#include <random>
#include <cstdio>
#include <cstdint>
int main() {
    std::mt19937_64 a{0xC0FFEEA4DBEEFull};
    for (int i = 0; i < 700; ++i) printf("%llu\n", (unsigned long long)a());
    std::mt19937_64 b{0xC0FFEEA4DBEEFull};
    std::uniform_int_distribution<uint32_t> g{0, 4095};
    for (int i = 0; i < 700; ++i) printf("%u\n", g(b));
}''')
    out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert np.array_equal(MT.draws(SEED, 700), np.array(out[:700], np.uint64))
    assert np.array_equal(MT.uniform_0_4095(SEED, 700), np.array(out[700:], np.uint32))


def test_scan_kat_oracle():
    x = MT.uniform_0_4095(SEED, N)
    assert x.max() <= 4095
    ref = np.cumsum(x.astype(np.uint64))  # std::inclusive_scan
    ex = oracle.exclusive_scan(x)
    assert ex[0] == 0
    assert np.array_equal(ex[1:].astype(np.uint64), ref[:-1])
    assert int(ex[-1]) + int(x[-1]) == int(ref[-1])


def test_scan_edge_sizes():
    rng = np.random.default_rng(1)
    for n in [1, 7, 8, 9, 15, 16, 17, 63, 64, 65, 2047, 2048, 2049]:
        x = rng.integers(0, 1 << 20, n).astype(np.uint32)
        ex = oracle.exclusive_scan(x)
        ref = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint32)
        assert np.array_equal(ex, ref), n


def test_ffz_truncates_like_gcc():
    # include/utils.hpp:25-45: __builtin_ffs takes an int, so GCC looks at the low 32 bits only
    exe = compile_cpp("ffz", r'''
#include <cstdio>
#include <cstdint>
static uint32_t ffz(uint64_t val) { int res = __builtin_ffs(~val); return res == 0 ? 64 : res - 1; }
int main() {
    uint64_t v;
    while (scanf("%llu", (unsigned long long*)&v) == 1) printf("%u\n", ffz(v));
}''', extra=("-w",))
    vals = [0x1F, 0xFFFFFFFF, 0xFFFFFFFF0000001F, 0x00000000FFFFFFFF, 0xFFFFFFFFFFFFFFFF, 0x7FFFFFFF, 0]
    out = subprocess.run([exe], input="\n".join(str(v) for v in vals), capture_output=True, text=True, check=True)
    got = [int(x) for x in out.stdout.split()]
    assert got == [5, 64, 5, 64, 64, 31, 0]
