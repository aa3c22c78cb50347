"""mt19937_64 (Matsumoto & Nishimura 2004 parameters, as std::mt19937_64) in numpy.

Test helper for the scan known-answer test, whose inputs the reference draws
with std::mt19937_64 and std::uniform_int_distribution<uint32_t>{0, 4095}
(benchmarks/prefix_sum.cpp:35-40, 215-219).  For a 2^64-range engine and a
power-of-two range libstdc++'s distribution returns (x * 4096) >> 64, the top
12 bits of each draw.
"""
import numpy as np

N, M = 312, 156
A = np.uint64(0xB5026F5AA96619E9)
UPPER = np.uint64(0xFFFFFFFF80000000)
LOWER = np.uint64(0x7FFFFFFF)


def _init(seed: int) -> np.ndarray:
    mt = np.zeros(N, np.uint64)
    mt[0] = seed
    f = 6364136223846793005
    prev = seed
    for i in range(1, N):
        prev = (f * (prev ^ (prev >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        mt[i] = prev
    return mt


def _twist(mt: np.ndarray) -> None:
    def step(lo, hi, nxt, far):
        x = (mt[lo:hi] & UPPER) | (nxt & LOWER)
        xa = x >> np.uint64(1)
        xa ^= np.where((x & np.uint64(1)) != 0, A, np.uint64(0))
        mt[lo:hi] = far ^ xa
    step(0, N - M, mt[1:N - M + 1].copy(), mt[M:N].copy())
    step(N - M, N - 1, mt[N - M + 1:N].copy(), mt[0:M - 1].copy())
    step(N - 1, N, mt[0:1].copy(), mt[M - 1:M].copy())


def _temper(y: np.ndarray) -> np.ndarray:
    y = y ^ ((y >> np.uint64(29)) & np.uint64(0x5555555555555555))
    y = y ^ ((y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000))
    y = y ^ ((y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000))
    return y ^ (y >> np.uint64(43))


def draws(seed: int, count: int) -> np.ndarray:
    mt = _init(seed)
    out = np.empty(count, np.uint64)
    k = 0
    while k < count:
        _twist(mt)
        take = min(N, count - k)
        out[k:k + take] = _temper(mt[:take])
        k += take
    return out


def uniform_0_4095(seed: int, count: int) -> np.ndarray:
    return (draws(seed, count) >> np.uint64(52)).astype(np.uint32)
