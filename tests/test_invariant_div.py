"""Division by an invariant integer as used by the linear-tile step kernels
(InvariantDiv in getting-air_amd/csrc/dense.h): m = ceil(2^(31+l) / d) with
l = ceil(log2 d), floor(n / d) = (n m) >> (31 + l) for 0 <= n < 2^31.

Checked against integer division: the Granlund-Montgomery condition
2^(31+l) <= m d <= 2^(31+l) + 2^l for every d up to 4096 and the extremes,
m < 2^32 (so n m fits 64 bits), and exact quotients at and around every
multiple of d near the top of the range and at random n.
"""
import numpy as np


def make(d):
    l = 0
    while (1 << l) < d:
        l += 1
    return ((1 << (31 + l)) + d - 1) // d, 31 + l


def test_condition_and_range():
    for d in list(range(1, 4097)) + [(1 << 20) + 3, (1 << 30) - 1, 1 << 30, (1 << 30) + 1, (1 << 31) - 1]:
        m, s = make(d)
        l = s - 31
        assert (1 << s) <= m * d <= (1 << s) + (1 << l), d
        assert m < (1 << 32), d


def test_quotients():
    rng = np.random.default_rng(3)
    top = (1 << 31) - 1
    for d in list(range(1, 1025)) + [1536, 2048, 2500, 3072, 4096, 65536 + 7]:
        m, s = make(d)
        k = top // d
        ns = [0, 1, d - 1, d, d + 1, top, top - 1]
        ns += [q * d + r for q in (k - 1, k) for r in (0, 1, d - 1) if q * d + r <= top]
        ns += [int(v) for v in rng.integers(0, top, 64)]
        for n in ns:
            assert (n * m) >> s == n // d, (n, d)
