"""The adaptive-partitioning kernel decides the reference's chi-square split test
(AdaptivePartitioningMI.hpp:608-619, double arithmetic, threshold 7.815) in integers: with
D = sum (4 n_i - n)^2 it splits when D > floor(n (2M + 1) / 2^49), M = 7.815 * 2^50.  This checks
the rule against the reference's double evaluation for every window size n up to 8191 around each
threshold, and the kernel's two-word evaluation of the floor against exact integer arithmetic."""
from fractions import Fraction

import numpy as np

C = 7.815


def reference_splits(counts):
    n = sum(counts)
    e = n / 4.0
    chi = 0.0
    for c in counts:
        d = c - e
        chi += d * d
    return chi / e > C


def test_threshold_constant_is_exact():
    M = Fraction(C) * 2 ** 50
    assert M.denominator == 1 and M.numerator.bit_length() == 53


def test_integer_rule_matches_double_rule():
    M = int(Fraction(C) * 2 ** 50)
    K = 2 * M + 1
    A, B = K >> 32, K & 0xFFFFFFFF
    for n in range(1, 8192):
        t = (n * K) >> 49
        assert t == ((n * A) + ((n * B) >> 32)) >> 17          # the kernel's evaluation
        assert (n * K) % (1 << 49) != 0                        # the midpoint is never hit exactly
        for D in range(max(0, t - 2), t + 3):
            # D / (4 n) is the reference's chi exactly; compare its rounded value with C
            assert ((D / 16.0) / (n / 4.0) > C) == (D > t), (n, D)


def test_rule_on_quadrant_counts():
    rng = np.random.default_rng(3)
    M = int(Fraction(C) * 2 ** 50)
    K = 2 * M + 1
    for _ in range(20000):
        n = int(rng.integers(3, 8192))
        cuts = np.sort(rng.integers(0, n + 1, 3))
        q = [int(cuts[0]), int(cuts[1] - cuts[0]), int(cuts[2] - cuts[1]), int(n - cuts[2])]
        D = sum((4 * c - n) ** 2 for c in q)
        assert D < 2 ** 31
        assert reference_splits(q) == (D > (n * K) >> 49)
