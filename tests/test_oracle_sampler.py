"""Pins the oracle's samplers (the checker for the HIP sampler).

Known answers and exact arithmetic, not a retyping of the formula:
  * prime table: count 1000, last prime 7919, prefix sums whose 999th value is
    3 674 994 (the last entry of m_prime_sums, src/RandomSampler.cuh:198) and
    whose total is 3 682 913;
  * MT19937: the C++ standard's known answer (10000th output of seed 5489);
  * permutation table: every segment is a permutation of 0..p-1;
  * radical inverse: base-2/3 van der Corput values, and the permuted
    inverse against its exact rational value (Fraction), which a dropped
    postfix term, reversed digit order or wrong permutation offset fails;
  * enumerator: brute force over tiny images -- every (pixel, sample) index
    maps back into its own pixel under ScaleX/ScaleY, and indices are distinct.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle


def test_prime_table_known_answers():
    L = oracle.lib()
    assert L.or_prime(0) == 2 and L.or_prime(1) == 3 and L.or_prime(999) == 7919
    assert L.or_prime_sum(999) == 3674994
    assert L.or_prime_sum(1000) == 3682913
    assert L.or_halton_table_size() == 3682913


def test_mt19937_known_answer():
    out = oracle.mt19937(5489, 10000)
    assert int(out[-1]) == 4123659995


def test_permutation_table_is_permutation():
    L = oracle.lib()
    t = oracle.halton_table(0)
    for i in list(range(0, 60)) + [500, 999]:
        p = L.or_prime(i)
        off = L.or_prime_sum(i)
        seg = t[off:off + p]
        assert np.array_equal(np.sort(seg), np.arange(p)), i
    # a different seed gives a different table
    assert not np.array_equal(t[:1000], oracle.halton_table(1)[:1000])


@pytest.mark.parametrize("index,expect", [(1, 0.5), (2, 0.25), (3, 0.75), (4, 0.125), (5, 0.625)])
def test_van_der_corput_base2(index, expect):
    assert oracle.lib().or_radical_inverse2(index) == np.float32(expect)


@pytest.mark.parametrize("index,num,den", [(1, 1, 3), (2, 2, 3), (3, 1, 9), (4, 4, 9), (5, 7, 9)])
def test_van_der_corput_base3(index, num, den):
    got = oracle.lib().or_radical_inverse3(index)
    assert abs(got - num / den) <= 2e-7


def test_radical_inverse_clamped():
    assert oracle.lib().or_radical_inverse2(0) == np.float32(0.000001)


def exact_permuted(index, p, perm):
    digits = []
    while index:
        digits.append(index % p)
        index //= p
    k = len(digits)
    s = 0
    for d in digits:
        s = s * p + int(perm[d])
    return Fraction(s, p ** k) + Fraction(int(perm[0]), p ** k * (p - 1))


def test_permuted_radical_inverse_exact():
    L = oracle.lib()
    t = oracle.halton_table(7)
    rng = np.random.RandomState(0)
    for dim in [2, 3, 4, 10, 50, 200, 999]:
        p = L.or_prime(dim)
        perm = t[L.or_prime_sum(dim):L.or_prime_sum(dim) + p]
        for index in list(rng.randint(1, 2 ** 31, size=40)) + [1, p - 1, p, p * p + 3, 2 ** 40 + 12345]:
            got = oracle.halton_value(int(index), dim, t)
            ex = float(exact_permuted(int(index), p, perm))
            ex = min(max(ex, 1e-6), 0.999999)
            # base_multiply accumulates one rounding per digit
            assert abs(got - ex) <= 1e-6 * max(ex, 1e-3), (dim, index, got, ex)


@pytest.mark.parametrize("w,h", [(5, 4), (16, 9), (33, 10)])
def test_enumerator_brute_force(w, h):
    L = oracle.lib()
    seen = set()
    for py in range(h):
        for px in range(w):
            for s in range(3):
                idx, e = oracle.enum_index(w, h, px, py, s)
                assert idx not in seen
                seen.add(idx)
                x = L.or_radical_inverse2(idx) * e.x_scale
                y = L.or_radical_inverse3(idx) * e.y_scale
                assert px <= x < px + 1 or (idx == 0 and px == 0)
                assert py - 1e-4 <= y < py + 1 + 1e-4
    # each sample period holds increment = 2^e2 * 3^e3 distinct indices
    _, e = oracle.enum_index(w, h, 0, 0, 0)
    assert e.increment >= w * h
