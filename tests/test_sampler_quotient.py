"""The quotient rule of the GPU radical inverse (dev_sampler.hpp).

For an index n < 2^32 and a Halton base p (the primes from 5 to 7919, Halton
dimensions 2..999) the digit loop takes floor(n / p) as
trunc(fma(n, RN(1/p), 2^-20)) in double precision.  The fma is evaluated here
exactly with rationals and rounded once to double, as the hardware does, and
the truncation is compared with integer division: at the values where a
slip would show (multiples of p, one below and one above them, 0 and the top
of the 32-bit range) for every base.
"""
import random
from fractions import Fraction

import oracle


def _fma_quotient(n, inv):
    return int(float(Fraction(n) * Fraction(inv) + Fraction(1, 1 << 20)))


def test_double_quotient_is_exact_for_32_bit_indices():
    L = oracle.lib()
    primes = [L.or_prime(i) for i in range(1000)]
    assert primes[2] == 5 and primes[-1] == 7919
    rng = random.Random(5)
    top = (1 << 32) - 1
    checked = 0
    for p in primes[2:]:
        inv = 1.0 / p
        qmax = top // p
        ns = {0, 1, p - 1, p, p + 1, top, top - 1, qmax * p, qmax * p - 1}
        for _ in range(12):
            k = rng.randint(1, qmax)
            ns.update((k * p - 1, k * p, k * p + 1))
        for n in ns:
            if 0 <= n <= top:
                assert _fma_quotient(n, inv) == n // p, (p, n)
                checked += 1
    assert checked > 30000


def test_quotient_offset_is_needed():
    """A multiple of 103 whose plain product n * RN(1/103) rounds below
    n / 103, so truncation alone would take the quotient one too small; the
    2^-20 offset lifts it."""
    n, p = 3010672387, 103
    assert n % p == 0
    assert int(float(Fraction(n) * Fraction(1.0 / p))) == n // p - 1
    assert _fma_quotient(n, 1.0 / p) == n // p
