"""The algorithm behind the kernels' division shortcuts, on the CPU, in exact
arithmetic (no GPU): with y = RN(1/b), q = RN(a y), r = RN(a - b q) and
q' = RN(q + r y) (both fused, one rounding each), q' is the correctly rounded
binary32 quotient RN(a/b).  The GPU check
(test_gpu_exact_division.py) covers the hardware reciprocal and the scaling
over exponents; this test pins the remainder step itself, independent of the
hardware, on random significand pairs and on quotients chosen to lie within
a few units in the last place of a rounding midpoint (where a dropped
correction or a wrong rounding would show)."""
from __future__ import annotations

import random
from fractions import Fraction

import numpy as np

P = 24  # binary32 significand bits


def rn(x: Fraction) -> Fraction:
    """Round a positive rational to the nearest binary32 value (ties to even),
    normal range only."""
    assert x > 0
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    # x in [2^e, 2^(e+1)); ulp = 2^(e - P + 1)
    scale = Fraction(2) ** (P - 1 - e)
    m = x * scale
    lo = m.numerator // m.denominator
    rem = m - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return Fraction(lo) / scale


def f32(x: Fraction) -> float:
    return float(np.float32(float(x)))


def rn_signed(x: Fraction) -> Fraction:
    if x == 0:
        return Fraction(0)
    return rn(x) if x > 0 else -rn(-x)


def quotient_by_remainder_step(a: Fraction, b: Fraction) -> tuple[Fraction, Fraction]:
    """The shortcut's operations, each rounded once as the hardware's multiply
    and fused multiply-adds do.  The remainder a - b q needs up to 25 bits,
    so the fused multiply-add can round it too."""
    y = rn(1 / b)
    q = rn(a * y)
    r = rn_signed(a - b * q)
    q1 = rn_signed(q + r * y)
    return q1, r


def sig(k: int) -> Fraction:
    """Significand in [1, 2) from a 23-bit integer."""
    return 1 + Fraction(k, 1 << 23)


def test_remainder_step_rounds_random_quotients_correctly():
    rng = random.Random(20261204)
    for _ in range(3000):
        a, b = sig(rng.getrandbits(23)), sig(rng.getrandbits(23))
        q1, _ = quotient_by_remainder_step(a, b)
        assert q1 == rn(a / b), (a, b)


def test_remainder_step_rounds_near_midpoint_quotients_correctly():
    """b random; a chosen as RN(b m) for a rounding midpoint m of the quotient's
    binade, so a/b lies within an ulp or two of the midpoint."""
    rng = random.Random(7)
    checked = 0
    for _ in range(3000):
        b = sig(rng.getrandbits(23))
        # midpoint m between two binary32 values in [1, 2)
        k = rng.getrandbits(23)
        m = 1 + Fraction(2 * k + 1, 1 << 24)
        a = rn(b * m)
        if not (1 <= a < 2):
            continue
        q1, _ = quotient_by_remainder_step(a, b)
        assert q1 == rn(a / b), (a, b)
        checked += 1
    assert checked > 1000


def test_a_plausible_mistake_fails():
    """Dropping the remainder step (q = RN(a y) alone) is wrong for some pairs:
    the check above would see it."""
    rng = random.Random(3)
    wrong = 0
    for _ in range(2000):
        a, b = sig(rng.getrandbits(23)), sig(rng.getrandbits(23))
        y = rn(1 / b)
        if rn(a * y) != rn(a / b):
            wrong += 1
    assert wrong > 0


def test_rounding_helper_agrees_with_numpy():
    rng = random.Random(11)
    for _ in range(500):
        x = Fraction(rng.getrandbits(60) + 1, rng.getrandbits(40) + 1)
        assert float(rn(x)) == f32(x)
