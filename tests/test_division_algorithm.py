"""The algorithm behind the kernels' division shortcuts, on the CPU, in exact
arithmetic (no GPU): with y = RN(1/b), q = RN(a y), r = RN(a - b q) and
q' = RN(q + r y) (both fused, one rounding each), q' is the correctly rounded
binary32 quotient RN(a/b).  The GPU check
(test_gpu_exact_division.py) covers the hardware reciprocal and the scaling
over exponents; this test pins the remainder step itself, independent of the
hardware, on random significand pairs and on quotients chosen to lie within
a few units in the last place of a rounding midpoint (where a dropped
correction or a wrong rounding would show).

The shortened square root (sqrt_rn) keeps the hardware root s and its two
neighbour tests: with d and u the binary32 values just below and above s, it
returns d if x - d s <= 0, u if x - u s > 0, else s (each residual one fused
multiply-add, whose sign is the exact residual's).  The tests below take s one
step below, at, or one step above the correctly rounded root and check that
the two tests land on RN(sqrt(x)), also at powers of four, where the value
below s is only half a step away."""
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


def next_up(v: np.float32) -> np.float32:
    return np.nextafter(v, np.float32(np.inf), dtype=np.float32)


def next_down(v: np.float32) -> np.float32:
    return np.nextafter(v, np.float32(0), dtype=np.float32)


def rn_sqrt(x: np.float32) -> np.float32:
    """RN(sqrt(x)) for a positive binary32 x, decided on exact squares."""
    X = Fraction(float(x))
    a = np.float32(np.sqrt(np.float64(x)))
    while Fraction(float(a)) ** 2 > X:
        a = next_down(a)
    while Fraction(float(next_up(a))) ** 2 <= X:
        a = next_up(a)
    b = next_up(a)
    m = (Fraction(float(a)) + Fraction(float(b))) / 2
    assert m * m != X  # no binary32 square root lies on a midpoint
    return a if X < m * m else b


def sqrt_by_neighbour_tests(x: np.float32, s: np.float32) -> np.float32:
    X, S = Fraction(float(x)), Fraction(float(s))
    d, u = next_down(s), next_up(s)
    r = d if X - Fraction(float(d)) * S <= 0 else s
    if X - Fraction(float(u)) * S > 0:
        r = u
    return r


def random_f32(rng: random.Random, lo_exp: int, hi_exp: int) -> np.float32:
    e = rng.randint(lo_exp, hi_exp)
    return np.float32(float(sig(rng.getrandbits(23))) * 2.0 ** e)


def test_neighbour_tests_round_the_square_root_correctly():
    rng = random.Random(5)
    xs = [random_f32(rng, -96, 100) for _ in range(1500)]
    # powers of four and their neighbours: the root is a power of two
    for e in range(-48, 50, 7):
        p = np.float32(4.0 ** e)
        xs += [p, next_up(p), next_down(p)]
    for x in xs:
        r = rn_sqrt(x)
        for s in (next_down(r), r, next_up(r)):
            assert sqrt_by_neighbour_tests(x, s) == r, (x, s, r)


def test_residual_magnitude_keeps_its_sign_above_the_domain_floor():
    """Each residual x - v s is a multiple of ulp(s)^2, at least 2^-142 for
    x >= 2^-96: above binary32's smallest subnormal, so the fused
    multiply-add's rounded residual has the exact residual's sign."""
    x = np.float32(2.0 ** -96)
    s = rn_sqrt(x)
    ulp = Fraction(float(next_up(s))) - Fraction(float(s))
    assert ulp * ulp >= Fraction(1, 2 ** 142) > Fraction(1, 2 ** 149)


def test_a_skipped_neighbour_test_fails():
    """Without the upward test, a hardware root one step low stays wrong."""
    rng = random.Random(9)
    wrong = 0
    for _ in range(300):
        x = random_f32(rng, -20, 20)
        r = rn_sqrt(x)
        s = next_down(r)
        X, S = Fraction(float(x)), Fraction(float(s))
        d = next_down(s)
        only_down = d if X - Fraction(float(d)) * S <= 0 else s
        wrong += only_down != r
    assert wrong > 0
