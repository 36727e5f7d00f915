"""Pin of the division the diffuse relaxation uses on the GPU.

The reference divides by the constant 1 + 4a with an IEEE float division
(src/fluid_solver_cpu.cpp:107-108).  The Gauss-Seidel kernels compute it as
Markstein's FMA correction (P. Markstein, "Computation of elementary functions
on the IBM RISC System/6000 processor", IBM J. R&D 34(1), 1990; Muller et al.,
Handbook of Floating-Point Arithmetic, FMA-based division):

    y  = RN(1/d)             (host, once per solve)
    q0 = RN(x * y)           (within ~1.5 ulp of x/d)
    q1 = RN(q0 + RN(x - q0 * d) * y)   (two FMAs; faithful, within 1 ulp)
    r  = x - q1 * d          (exact, one FMA: q1 is faithful)
    q2 = RN(q1 + r * y)      (one FMA)  ==  RN(x / d)

valid when no intermediate underflows or overflows; the kernel takes the IEEE
division for |x| outside [2^-100, 2^100], zero, inf and NaN.  Here every step
is evaluated in exact rational arithmetic and the result compared with the
correctly rounded quotient, for the denominators the configurations produce
and for random, near-tie and extreme dividends.
"""
from __future__ import annotations

import random
from fractions import Fraction

import numpy as np
import pytest


def rn32(q: Fraction) -> Fraction:
    """Round a rational to the nearest float32 (ties to even), subnormals included."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    # a in [2^e, 2^(e+1)); 24-bit significand -> quantum 2^(e-23), at least 2^-149
    quantum = Fraction(2) ** (max(e, -126) - 23)
    m = a / quantum
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * fl * quantum


def f32(x: float) -> Fraction:
    return Fraction(float(np.float32(x)))


def markstein(x: Fraction, d: Fraction, y: Fraction) -> Fraction:
    q0 = rn32(x * y)
    q1 = rn32(q0 + rn32(x - q0 * d) * y)
    r = x - q1 * d
    assert rn32(r) == r, "remainder not exactly representable"
    return rn32(q1 + r * y)


def denominators():
    """1 + 4a, a = (dt * float(R*R)) * rate, as the host computes them."""
    out = []
    for R, dt, rate in ((4098, 0.05, 1e-5), (4098, 0.05, 1e-6), (1026, 0.05, 1e-5), (1026, 0.05, 1e-6),
                        (8194, 0.05, 1e-6), (130, 0.1, 1e-3), (34, 0.1, 5e-2), (258, 0.1, 1e-3),
                        (67, 0.07, 2e-2), (1026, 0.1, 1.0)):
        a = np.float32(dt) * np.float32(R * R) * np.float32(rate)
        d = np.float32(1.0) + np.float32(4.0) * a
        out.append(d)
    return out


@pytest.mark.parametrize("d", denominators(), ids=lambda d: f"{float(d):.6g}")
def test_markstein_division_is_correctly_rounded(d):
    rng = random.Random(int(np.float32(d).view(np.uint32)))
    D = Fraction(float(d))
    y = Fraction(float(np.float32(1.0) / d))
    xs = []
    for _ in range(6000):  # random significands over many binades
        m = rng.randrange(1 << 23, 1 << 24)
        e = rng.randrange(-100, 100)
        xs.append(Fraction(m) * Fraction(2) ** (e - 23) * (1 if rng.random() < 0.5 else -1))
    for _ in range(1000):  # dividends whose quotient sits next to a rounding tie
        q = rn32(Fraction(rng.uniform(0.01, 100.0)))
        e = q.numerator.bit_length() - q.denominator.bit_length()
        half = Fraction(2) ** (e - 24)
        for t in (q + half, q - half):
            x = rn32(t * D) if abs(t * D) > 0 else Fraction(0)
            xs.append(x)
    for e in (-100, -60, 0, 60, 99):  # extremes of the fast-path range
        xs.append(Fraction(2) ** e)
        xs.append(Fraction((1 << 24) - 1) * Fraction(2) ** (e - 24))
    for x in xs:
        if x == 0:
            continue
        assert markstein(x, D, y) == rn32(x / D), (float(x), float(d))


def test_exact_quotient_and_unit_divisor():
    d = np.float32(1.0)
    for xv in (1.0, -3.5, 1e-20, 7.25e20):
        x = f32(xv)
        assert markstein(x, Fraction(1), Fraction(1)) == x
    d = np.float32(4.0)
    assert markstein(f32(10.0), Fraction(4), Fraction(1, 4)) == Fraction(5, 2)


def division_floor(d: np.float32) -> float:
    """Host rule for the smallest nonzero |dividend| of the FMA-corrected
    division (fluid-2d_amd/csrc/fs2d_internal.hpp division_floor)."""
    if not d <= np.float32(2.0 ** 100):
        return float("inf")
    return max(2.0 ** -100, float(np.float32(d) * np.float32(2.0 ** -124)))


@pytest.mark.parametrize("d", [np.float32(1.8e7), np.float32(2.0 ** 26 * 1.5), np.float32(2.0 ** 40 + 12345 * 2.0 ** 17),
                               np.float32(3.3e20), np.float32(2.0 ** 99 * 1.75)], ids=lambda d: f"{float(d):.3g}")
def test_markstein_exact_from_the_division_floor_for_large_denominators(d):
    """Large 1 + 4a (huge diffusion rates): from lo = max(2^-100, d 2^-124)
    upward the quotient stays normal and the correction exact."""
    D = Fraction(float(d))
    y = Fraction(float(np.float32(1.0) / d))
    lo = division_floor(d)
    rng = random.Random(int(d.view(np.uint32)))
    e_lo = int(np.floor(np.log2(lo)))
    for e in list(range(e_lo, e_lo + 6)) + [rng.randrange(e_lo, 100) for _ in range(300)]:
        for _ in range(20):
            m = rng.randrange(1 << 23, 1 << 24)
            x = Fraction(m) * Fraction(2) ** (e - 23)
            if x < Fraction(lo) or x > Fraction(2) ** 100:
                continue
            assert markstein(x, D, y) == rn32(x / D), (float(x), float(d))


def test_division_floor_rule():
    assert division_floor(np.float32(1.0)) == 2.0 ** -100
    assert division_floor(np.float32(2.0 ** 24)) == 2.0 ** -100
    assert division_floor(np.float32(2.0 ** 30)) == 2.0 ** -94
    assert division_floor(np.float32(2.0 ** 101)) == float("inf")
    assert division_floor(np.float32(np.nan)) == float("inf")
