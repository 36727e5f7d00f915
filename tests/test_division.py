"""Pin of the division the diffuse relaxation uses on the GPU.

The reference divides by the constant d = 1 + 4a with an IEEE float division
(src/fluid_solver_cpu.cpp:107-108).  The kernels compute it from the host's
y = RN(1/d) and y_lo = RN(1/d - y) (so y + y_lo is 1/d to about 2^-48):

    q1 = RN(x y + RN(x y_lo))   (one multiply, one FMA: within 2^-47 of x/d
                                 before rounding, so faithful)
    r  = x - q1 d               (exact, one FMA: q1 is faithful)
    q2 = RN(q1 + r y)           (one FMA; Markstein, IBM J. R&D 34(1), 1990)
                                 ==  RN(x / d)

valid when no intermediate underflows or overflows; dividends outside
[lo, 2^100] (zero, tiny, huge, inf, NaN) take the IEEE division.  For most
divisors q1 alone is already RN(x / d) for every dividend (Brisebarre, Muller
and Raina, IEEE Trans. Computers 53(8), 2004); the host decides that per
divisor by trying every significand of one binade, and the Gauss-Seidel unit
kernels then skip the correction.  Here every step is evaluated in exact
rational arithmetic against the correctly rounded quotient, and the host's
per-divisor decision is checked against an independent numpy enumeration.
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


def inv_lo(d: np.float32, y: np.float32) -> Fraction:
    """RN(1/d - y) as the host computes it: (1 - d y) / d in double (1 - d y
    exact), rounded to float."""
    return Fraction(float(np.float32((1.0 - float(d) * float(y)) / float(d))))


def two_product(x: Fraction, y: Fraction, yl: Fraction) -> Fraction:
    return rn32(x * y + rn32(x * yl))


def markstein(x: Fraction, d: Fraction, y: Fraction, yl: Fraction) -> Fraction:
    q1 = two_product(x, y, yl)
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
    yf = np.float32(1.0) / d
    y, yl = Fraction(float(yf)), inv_lo(d, yf)
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
        assert markstein(x, D, y, yl) == rn32(x / D), (float(x), float(d))


def test_exact_quotient_and_unit_divisor():
    d = np.float32(1.0)
    for xv in (1.0, -3.5, 1e-20, 7.25e20):
        x = f32(xv)
        assert markstein(x, Fraction(1), Fraction(1), Fraction(0)) == x
        assert two_product(x, Fraction(1), Fraction(0)) == x
    assert markstein(f32(10.0), Fraction(4), Fraction(1, 4), Fraction(0)) == Fraction(5, 2)


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
    yf = np.float32(1.0) / d
    y, yl = Fraction(float(yf)), inv_lo(d, yf)
    lo = division_floor(d)
    rng = random.Random(int(d.view(np.uint32)))
    e_lo = int(np.floor(np.log2(lo)))
    for e in list(range(e_lo, e_lo + 6)) + [rng.randrange(e_lo, 100) for _ in range(300)]:
        for _ in range(20):
            m = rng.randrange(1 << 23, 1 << 24)
            x = Fraction(m) * Fraction(2) ** (e - 23)
            if x < Fraction(lo) or x > Fraction(2) ** 100:
                continue
            assert markstein(x, D, y, yl) == rn32(x / D), (float(x), float(d))


def test_division_floor_rule():
    assert division_floor(np.float32(1.0)) == 2.0 ** -100
    assert division_floor(np.float32(2.0 ** 24)) == 2.0 ** -100
    assert division_floor(np.float32(2.0 ** 30)) == 2.0 ** -94
    assert division_floor(np.float32(2.0 ** 101)) == float("inf")
    assert division_floor(np.float32(np.nan)) == float("inf")


# ---- the two-product quotient and the host's per-divisor check ----------

def binade_misses(d: np.float32, y: np.float32, yl: np.float32, E: int = 0) -> np.ndarray:
    """Dividends x = 2^E m, m over every significand of [1, 2), at which
    RN(x y + RN(x yl)) differs from RN(x / d): numpy float32 for the products
    and the quotient, the FMA from an exact double product, an exact two-sum
    and a tie correction (the double sum is off the exact one only by its own
    rounding error, which decides a float32 rounding only when the double sum
    sits exactly on a float32 midpoint)."""
    m = np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x3F800000)
    x = (m.view(np.float32) * np.float32(2.0 ** E)).astype(np.float32)
    t = (x * yl).astype(np.float32)
    p = x.astype(np.float64) * np.float64(y)  # exact: 24 x 24 bits
    tt = t.astype(np.float64)
    s_ = p + tt
    bp = s_ - tt
    err = (p - bp) + (tt - (s_ - bp))  # s_ + err == p + tt exactly
    q = s_.astype(np.float32)
    lo_n = np.nextafter(q, np.float32(-np.inf)).astype(np.float64)
    hi_n = np.nextafter(q, np.float32(np.inf)).astype(np.float64)
    qd = q.astype(np.float64)
    # s_ on a midpoint between q and a neighbour: the exact sum's side decides
    up = (s_ == (qd + hi_n) / 2) & (err > 0)
    dn = (s_ == (qd + lo_n) / 2) & (err < 0)
    q = np.where(up, np.nextafter(q, np.float32(np.inf)), q)
    q = np.where(dn, np.nextafter(q, np.float32(-np.inf)), q)
    want = (x / d).astype(np.float32)
    return x[q != want]


# Divisors the configurations produce, and ones the two-product quotient
# misses (a single significand each, found by the enumeration below).
TWO_PRODUCT_CASES = [(float(d), True) for d in denominators()[:5]] + [
    (float.fromhex("0x1.a40e5ep+9"), False),  # 840.11
    (float.fromhex("0x1.de71cep+0"), False),  # 1.87
    (3.832235097885132, False),  # 1 + 4 (0.1 67^2 rate), rate ~ 1.577e-3
    (12.917113304138184, False),  # R = 130, rate ~ 1.763e-3
]


def two_product_pair(d: np.float32):
    """The pair of the two-product quotient as the host forms it: y2 = RN(1/d),
    or the float below it when that is above 1/d, and y2_lo = RN((1 - d y2) / d)
    (1 - d y2 exact in double), positive."""
    y = np.float32(1.0) / d
    if 1.0 - float(d) * float(y) < 0.0:
        y = np.nextafter(y, np.float32(0.0))
    return y, np.float32((1.0 - float(d) * float(y)) / float(d))


@pytest.mark.parametrize("d,exact", TWO_PRODUCT_CASES, ids=lambda v: f"{v}")
def test_host_two_product_check_agrees_with_enumeration(d, exact):
    import fs2d_amd
    d32 = np.float32(d)
    plan = fs2d_amd.division_plan(float(d32))
    y = np.float32(1.0) / d32
    assert np.float32(plan["inv"]) == y
    assert np.float32(plan["inv_lo"]) == np.float32(float(inv_lo(d32, y)))
    y2, y2l = two_product_pair(d32)
    misses = binade_misses(d32, y2, y2l)
    assert (len(misses) == 0) == exact
    assert plan["checked"] == exact
    assert plan["lo"] >= division_floor(d32)
    if plan["checked"]:
        assert np.float32(plan["two_hi"]) == y2 and np.float32(plan["two_lo"]) == y2l and y2l > 0
        assert Fraction(float(y2)) <= 1 / Fraction(float(d32))
        # from lo up RN(x y2_lo) is normal, so every binade scales the checked one
        assert float(np.float32(plan["lo"]) * y2l) >= 2.0 ** -126
        # an infinite dividend: RN(inf y2 + RN(inf y2_lo)) = inf, as the IEEE division
        with np.errstate(over="ignore", invalid="ignore"):
            inf = np.float32(np.inf)
            assert np.float32(inf * y2 + inf * y2l) == inf and np.float32(-inf * y2 + -inf * y2l) == -inf
    for x in misses[:4]:
        # where the two-product quotient misses, the Markstein step still rounds correctly
        X, D = Fraction(float(x)), Fraction(float(d32))
        Y, YL = Fraction(float(y)), Fraction(float(np.float32(plan["inv_lo"])))
        assert two_product(X, Fraction(float(y2)), Fraction(float(y2l))) != rn32(X / D)
        assert markstein(X, D, Y, YL) == rn32(X / D)


def test_enumeration_is_exact_on_known_cases():
    """The numpy enumeration against exact rational arithmetic on sampled
    significands, so that its two-sum and tie handling are pinned too."""
    rng = random.Random(7)
    for d in (np.float32(4.358721), np.float32(float.fromhex("0x1.a40e5ep+9")), np.float32(3.0)):
        y, yl = two_product_pair(d)
        misses = {float(v) for v in binade_misses(d, y, yl)}
        D, Y, YL = Fraction(float(d)), Fraction(float(y)), Fraction(float(yl))
        ms = [rng.randrange(1 << 23) for _ in range(1500)] + [int(np.float32(v).view(np.uint32)) & 0x7FFFFF for v in misses]
        for m in ms:
            x = float(np.uint32(0x3F800000 | m).view(np.float32))
            X = Fraction(x)
            assert (two_product(X, Y, YL) != rn32(X / D)) == (x in misses), (float(d), x)


def test_two_product_check_needs_a_positive_tail():
    """A divisor whose reciprocal is a float (a power of two, d = 1 when a = 0)
    has no positive tail: the kernels keep the corrected quotient, exact there
    too, and the plan leaves the division floor as it is."""
    import fs2d_amd
    for d in (1.0, 2.0, 1024.0):
        plan = fs2d_amd.division_plan(d)
        assert plan["inv_lo"] == 0.0 and not plan["checked"] and plan["lo"] == division_floor(np.float32(d))
