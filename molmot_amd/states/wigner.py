"""Wigner 3j, 6j and 9j symbols.

Every angular momentum is handled as twice its value (an int), so half-integer
arithmetic stays exact; the Racah sums are accumulated as integers over a
common denominator and only the final square root is taken in floating point
(the result is sign * sqrt(p / q) with p / q the exact square of the symbol).
The public functions take ordinary numbers (int, float, Fraction; 0.5, 1.5 ...)
like the reference's ``molmot/wigner.py`` (wigner3j / wigner6j / wigner9j);
arguments that violate a selection rule give 0.0.

Formulas: Racah's single-sum forms of the 3j and 6j symbols (Edmonds,
Angular Momentum in Quantum Mechanics, eqs. 3.7.3 and 6.3.7) and the 9j as a
sum over products of three 6j symbols (Edmonds eq. 6.4.3).
"""
from __future__ import annotations

import functools
import math
from fractions import Fraction

__all__ = ["wigner3j", "wigner6j", "wigner9j", "two", "fac2"]


def two(x) -> int:
    """Twice an angular momentum as an int; raises for non-half-integers."""
    if isinstance(x, int):
        return 2 * x
    t = 2 * Fraction(x).limit_denominator(1000)
    if t.denominator != 1:
        raise ValueError(f"{x} is not a multiple of 1/2")
    return int(t)


def fac2(t: int) -> int:
    """(t/2)! for an even t >= 0."""
    return math.factorial(t >> 1)


def _tri_ok(a: int, b: int, c: int) -> bool:
    """Triangle rule for doubled momenta (a, b, c), perimeter even."""
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


def _delta_sq(a: int, b: int, c: int) -> Fraction:
    """Triangle coefficient Delta(a b c) for doubled momenta (exact)."""
    return Fraction(fac2(a + b - c) * fac2(a - b + c) * fac2(-a + b + c), fac2(a + b + c + 2))


def _signed_sqrt(s: Fraction, sq: Fraction) -> float:
    """s * sqrt(sq) as a float, from the exact square s^2 * sq."""
    if s == 0:
        return 0.0
    v = s * s * sq
    r = math.sqrt(v.numerator / v.denominator)   # int / int rounds once
    return r if s > 0 else -r


@functools.lru_cache(maxsize=1 << 16)
def _3j(a: int, b: int, c: int, ma: int, mb: int, mc: int) -> float:
    if ma + mb + mc != 0 or not _tri_ok(a, b, c):
        return 0.0
    if abs(ma) > a or abs(mb) > b or abs(mc) > c:
        return 0.0
    if (a + ma) % 2 or (b + mb) % 2 or (c + mc) % 2:
        return 0.0
    # Racah: sum over k of (-1)^k / [k! (c-b+k+ma)/.. ...], doubled indices
    lo = max(0, b - c - ma, a - c + mb)
    hi = min(a + b - c, a - ma, b + mb)
    s = Fraction(0)
    for k in range(lo, hi + 1, 2):
        den = (fac2(k) * fac2(c - b + k + ma) * fac2(c - a + k - mb) *
               fac2(a + b - c - k) * fac2(a - k - ma) * fac2(b - k + mb))
        s += Fraction(-1 if (k >> 1) & 1 else 1, den)
    pre = _delta_sq(a, b, c) * (fac2(a + ma) * fac2(a - ma) * fac2(b + mb) *
                                fac2(b - mb) * fac2(c + mc) * fac2(c - mc))
    if ((a - b - mc) >> 1) & 1:
        s = -s
    return _signed_sqrt(s, pre)


@functools.lru_cache(maxsize=1 << 16)
def _6j(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    # triads (a b c), (a e f), (d b f), (d e c)
    triads = ((a, b, c), (a, e, f), (d, b, f), (d, e, c))
    if not all(_tri_ok(*t) for t in triads):
        return 0.0
    sums = [sum(t) for t in triads]
    bounds = (a + b + d + e, b + c + e + f, c + a + f + d)
    lo, hi = max(sums), min(bounds)
    s = Fraction(0)
    for t in range(lo, hi + 1, 2):
        den = fac2(t - sums[0]) * fac2(t - sums[1]) * fac2(t - sums[2]) * fac2(t - sums[3])
        for bnd in bounds:
            den *= fac2(bnd - t)
        s += Fraction((-1 if (t >> 1) & 1 else 1) * fac2(t + 2), den)
    pre = Fraction(1)
    for tr in triads:
        pre *= _delta_sq(*tr)
    return _signed_sqrt(s, pre)


@functools.lru_cache(maxsize=1 << 14)
def _9j(a, b, c, d, e, f, g, h, i) -> float:
    for tr in ((a, b, c), (d, e, f), (g, h, i), (a, d, g), (b, e, h), (c, f, i)):
        if not _tri_ok(*tr):
            return 0.0
    lo = max(abs(a - i), abs(d - h), abs(b - f))
    hi = min(a + i, d + h, b + f)
    tot = 0.0
    for x in range(lo, hi + 1, 2):
        w = (x + 1) * _6j(a, d, g, h, i, x) * _6j(b, e, h, d, x, f) * _6j(c, f, i, x, a, b)
        tot += -w if x & 1 else w
    return tot


def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    """( j1 j2 j3 ; m1 m2 m3 )."""
    return _3j(two(j1), two(j2), two(j3), two(m1), two(m2), two(m3))


def wigner6j(j1, j2, j3, j4, j5, j6) -> float:
    """{ j1 j2 j3 ; j4 j5 j6 }."""
    return _6j(two(j1), two(j2), two(j3), two(j4), two(j5), two(j6))


def wigner9j(j1, j2, j3, j4, j5, j6, j7, j8, j9) -> float:
    """{ j1 j2 j3 ; j4 j5 j6 ; j7 j8 j9 }."""
    return _9j(two(j1), two(j2), two(j3), two(j4), two(j5), two(j6),
               two(j7), two(j8), two(j9))
