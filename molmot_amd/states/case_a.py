"""Hund's case (a) operators of a linear molecule (2 Pi excited states).

Matrix elements <s1| O |s2> in the |Lambda Sigma J P F M> basis, from doubled
integers as in ``case_b``.  Each function states the reference operator whose
values it reproduces.
"""
from __future__ import annotations

import math

from .basis import HundsCaseA_LinearMolecule, delta
from .case_b import phase2
from .wigner import _3j, two

__all__ = ["HundsCaseA_LinearMolecule", "Rotation", "SpinOrbit", "LambdaDoubling_p2q"]

_VIB = ("v1", "v2", "v3")


def Rotation(s1, s2) -> float:
    """B [J(J+1) - P^2] on the diagonal, and the spin-uncoupling term
    -sqrt(J(J+1) - P P') sqrt(S(S+1) - Sigma Sigma') between states with
    Sigma and P both one apart (molmot/states/case_a.py:102-140)."""
    if not delta(s1, s2, *_VIB, "I", "S", "Lambda", "ell", "K", "F", "M"):
        return 0.0
    J2, Jp2 = two(s1.J), two(s2.J)
    if J2 != Jp2:
        return 0.0
    S2 = two(s1.S)
    P1, P2 = two(s1.P), two(s2.P)
    G1, G2 = two(s1.Sigma), two(s2.Sigma)
    if P1 == P2 and G1 == G2:
        return (J2 * (J2 + 2) - P1 * P1) / 4.0
    if abs(G1 - G2) == 2 and abs(P1 - P2) == 2:
        return -math.sqrt((J2 * (J2 + 2) - P1 * P2) / 4.0) * \
            math.sqrt((S2 * (S2 + 2) - G1 * G2) / 4.0)
    return 0.0


def SpinOrbit(s1, s2) -> float:
    """A Lambda Sigma, diagonal (molmot/states/case_a.py:213-225)."""
    if not delta(s1, s2, *_VIB, "I", "S", "Lambda", "ell", "K", "Sigma", "J", "P", "F", "M"):
        return 0.0
    return s1.Lambda * s1.Sigma


def LambdaDoubling_p2q(s1, s2) -> float:
    """(p + 2q) Lambda-doubling: couples Lambda with Lambda' = Lambda + 2q,
    Sigma' = Sigma - q, P' = P + q for q = +-1
    (molmot/states/case_a.py:274-323):
    sum_q (-1)^(J-P) (-1)^(S-Sigma) sqrt(J(J+1)(2J+1) S(S+1)(2S+1))
    (J 1 J; -P -q P') (S 1 S; -Sigma q Sigma')."""
    if not delta(s1, s2, *_VIB, "I", "S", "ell", "J", "F", "M"):
        return 0.0
    L1, L2 = two(s1.Lambda), two(s2.Lambda)
    if L1 == 0 or L2 == 0:
        return 0.0
    J2, S2 = two(s1.J), two(s1.S)
    G1, G2 = two(s1.Sigma), two(s2.Sigma)
    P1, P2 = two(s1.P), two(s2.P)
    red = math.sqrt(J2 * (J2 + 2) * (J2 + 1) * S2 * (S2 + 2) * (S2 + 1) / 16.0)
    val = 0.0
    for q in (-1, 1):
        q2 = 2 * q
        if L2 != L1 + 2 * q2 or G2 != G1 - q2 or P2 != P1 + q2:
            continue
        w = _3j(J2, 2, J2, -P1, -q2, P2) * _3j(S2, 2, S2, -G1, q2, G2)
        val += phase2(J2 - P1) * phase2(S2 - G1) * red * w
    return val
