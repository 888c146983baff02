"""Hund's case (b) operators of a linear molecule (2S+1 Sigma ground states).

Matrix elements <s1| O |s2> in the |N S J I F M> basis; every angular
momentum is converted to a doubled integer first, phases are (-1)^(x/2) of
doubled exponents, and the reduced matrix elements are the standard
spherical-tensor forms (Brown & Carrington, Rotational Spectroscopy of
Diatomic Molecules, ch. 5 and 8).  Each function states the reference
operator whose values it reproduces.
"""
from __future__ import annotations

import math

from .basis import HundsCaseB_LinearMolecule, delta
from .wigner import _3j, _6j, two

__all__ = ["HundsCaseB_LinearMolecule", "Rotation", "RotationDistortion", "SpinRotation",
           "Hyperfine_IS", "Zeeman", "zeeman_nuclear", "TDM", "phase2"]

_VIB = ("v1", "v2", "v3")


def phase2(x2: int) -> int:
    """(-1)^(x2/2) for a doubled exponent x2 (must be even)."""
    if x2 & 1:
        raise ValueError("half-integer phase exponent")
    return -1 if (x2 >> 1) & 1 else 1


def _red(j2: int) -> float:
    """sqrt(j(j+1)(2j+1)) for doubled j: the reduced element of J^(1)."""
    return math.sqrt(j2 * (j2 + 2) * (j2 + 1) / 4.0)


def Rotation(s1, s2) -> float:
    """B [N(N+1) - Lambda^2], diagonal (molmot/states/case_b.py:97-115)."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N", "J", "F", "M"):
        return 0.0
    return s1.N * (s1.N + 1) - s1.Lambda ** 2


def RotationDistortion(s1, s2) -> float:
    """-D [N(N+1) - Lambda^2]^2, diagonal (molmot/states/case_b.py:118-136)."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N", "J", "F", "M"):
        return 0.0
    r = s1.N * (s1.N + 1) - s1.Lambda ** 2
    return -r * r


def SpinRotation(s1, s2) -> float:
    """gamma N.S: diagonal, [J(J+1) - N(N+1) - S(S+1)] / 2
    (molmot/states/case_b.py:139-204, whose J' != J terms are zero)."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N", "J", "F", "M"):
        return 0.0
    if two(s1.S) == 0:
        return 0.0
    J2, N2, S2 = two(s1.J), two(s1.N), two(s1.S)
    return (J2 * (J2 + 2) - N2 * (N2 + 2) - S2 * (S2 + 2)) / 8.0


def Hyperfine_IS(s1, s2) -> float:
    """Fermi contact b_F I.S, diagonal in N, F, M, off-diagonal in J
    (molmot/states/case_b.py:207-248):
    (-1)^(J'+I+F) (-1)^(N+S+J+1) [J][J'] <S||S||S> <I||I||I>
    {J J' 1; I I F} {S S 1; J' J N}, with [j] = sqrt(2j+1)."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N", "F", "M"):
        return 0.0
    N2, S2, I2 = two(s1.N), two(s1.S), two(s1.I)
    J2, Jp2, F2 = two(s1.J), two(s2.J), two(s1.F)
    if I2 == 0 or S2 == 0:
        return 0.0
    w = _6j(J2, Jp2, 2, I2, I2, F2) * _6j(S2, S2, 2, Jp2, J2, N2)
    if w == 0.0:
        return 0.0
    ph = phase2(Jp2 + I2 + F2) * phase2(N2 + S2 + J2 + 2)
    return ph * math.sqrt((J2 + 1) * (Jp2 + 1)) * _red(S2) * _red(I2) * w


def _f_coupling(s1, s2, p: int):
    """Common space-fixed part of a rank-1 operator acting on J (not I):
    (-1)^(F-M) (F 1 F'; -M p M') (-1)^(J+I+F'+1) [F][F'] {J' F' I; F J 1},
    or None when a selection rule fails."""
    F2, Fp2 = two(s1.F), two(s2.F)
    M2, Mp2 = two(s1.M), two(s2.M)
    J2, Jp2, I2 = two(s1.J), two(s2.J), two(s1.I)
    if abs(F2 - Fp2) > 2 or abs(J2 - Jp2) > 2 or M2 - Mp2 != 2 * p:
        return None
    w = _3j(F2, 2, Fp2, -M2, 2 * p, Mp2) * _6j(Jp2, Fp2, I2, F2, J2, 2)
    if w == 0.0:
        return 0.0
    return (phase2(F2 - M2) * phase2(J2 + I2 + Fp2 + 2) *
            math.sqrt((F2 + 1) * (Fp2 + 1)) * w)


def Zeeman(s1, s2, p: int) -> float:
    """Electron-spin Zeeman T^1_p(S), p in {-1, 0, 1}
    (molmot/states/case_b.py:443-490): (-1)^p times the F coupling times
    (-1)^(S+N+J'+1) [J][J'] <S||S||S> {S J' N; J S 1}."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N"):
        return 0.0
    S2, N2 = two(s1.S), two(s1.N)
    if S2 == 0:
        return 0.0
    fc = _f_coupling(s1, s2, p)
    if not fc:
        return 0.0
    J2, Jp2 = two(s1.J), two(s2.J)
    w = _6j(S2, Jp2, N2, J2, S2, 2)
    return ((-1) ** p * fc * phase2(S2 + N2 + Jp2 + 2) *
            math.sqrt((J2 + 1) * (Jp2 + 1)) * _red(S2) * w)


def zeeman_nuclear(s1, s2, p: int) -> float:
    """Nuclear-spin Zeeman T^1_p(I) (molmot/states/case_b.py:493-530):
    (-1)^p (-1)^(F-M) (F 1 F'; -M p M') (-1)^(I+J+F'+1) [F][F'] <I||I||I>
    {I F' J; F I 1}, diagonal in J."""
    if not delta(s1, s2, *_VIB, "S", "I", "Lambda", "ell", "K", "N", "J"):
        return 0.0
    I2, J2 = two(s1.I), two(s1.J)
    if I2 == 0:
        return 0.0
    F2, Fp2, M2, Mp2 = two(s1.F), two(s2.F), two(s1.M), two(s2.M)
    if abs(F2 - Fp2) > 2 or M2 - Mp2 != 2 * p:
        return 0.0
    w = _3j(F2, 2, Fp2, -M2, 2 * p, Mp2) * _6j(I2, Fp2, J2, F2, I2, 2)
    if w == 0.0:
        return 0.0
    return ((-1) ** p * phase2(F2 - M2) * phase2(I2 + J2 + Fp2 + 2) *
            math.sqrt((F2 + 1) * (Fp2 + 1)) * _red(I2) * w)


def TDM(s1, s2, p: int) -> float:
    """Electric-dipole T^1_p(d) between case (b) states, p in {-1, 0, 1}
    (molmot/states/case_b.py:926-983): -(-1)^p times the F coupling times
    (-1)^(N+S+J'+1) [J][J'] {N' J' S; J N 1} (-1)^(N-K) [N][N']
    (N 1 N'; -K q K') with q = K - K'."""
    if not delta(s1, s2, *_VIB, "S", "I", "ell"):
        return 0.0
    N2, Np2 = two(s1.N), two(s2.N)
    if abs(N2 - Np2) > 2:
        return 0.0
    K2, Kp2 = two(s1.K), two(s2.K)
    q2 = K2 - Kp2
    if abs(q2) > 2:
        return 0.0
    body = _3j(N2, 2, Np2, -K2, q2, Kp2)
    if abs(body) < 1e-15:
        return 0.0
    fc = _f_coupling(s1, s2, p)
    if not fc:
        return 0.0
    S2, J2, Jp2 = two(s1.S), two(s1.J), two(s2.J)
    w = _6j(Np2, Jp2, S2, J2, N2, 2)
    return (-((-1) ** p) * fc * phase2(N2 + S2 + Jp2 + 2) *
            math.sqrt((J2 + 1) * (Jp2 + 1)) * w * phase2(N2 - K2) *
            math.sqrt((N2 + 1) * (Np2 + 1)) * body)
