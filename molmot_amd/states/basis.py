"""Coupled angular-momentum basis states of a linear molecule.

The two bases the molecular tables are built in:

* ``HundsCaseB_LinearMolecule`` |v1 v2 v3, S, I, Lambda, ell, K, N, J, F, M>
  (molmot/states/case_b.py:27-78): K = Lambda + ell, J = N + S, F = J + I.
* ``HundsCaseA_LinearMolecule`` |v1 v2 ell v3, Lambda, K, I, S, Sigma, J, P, F, M>
  (molmot/states/case_a.py:26-90): Sigma the projection of S on the axis,
  P = Lambda + Sigma.

``enumerate_states(cls, bounds)`` (molmot/states/basis.py:69-130) expands the
quantum numbers in the class's field order: a number in ``bounds`` is held
fixed, a list or range is iterated, and the rest follow from the class's
coupling rules (a value, or an inclusive range in unit steps).  The reference
leaves N (case b) and J (case a) unbounded by default (100 values above the
minimum); here they must be given, and a missing one raises ValueError.
Values are stored as floats, as in the reference, and every operator
converts them to doubled integers (``wigner.two``) before any arithmetic.
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from fractions import Fraction
from typing import Dict, List

from .wigner import two

__all__ = ["BasisState", "HundsCaseB_LinearMolecule", "HundsCaseA_LinearMolecule",
           "enumerate_states", "delta", "order_basis_by_m"]


def _fmt(x) -> str:
    return str(Fraction(x).limit_denominator(10))


@dataclass
class BasisState:
    """Base class: quantum numbers are the dataclass fields."""

    def qn_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)}

    def key(self) -> tuple:
        """Doubled-integer tuple of every quantum number (hashable)."""
        return tuple(two(getattr(self, f.name)) for f in fields(self))


@dataclass
class HundsCaseB_LinearMolecule(BasisState):
    v1: float = 0
    v2: float = 0
    v3: float = 0
    S: float = 0
    I: float = 0
    Lambda: float = 0
    ell: float = 0
    K: float = 0
    N: float = 0
    J: float = 0
    F: float = 0
    M: float = 0

    # coupling rules in field order: name -> f(assigned) giving a value or (lo, hi)
    rules = {
        "K": lambda q: q["Lambda"] + q["ell"],
        "J": lambda q: (abs(q["N"] - q["S"]), q["N"] + q["S"]),
        "F": lambda q: (abs(q["J"] - q["I"]), q["J"] + q["I"]),
        "M": lambda q: (-q["F"], q["F"]),
    }
    required = ("N",)

    def __repr__(self):
        return (f"|v=({self.v1},{self.v2},{self.v3}), N={_fmt(self.N)}, J={_fmt(self.J)}, "
                f"F={_fmt(self.F)}, M={_fmt(self.M)}>")


@dataclass
class HundsCaseA_LinearMolecule(BasisState):
    v1: float = 0
    v2: float = 0
    ell: float = 0
    v3: float = 0
    Lambda: float = 0
    K: float = 0
    I: float = 0
    S: float = 0
    Sigma: float = 0
    J: float = 0
    P: float = 0
    F: float = 0
    M: float = 0

    rules = {
        "K": lambda q: q["Lambda"] + q["ell"],
        "Sigma": lambda q: (-q["S"], q["S"]),
        "P": lambda q: q["Lambda"] + q["Sigma"],
        "F": lambda q: (abs(q["J"] - q["I"]), q["J"] + q["I"]),
        "M": lambda q: (-q["F"], q["F"]),
    }
    required = ("J",)

    def __repr__(self):
        return (f"|v=({self.v1},{self.v2},{self.v3}), Lambda={_fmt(self.Lambda)}, "
                f"Sigma={_fmt(self.Sigma)}, J={_fmt(self.J)}, P={_fmt(self.P)}, "
                f"F={_fmt(self.F)}, M={_fmt(self.M)}>")


def _values(spec) -> List[Fraction]:
    if isinstance(spec, (int, float, Fraction)):
        return [Fraction(spec).limit_denominator(1000)]
    return [Fraction(v).limit_denominator(1000) for v in spec]


def enumerate_states(state_class, qn_bounds: Dict[str, object]) -> list:
    """All basis states of ``state_class`` within ``qn_bounds``, in the
    order of molmot/states/basis.py:69-190 (the first field varies slowest)."""
    names = [f.name for f in fields(state_class)]
    rules = state_class.rules
    for r in state_class.required:
        if r not in qn_bounds:
            raise ValueError(f"{state_class.__name__}: {r} must be given in qn_bounds")
    partial = [dict()]
    for name in names:
        nxt = []
        for q in partial:
            if name in qn_bounds:
                choices = _values(qn_bounds[name])
            elif name in rules:
                r = rules[name](q)
                if isinstance(r, tuple):
                    lo, hi = (Fraction(x).limit_denominator(1000) for x in r)
                    choices = []
                    v = lo
                    while v <= hi:
                        choices.append(v)
                        v += 1
                else:
                    choices = [Fraction(r).limit_denominator(1000)]
            else:
                choices = [Fraction(0)]
            for v in choices:
                d = dict(q)
                d[name] = v
                nxt.append(d)
        partial = nxt
    return [state_class(**{k: float(v) for k, v in q.items()}) for q in partial]


def delta(s1, s2, *names) -> bool:
    """True when s1 and s2 agree on every named quantum number."""
    return all(two(getattr(s1, n)) == two(getattr(s2, n)) for n in names)


def order_basis_by_m(basis: list) -> list:
    return sorted(basis, key=lambda s: float(getattr(s, "M", 0)))
