"""The kernels divide by 2 pi and by the mass as RN(x * RN(1/y)) plus one
FMA correction (Markstein), which must equal the IEEE quotient the
reference's detunings and Euler updates use (exact detuning rounding).
Checked in C on 2e6 pseudo-random x per divisor over 140 binades; the same
count with the correction dropped (a planted mistake) is far from zero."""
import math

import pytest

import oracle
from molmot_amd.constants import amu


@pytest.mark.parametrize("y", [2.0 * math.pi, 105.0 * amu, 87.0 * amu, 3.7e7, 1.0 / 3.0, 7.0])
def test_one_correction_quotient_is_correctly_rounded(y, mol):
    assert oracle.div_rn_mismatches(y, 2_000_000, 12345) == 0
    assert oracle.div_rn_mismatches(mol.mass, 200_000, 7) == 0


def test_dropping_the_correction_is_detected():
    assert oracle.div_rn_mismatches(2.0 * math.pi, 200_000, 3, correct=False) > 10_000
