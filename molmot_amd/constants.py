"""Physical constants (SI) used on the rate-equation / trajectory path.

Values are CODATA 2018, as in molmot/constants.py:8-33.  ``HBAR_LEGACY`` is
the truncated value that the reference's Numba and 1D-CUDA kernels hard-code
(molmot/obe/rate_equations_jit.py:27, molmot_cuda/rate_equations_cuda.py:427);
the two differ by 1.612e-8 relative, which is the whole CPU/GPU gap the
reference reports.  Every kernel here takes hbar as an argument.
"""
import math

gS = 2.0023193043617
c = 299792458.0
h = 6.62607015e-34
hbar = 1.054571817e-34          # CODATA (molmot/constants.py:18)
HBAR_LEGACY = 1.0545718e-34     # rate_equations_jit.py:27
mu_B = 9.2740100783e-24
k_B = 1.380649e-23
amu = 1.66053906660e-27


def wavenumber(wavelength):
    """k = 2*pi/lambda in 1/m (molmot/constants.py:38-41)."""
    return 2.0 * math.pi / wavelength


def v_recoil(k, mass, hbar_value=hbar):
    """Photon recoil velocity hbar*k/m in m/s (molmot/constants.py:44-46)."""
    return hbar_value * k / mass


def I_sat(Gamma_hz, wavelength):
    """Saturation intensity pi*h*c*Gamma/(3 lambda^3) in W/m^2
    (molmot/constants.py:54-66)."""
    return math.pi * h * c * Gamma_hz / (3.0 * wavelength ** 3)
