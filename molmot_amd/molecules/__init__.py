"""Molecular tables consumed by the solver kernels."""
from .sroh import (MolecularData, load_sroh_from_julia, load_sroh,
                   synthetic_molecule, SROH_DATA_DIR,
                   GAMMA_RAD, K_WAVE, MASS_KG, LAMBDA_M)

__all__ = ["MolecularData", "load_sroh_from_julia", "load_sroh",
           "synthetic_molecule", "SROH_DATA_DIR", "GAMMA_RAD", "K_WAVE",
           "MASS_KG", "LAMBDA_M"]
