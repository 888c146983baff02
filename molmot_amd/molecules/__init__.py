"""Molecular tables consumed by the solver kernels: the SrOH CSV tables, a
seeded synthetic builder, and tables built from spectroscopic constants
(SrOH, CaOH, CaF) with their CSV wire format."""
from .sroh import (MolecularData, load_sroh_from_julia, load_sroh,
                   synthetic_molecule, SROH_DATA_DIR,
                   GAMMA_RAD, K_WAVE, MASS_KG, LAMBDA_M)
from .builders import (LinearMoleculeConstants, SROH, CAOH, CAF, build_tables,
                       build_sroh_hamiltonian, build_caoh_hamiltonian, build_caf_hamiltonian,
                       load_caoh_from_sroh_structure, load_caf_from_sroh_structure,
                       save_tables_csv, load_tables_csv, zeeman_per_gauss)

__all__ = ["MolecularData", "load_sroh_from_julia", "load_sroh",
           "synthetic_molecule", "SROH_DATA_DIR", "GAMMA_RAD", "K_WAVE",
           "MASS_KG", "LAMBDA_M", "LinearMoleculeConstants", "SROH", "CAOH", "CAF",
           "build_tables", "build_sroh_hamiltonian", "build_caoh_hamiltonian",
           "build_caf_hamiltonian", "load_caoh_from_sroh_structure",
           "load_caf_from_sroh_structure", "save_tables_csv", "load_tables_csv",
           "zeeman_per_gauss"]
